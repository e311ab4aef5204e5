// Background rendering job (reference src/render_job.cpp:9-59), with the frame-count and
// data-race defects of the reference fixed (see render_job.h).
#include <chrono>

#include "sol/render_job.h"
#include "sol/renderer.h"

namespace sol {

RenderJob::RenderJob(const Renderer& renderer, Image& output) : renderer(renderer), output(output) {}

RenderJob::~RenderJob() {
    cancel();
    if (render_thread_.joinable()) render_thread_.join();
}

void RenderJob::start(std::function<bool(const RenderJob&)>&& frame_end) {
    if (render_thread_.joinable()) render_thread_.join();
    is_done_ = false;
    cancelled_ = false;
    frames_ = 0;
    samples_ = 0;
    render_thread_ = std::thread([this, callback = std::move(frame_end)] {
        size_t per_frame = samples_per_frame ? samples_per_frame : 1;
        for (size_t i = 0; sample_count == 0 || i < sample_count; i += per_frame) {
            size_t count = per_frame;
            if (sample_count != 0 && count > sample_count - i) count = sample_count - i;
            renderer.render(output, i, count);
            samples_ += count;
            frames_ += 1;
            if ((callback && !callback(*this)) || cancelled_) break;
        }
        std::lock_guard<std::mutex> lock(mutex_);
        is_done_ = true;
        done_cond_.notify_all();
    });
}

bool RenderJob::wait(size_t timeout_ms) {
    if (!render_thread_.joinable()) return true;
    {
        std::unique_lock<std::mutex> lock(mutex_);
        if (timeout_ms != 0) {
            auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
            if (!done_cond_.wait_until(lock, until, [&] { return is_done_.load(); })) return false;
        } else {
            done_cond_.wait(lock, [&] { return is_done_.load(); });
        }
    }
    render_thread_.join();
    return true;
}

void RenderJob::cancel() { cancelled_ = true; }

} // namespace sol
