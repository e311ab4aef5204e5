// Background rendering job (reference src/render_job.cpp:9-59), with the frame-count, data-race
// and move defects of the reference fixed (see render_job.h).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>

#include "sol/render_job.h"
#include "sol/renderer.h"

namespace sol {

// What the render thread touches: never the RenderJob itself, except through `owner` for the
// callback's argument.
struct RenderJob::State {
    std::thread thread;
    std::mutex mutex;
    std::condition_variable done_cond;
    std::atomic<bool> done{true};
    std::atomic<bool> cancelled{false};
    std::atomic<size_t> frames{0};
    std::atomic<size_t> samples{0};
    std::atomic<RenderJob*> owner{nullptr};

    void join() {
        if (thread.joinable()) thread.join();
    }
};

RenderJob::RenderJob(const Renderer& renderer, Image& output)
    : renderer(renderer), output(output), state_(std::make_unique<State>()) {
    state_->owner = this;
}

RenderJob::RenderJob(RenderJob&& other)
    : sample_count(other.sample_count)
    , samples_per_frame(other.samples_per_frame)
    , renderer(other.renderer)
    , output(other.output)
    , state_(std::move(other.state_)) {
    state_->owner = this;
    other.state_ = std::make_unique<State>();
    other.state_->owner = &other;
}

RenderJob::~RenderJob() {
    if (!state_) return;
    cancel();
    state_->join();
}

void RenderJob::start(std::function<bool(const RenderJob&)>&& frame_end) {
    State* st = state_.get();
    st->join();
    st->done = false;
    st->cancelled = false;
    st->frames = 0;
    st->samples = 0;
    const size_t total = sample_count;
    const size_t per_frame = samples_per_frame ? samples_per_frame : 1;
    const Renderer* r = &renderer;
    Image* out = &output;
    st->thread = std::thread([st, total, per_frame, r, out, callback = std::move(frame_end)] {
        for (size_t i = 0; total == 0 || i < total; i += per_frame) {
            size_t count = per_frame;
            if (total != 0 && count > total - i) count = total - i;
            r->render(*out, i, count);
            st->samples += count;
            st->frames += 1;
            if ((callback && !callback(*st->owner.load())) || st->cancelled) break;
        }
        std::lock_guard<std::mutex> lock(st->mutex);
        st->done = true;
        st->done_cond.notify_all();
    });
}

bool RenderJob::wait(size_t timeout_ms) {
    State* st = state_.get();
    if (!st->thread.joinable()) return true;
    {
        std::unique_lock<std::mutex> lock(st->mutex);
        if (timeout_ms != 0) {
            auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
            if (!st->done_cond.wait_until(lock, until, [&] { return st->done.load(); })) return false;
        } else {
            st->done_cond.wait(lock, [&] { return st->done.load(); });
        }
    }
    st->join();
    return true;
}

void RenderJob::cancel() { state_->cancelled = true; }

size_t RenderJob::frames_rendered() const { return state_->frames.load(); }
size_t RenderJob::samples_rendered() const { return state_->samples.load(); }

} // namespace sol
