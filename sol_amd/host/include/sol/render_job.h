// Background rendering job, as sol::RenderJob (reference include/sol/render_job.h:18-49,
// src/render_job.cpp:9-59). Same interface. Differences, all fixes of reference defects:
//  - each frame renders min(samples_per_frame, remaining) samples (the reference passes
//    i + samples_per_frame as the count, render_job.cpp:27-31);
//  - the done flag is atomic and the frame callback is owned by the thread (the reference
//    races on a plain bool and captures a dangling reference, render_job.cpp:25,32,43,58).
#ifndef SOL_RENDER_JOB_H
#define SOL_RENDER_JOB_H

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace sol {

struct Image;
class Renderer;

struct RenderJob {
    size_t sample_count = 0;       ///< Samples to render (0 = until cancelled).
    size_t samples_per_frame = 1;  ///< Samples per render() call.
    const Renderer& renderer;
    Image& output;

    RenderJob(const Renderer& renderer, Image& output);
    RenderJob(const RenderJob&) = delete;
    ~RenderJob();

    void start(std::function<bool(const RenderJob&)>&& frame_end = {});
    /// Waits for the job (0 = no timeout). Returns true if it is over.
    bool wait(size_t timeout_ms = 0);
    void cancel();

    size_t frames_rendered() const { return frames_.load(); }
    size_t samples_rendered() const { return samples_.load(); }

private:
    std::thread render_thread_;
    std::mutex mutex_;
    std::condition_variable done_cond_;
    std::atomic<bool> is_done_{true};
    std::atomic<bool> cancelled_{false};
    std::atomic<size_t> frames_{0};
    std::atomic<size_t> samples_{0};
};

} // namespace sol

#endif
