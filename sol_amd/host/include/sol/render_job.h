// Background rendering job, as sol::RenderJob (reference include/sol/render_job.h:18-49,
// src/render_job.cpp:9-59). Same interface, including the move constructor. Differences, all fixes
// of reference defects:
//  - each frame renders min(samples_per_frame, remaining) samples (the reference passes
//    i + samples_per_frame as the count, render_job.cpp:27-31);
//  - the done flag is atomic and the frame callback is owned by the render thread (the reference
//    races on a plain bool and captures a dangling reference, render_job.cpp:25,32,43,58);
//  - the render thread works on a heap-held state, not on the RenderJob object, so a job can be
//    moved while it runs (the reference's move leaves the thread writing into the moved-from
//    object): the frame callback receives the object that owns the job when the frame ends.
#ifndef SOL_RENDER_JOB_H
#define SOL_RENDER_JOB_H

#include <cstddef>
#include <functional>
#include <memory>

namespace sol {

struct Image;
class Renderer;

/// Controlled from one thread (start / wait / cancel / move), as in the reference.
struct RenderJob {
    size_t sample_count = 0;       ///< Samples to render (0 = until cancelled).
    size_t samples_per_frame = 1;  ///< Samples per render() call.
    const Renderer& renderer;
    Image& output;

    RenderJob(const Renderer& renderer, Image& output);
    RenderJob(const RenderJob&) = delete;
    /// Takes over other's settings and its job, running or not; other is left an idle job of the
    /// same renderer and image.
    RenderJob(RenderJob&& other);
    ~RenderJob();   ///< cancels a running job and waits for it

    /// Renders frames of samples_per_frame samples on a background thread; after each frame the
    /// callback (if any) runs there, and the job stops when it returns false or on cancel().
    void start(std::function<bool(const RenderJob&)>&& frame_end = {});
    /// Waits for the job (0 = no timeout). Returns true if it is over.
    bool wait(size_t timeout_ms = 0);
    /// Asks the job to stop after the frame in progress.
    void cancel();

    size_t frames_rendered() const;
    size_t samples_rendered() const;

private:
    struct State;
    std::unique_ptr<State> state_;
};

} // namespace sol

#endif
