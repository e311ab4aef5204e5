// Host-side checks built with sanitizers by tests/test_host_sanitizers.py (CPU only):
//  - sol::RenderJob (reference src/render_job.cpp:9-59, include/sol/render_job.h:18-49) over a
//    fake renderer whose speed the test sets: frame ranges, the callback stopping the job after k
//    frames, cancel() of an unbounded job, wait(timeout) returning false then true, the move
//    constructor on an idle and on a running job, the destructor of a running job, a restart.
//    Built with ThreadSanitizer (these alone) and with AddressSanitizer + UBSan;
//  - the host library and the oracle (AddressSanitizer + UBSan build only): sol's four test scenes
//    loaded from their TOML / OBJ / MTL files, procedural scenes with their BVH builds and
//    flattening, image files written, ray batches, and oracle renders and intersections over a
//    flattened scene.
// Prints one "ok <name>" line per check and exits non-zero at the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "sol/image.h"
#include "sol/render_job.h"
#include "sol/renderer.h"
#include "sol/scene.h"

#ifndef RENDER_JOB_ONLY
#include "oracle.h"
#include "sol_host.h"
#endif

namespace {

#define CHECK(cond)                                                                             \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);              \
            std::exit(1);                                                                       \
        }                                                                                       \
    } while (0)

using ms = std::chrono::milliseconds;

const sol::Scene& empty_scene() {
    static const sol::Scene scene;
    return scene;
}

// A renderer that adds `count` to pixel 0 of channel 0 and records every call's sample range,
// sleeping `delay` per call.
class FakeRenderer final : public sol::Renderer {
public:
    explicit FakeRenderer(int delay_ms) : Renderer("fake", empty_scene()), delay_(delay_ms) {}

    void render(sol::Image& image, size_t sample_index, size_t sample_count) const override {
        if (delay_) std::this_thread::sleep_for(ms(delay_));
        image.channel(0)[0] += static_cast<float>(sample_count);
        std::lock_guard<std::mutex> lock(mutex_);
        ranges_.push_back({sample_index, sample_count});
    }

    std::vector<std::pair<size_t, size_t>> ranges() const {
        std::lock_guard<std::mutex> lock(mutex_);
        return ranges_;
    }

private:
    int delay_;
    mutable std::mutex mutex_;
    mutable std::vector<std::pair<size_t, size_t>> ranges_;
};

void frames_cover_the_samples_once() {
    FakeRenderer r(0);
    sol::Image img(2, 2, 3);
    sol::RenderJob job(r, img);
    job.sample_count = 10;
    job.samples_per_frame = 3;
    job.start();
    CHECK(job.wait());
    auto rs = r.ranges();
    CHECK(rs.size() == 4);
    CHECK(rs[0] == std::make_pair(size_t(0), size_t(3)) && rs[1] == std::make_pair(size_t(3), size_t(3)));
    CHECK(rs[2] == std::make_pair(size_t(6), size_t(3)) && rs[3] == std::make_pair(size_t(9), size_t(1)));
    CHECK(job.frames_rendered() == 4 && job.samples_rendered() == 10);
    CHECK(img.channel(0)[0] == 10.0f);
    std::puts("ok frames_cover_the_samples_once");
}

void callback_false_stops_after_k_frames() {
    FakeRenderer r(0);
    sol::Image img(1, 1, 3);
    sol::RenderJob job(r, img);
    job.sample_count = 100;
    job.samples_per_frame = 2;
    std::atomic<int> calls{0};
    job.start([&](const sol::RenderJob& j) {
        ++calls;
        return j.frames_rendered() < 3;   // false after the third frame
    });
    CHECK(job.wait());
    CHECK(calls == 3 && job.frames_rendered() == 3 && job.samples_rendered() == 6);
    CHECK(r.ranges().size() == 3);
    std::puts("ok callback_false_stops_after_k_frames");
}

void cancel_stops_an_unbounded_job() {
    FakeRenderer r(2);
    sol::Image img(1, 1, 3);
    sol::RenderJob job(r, img);
    job.sample_count = 0;   // until cancelled
    job.start();
    std::this_thread::sleep_for(ms(30));
    job.cancel();
    CHECK(job.wait());
    size_t frames = job.frames_rendered();
    CHECK(frames >= 1);
    std::this_thread::sleep_for(ms(20));
    CHECK(job.frames_rendered() == frames);    // nothing renders after the job is over
    CHECK(r.ranges().size() == frames);
    std::puts("ok cancel_stops_an_unbounded_job");
}

void wait_times_out_on_a_slow_job() {
    FakeRenderer r(40);
    sol::Image img(1, 1, 3);
    sol::RenderJob job(r, img);
    job.sample_count = 4;
    job.start();
    CHECK(!job.wait(10));                      // 4 frames of 40 ms cannot end within 10 ms
    CHECK(job.wait());
    CHECK(job.frames_rendered() == 4 && img.channel(0)[0] == 4.0f);
    CHECK(job.wait(1));                        // an ended job waits no more
    std::puts("ok wait_times_out_on_a_slow_job");
}

void move_an_idle_and_a_running_job() {
    FakeRenderer r(3);
    sol::Image img(1, 1, 3);
    sol::RenderJob a(r, img);
    a.sample_count = 5;
    a.samples_per_frame = 5;
    sol::RenderJob b(std::move(a));            // idle
    CHECK(b.sample_count == 5 && b.samples_per_frame == 5 && &b.output == &img);
    b.start();
    CHECK(b.wait() && b.samples_rendered() == 5);

    sol::RenderJob c(r, img);
    c.sample_count = 0;
    std::atomic<const sol::RenderJob*> seen{nullptr};
    c.start([&](const sol::RenderJob& j) {
        seen = &j;
        return true;
    });
    std::this_thread::sleep_for(ms(15));
    sol::RenderJob d(std::move(c));            // running: the job now belongs to d
    CHECK(c.wait(1) && c.frames_rendered() == 0);   // the moved-from job is idle
    std::this_thread::sleep_for(ms(15));
    CHECK(seen.load() == &d);                  // later frames report the new owner
    d.cancel();
    CHECK(d.wait() && d.frames_rendered() >= 2);
    std::puts("ok move_an_idle_and_a_running_job");
}

void destructor_ends_a_running_job() {
    FakeRenderer r(2);
    sol::Image img(1, 1, 3);
    {
        sol::RenderJob job(r, img);
        job.sample_count = 0;
        job.start();
        std::this_thread::sleep_for(ms(10));
    }   // cancels and joins
    size_t n = r.ranges().size();
    std::this_thread::sleep_for(ms(10));
    CHECK(r.ranges().size() == n && n >= 1);
    std::puts("ok destructor_ends_a_running_job");
}

void restart_resets_the_counters() {
    FakeRenderer r(0);
    sol::Image img(1, 1, 3);
    sol::RenderJob job(r, img);
    job.sample_count = 3;
    job.start();
    CHECK(job.wait() && job.samples_rendered() == 3);
    job.sample_count = 2;
    job.start();
    CHECK(job.wait() && job.samples_rendered() == 2 && job.frames_rendered() == 2);
    CHECK(img.channel(0)[0] == 5.0f);
    std::puts("ok restart_resets_the_counters");
}

#ifndef RENDER_JOB_ONLY
void host_library_and_oracle(const std::string& data_dir, const std::string& tmp_dir) {
    const char* names[] = {"cornell_box", "cornell_box_glossy", "cornell_box_specular", "cornell_box_water"};
    for (const char* name : names) {
        std::string path = data_dir + "/" + name + ".toml";
        sol_scene* s = nullptr;
        CHECK(sol_scene_load(path.c_str(), nullptr, nullptr, &s) == SOL_OK);
        sol_mesh_view mv;
        CHECK(sol_scene_mesh(s, &mv) == SOL_OK && mv.num_tris > 0 && mv.num_lights > 0);
        sol_scene_destroy(s);
    }
    sol_scene* bad = nullptr;
    CHECK(sol_scene_load((data_dir + "/missing.toml").c_str(), nullptr, nullptr, &bad) == SOL_E_LOAD && !bad);
    std::puts("ok scene_files");

    sol_sphere_field_params p;
    sol_default_sphere_field_params(&p);
    p.num_spheres = 12;
    p.subdivision = 2;
    sol_scene* field = nullptr;
    CHECK(sol_scene_sphere_field(&p, nullptr, &field) == SOL_OK);
    sol_scene* box = nullptr;
    CHECK(sol_scene_cornell_box(1.0f, nullptr, &box) == SOL_OK);
    sol_scene* soup = nullptr;
    CHECK(sol_scene_triangle_soup(2000, 3, nullptr, &soup) == SOL_OK);
    std::puts("ok procedural_scenes");

    // oracle over each scene's product BVH: a small render and ray batches
    std::vector<sol_ray> rays(2048);
    CHECK(sol_gen_rays(0, rays.size(), 5, 1e-5f, rays.data()) == SOL_OK);
    for (sol_scene* s : {field, box, soup}) {
        sol_mesh_view mv;
        sol_bvh_view bv;
        CHECK(sol_scene_mesh(s, &mv) == SOL_OK && sol_scene_bvh(s, &bv) == SOL_OK);
        oracle_scene* o = nullptr;
        CHECK(oracle_scene_create(&mv, &bv, &o) == 0);
        sol_pt_config cfg;
        sol_default_pt_config(&cfg);
        cfg.max_path_len = 6;
        std::vector<float> img(3 * 16 * 12, 0.0f);
        CHECK(oracle_render(o, &cfg, 16, 12, 0, 2, img.data(), 2, 0, 0, nullptr) == 0);
        std::vector<sol_hit> hits(rays.size());
        CHECK(oracle_intersect(o, rays.data(), rays.size(), 0, 0, hits.data(), nullptr) == 0);
        std::vector<sol_surface_hit> surf(rays.size());
        CHECK(oracle_intersect_surface(o, rays.data(), rays.size(), surf.data()) == 0);
        for (size_t i = 0; i < rays.size(); ++i) CHECK((hits[i].prim >= 0) == (surf[i].hit != 0));
        oracle_scene_destroy(o);
    }
    std::puts("ok oracle_over_flattened_scenes");

    std::vector<float> pixels(3 * 8 * 4, 0.25f);
    CHECK(sol_image_save((tmp_dir + "/host_checks.exr").c_str(), SOL_IMAGE_EXR, pixels.data(), 8, 4, 3) == SOL_OK);
    CHECK(sol_image_save((tmp_dir + "/host_checks.pfm").c_str(), SOL_IMAGE_PFM, pixels.data(), 8, 4, 3) == SOL_OK);
    CHECK(sol_image_save((tmp_dir + "/host_checks.png").c_str(), SOL_IMAGE_PNG, pixels.data(), 8, 4, 3) != SOL_OK);
    std::puts("ok image_files");
    sol_scene_destroy(field);
    sol_scene_destroy(box);
    sol_scene_destroy(soup);
}
#endif

} // namespace

int main(int argc, char** argv) {
    frames_cover_the_samples_once();
    callback_false_stops_after_k_frames();
    cancel_stops_an_unbounded_job();
    wait_times_out_on_a_slow_job();
    move_an_idle_and_a_running_job();
    destructor_ends_a_running_job();
    restart_resets_the_counters();
#ifndef RENDER_JOB_ONLY
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s DATA_DIR TMP_DIR\n", argv[0]);
        return 2;
    }
    host_library_and_oracle(argv[1], argv[2]);
#else
    (void)argc;
    (void)argv;
#endif
    std::puts("all host checks passed");
    return 0;
}
