// Host-side launchers of the kernels, one translation unit per kernel family. Every launcher
// enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"
#include "wave.h"

namespace solhip {

// Traversal scratch shared by every traversal kernel of a context.
struct TraversalScratch {
    int* spill;                       // spill_entries * max_threads ints
    int spill_entries;                // per lane
    int max_threads;                  // capacity of the spill area, in lanes
    int* overflow;                    // set when a stack would overflow
};

struct LaunchShape {
    int block;       // threads per workgroup: 64, 128 or 256
    int lds_stack;   // LDS stack entries per lane: 8, 16 or 32
    int grid;        // workgroups (persistent: work is fetched from an atomic counter)
    bool wide;       // traverse the 4-wide layout where the scene has one
};

// Counters a kernel adds to: [0] extension rays, [1] shadow rays, [2] samples.
using StatCounters = unsigned long long*;

// Work counters of the intersection kernel: kTraceParts counters kCounterStride words apart (an
// even number of 32-bit words; launch_intersect clears them before each launch).
constexpr int kIntersectCounterWords = kTraceParts * kCounterStride;
static_assert(kIntersectCounterWords % 2 == 0, "cleared as 64-bit words");

hipError_t launch_intersect(const SceneView& s, const sol_ray* rays, uint64_t n, bool any_hit, sol_hit* hits,
                            int32_t* occluded, unsigned* work_counters, const TraversalScratch& scratch,
                            const LaunchShape& shape, hipStream_t stream);

hipError_t launch_megakernel(const SceneView& s, const sol_pt_config& cfg, uint32_t width, uint32_t height,
                             uint64_t sample_index, uint32_t sample_count, float* accum,
                             unsigned long long* work_counter, StatCounters stats, const TraversalScratch& scratch,
                             const LaunchShape& shape, hipStream_t stream);

// Adjoint path replay: adds dLoss/dparameter into d_mat (num_materials x 4) and d_light
// (num_lights x 3) for the loss gradient d_image (planar RGB, width x height); radiance holds each
// sample's radiance, pixel-major (sample_count per pixel).
hipError_t launch_adjoint(const SceneView& s, const sol_pt_config& cfg, uint32_t width, uint32_t height,
                          uint64_t sample_index, uint32_t sample_count, const float4* radiance, const float* d_image,
                          uint32_t num_materials,
                          float* d_mat, float* d_light, unsigned long long* work_counter, StatCounters stats,
                          const TraversalScratch& scratch, const LaunchShape& shape, hipStream_t stream);

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t stream);
// Clears n floats (a kernel, not a memset, so that it can be captured in a graph).
hipError_t launch_zero_f32(float* p, size_t n, hipStream_t stream);

// Workgroups for a persistent launch: enough for every CU at full occupancy.
int persistent_grid(int compute_units, int block);

} // namespace solhip
