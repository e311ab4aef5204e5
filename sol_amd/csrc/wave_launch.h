// Launchers of the wavefront kernels.
#pragma once

#include "launch.h"
#include "wave.h"

namespace solhip {

struct WaveLaunch {
    SceneView single;         // the scene of a single-scene run (unused when multi)
    WaveTables tabs;
    WaveBuf buf;
    sol_pt_config cfg;
    TraversalScratch scratch;
    int block;                // traversal kernels
    int lds_stack;
    int grid_traverse;
    int grid_shade;           // workgroups of k_sort, k_shade, k_start before wave_set_grids
    int grid_sort;            // resident workgroups of each kernel (set by wave_set_grids): these
    int grid_kshade[4];       // kernels hand out rounds by workgroup index, so every workgroup
    int grid_start;           // of the grid must be resident at once
    bool multi;
    bool wide;                // every scene has the 4-wide layout: k_trace traverses it
    uint32_t shade_mask;      // bit q: material queue q can receive entries (its kernel is launched)
    hipStream_t stream;
};

// Sizes the grids of the sort, shade and start kernels to the workgroups the device keeps
// resident (occupancy x compute units), capped at grid_shade.
hipError_t wave_set_grids(WaveLaunch& L, int compute_units);

hipError_t wave_launch_trace(const WaveLaunch& L, int cur);
hipError_t wave_launch_sort(const WaveLaunch& L, int cur);
// queue: kQueueDiffuse, kQueuePhong, kQueueSpecular, or kQueueAll (the merged queue)
hipError_t wave_launch_shade(const WaveLaunch& L, int queue, int nxt);
hipError_t wave_launch_start(const WaveLaunch& L, int nxt);
hipError_t wave_launch_set_chunk(const WaveLaunch& L, unsigned long long begin, unsigned long long end);
hipError_t wave_launch_check_drained(const WaveLaunch& L);
hipError_t wave_launch_reduce(const WaveLaunch& L, uint32_t first_tile, uint32_t num_tiles, uint32_t max_tile_pixels,
                              unsigned long long chunk_begin, uint32_t layer);

} // namespace solhip
