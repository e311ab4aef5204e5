// Data layout of the wavefront renderer (device side and the host code that sets it up).
//
// Work is a list of tiles; a tile covers tile_w x tile_h pixels of one image ("job") and all the
// samples the render call asks for. Tiles own consecutive global work items, sample-major inside
// the tile: item = tile_item0 + s * (tw * th) + (y - y0) * tw + (x - x0). A path slot takes the
// next item from a global atomic counter whenever its path ends, so slots never idle while work
// is left. Each finished path writes its colour to the item's entry of the sample buffer; a
// reduction then adds, per pixel, its samples in sample order - the order of sol's per-pixel loop
// (path_tracer.cpp:14-21) - and accumulates the sum into the image once.
#pragma once

#include "common.h"

namespace solhip {

enum : int {
    kQueueDiffuse = 0,
    kQueuePhong = 1,
    kQueueSpecular = 2,
    kNumShadeQueues = 3
};

// Device counters, one unsigned int each.
enum : int {
    kCntExtend0 = 0,    // extension queue 0 length
    kCntExtend1 = 1,    // extension queue 1 length
    kCntShade0 = 2,     // shade queue lengths (3)
    kCntShadow = 5,     // shadow queue length
    kCntRegen = 6,      // terminated-path queue length
    kCntWork = 8,       // per-kernel work-fetch counters: extend, shade x3, shadow, regen
    kNumWaveCounters = 16
};

struct WaveBuf {
    uint32_t slots;
    // path state [slots]
    float *tx, *ty, *tz;         // throughput
    float *cx, *cy, *cz;         // radiance gathered so far
    float* pdf;                  // pdf of the previous bounce (0 after a specular one)
    unsigned long long* rng;     // PCG state
    unsigned long long* item;    // global work item
    uint32_t* depth;             // bounces done
    uint32_t* scene;             // scene index (batched runs)
    // hit record [slots], written by extend, read by shade
    float *hpx, *hpy, *hpz;      // point
    float *hnx, *hny, *hnz;      // shading normal
    float *hfx, *hfy, *hfz;      // geometric normal (flipped to the ray side)
    float *hdx, *hdy, *hdz;      // ray direction
    int* hmat;                   // material index, bit 30 set for a front-side hit
    // extension queues [2][slots]: org, dir, tmin, slot
    float* eq[2][7];
    uint32_t* es[2];
    // shade queues [3][slots]: slot
    uint32_t* mq[kNumShadeQueues];
    // shadow queue [slots]: org, dir, contribution, slot
    float* sq[9];
    uint32_t* ss;
    // terminated paths [slots]
    uint32_t* rq;
    unsigned* counts;            // kNumWaveCounters
    unsigned long long* items;   // [0] next global item, [1] chunk end, [2] chunk begin
    unsigned long long* stats;   // [0] extension rays [1] shadow rays [2] samples
    float* samples;              // 3 planes of sample_capacity floats
    unsigned long long sample_capacity;
};

struct WaveTables {
    const SceneView* scenes;          // per scene
    const uint32_t* job_scene;
    float* const* job_accum;          // device image of each job
    const uint32_t* job_w;
    const uint32_t* job_h;
    const uint32_t* tile_job;
    const uint32_t* tile_x0;
    const uint32_t* tile_y0;
    const uint32_t* tile_w;
    const uint32_t* tile_h;
    const unsigned long long* tile_item0;   // first global item, plus one entry = total items
    const unsigned long long* tile_sample0; // absolute sample index of the tile's first sample
    uint32_t num_tiles;
};

} // namespace solhip
