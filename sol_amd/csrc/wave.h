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

// Workgroup size of the sort, shade and start kernels.
constexpr int kWaveBlock = 512;
// Entries per lane and round of the sort: its rounds hold kWaveSortItems x kWaveBlock entries.
constexpr int kWaveSortItems = 4;
// The largest round of any pass (entries appended by one workgroup between two reservations).
constexpr int kWaveRound = kWaveSortItems * kWaveBlock;

enum : int {
    kQueueDiffuse = 0,
    kQueuePhong = 1,
    kQueueSpecular = 2,
    kNumShadeQueues = 3,
    kQueueAll = 3          // the merged queue's kernel (its entries are stored as queue 0)
};

// Diffuse / glossy (Phong, and the diffuse-Phong mix) / specular: the queues follow the BSDF's
// shading work, the mixed BSDF shares the Phong queue (it runs both lobes).
SOL_HD int queue_of(int type) {
    if (type == SOL_BSDF_DIFFUSE) return kQueueDiffuse;
    if (type == SOL_BSDF_PHONG || type == SOL_BSDF_INTERP) return kQueuePhong;
    return kQueueSpecular;
}

// Queues are split into nseg segments (WaveBuf::nseg: 1, 2, 4 or 8, a launch parameter), each with
// its own length counter: a workgroup appends to segment blockIdx % nseg only, so the device
// atomics of the appends spread over nseg addresses. Readers see a queue as the concatenation of
// its segments. Counters and item ranges are laid out for kMaxSegments.
constexpr int kMaxSegments = 8;

// The shading queues are either material-sorted (kNumShadeQueues: diffuse, glossy, specular, one
// kernel each) or merged into queue 0 with one kernel for every BSDF (kQueueAll), a launch
// parameter (WaveBuf::merged).

enum : int {
    kQExtend0 = 0,      // extension queues (two: one read, one written per bounce)
    kQExtend1 = 1,
    kQShade0 = 2,       // shade queues (kNumShadeQueues)
    kQShadow = 5,       // shadow queue
    kNumQueues = 6
};

// A small launch of the traversal kernel hands its entries out from kTraceParts consecutive
// parts, each with a work-fetch counter of its own (see trace_fetch in wave_kernels.hip).
constexpr int kTraceParts = 8;

// Device counters: the length of segment s of queue q is counter q * kMaxSegments + s, then the
// kTraceParts work-fetch counters of the traversal kernel (these kNumWaveCounters are cleared for
// every pass), then the start flag: nonzero while extension queue 0 holds the start fill of a
// chunk, whose entries leave out their first word (the ray origin is the camera's eye, tmin 0).
// Counters sit kCounterStride words (256 bytes) apart: device atomics on one cache line
// serialise, so each counter gets a line of its own.
constexpr int kCounterStride = 64;
SOL_HD constexpr int qcount(int queue, int seg) { return (queue * kMaxSegments + seg) * kCounterStride; }
enum : int {
    kNumWaveCounters = kNumQueues * kMaxSegments + kTraceParts,
    kCntTraceWork = kNumQueues * kMaxSegments * kCounterStride,
    kCntStartFill = kNumWaveCounters * kCounterStride,
    kCounterWords = (kNumWaveCounters + 1) * kCounterStride
};

// Statistics (rays traced, paths started) are summed over kStatSlots copies, kStatStride 64-bit
// words (256 bytes) apart, each workgroup adding to copy blockIdx % kStatSlots: with one copy the
// atomic of every wave of every launch landed on the same cache line.
constexpr int kStatSlots = 32;
constexpr int kStatStride = 32;
constexpr int kStatWords = kStatSlots * kStatStride;
// Word 3 of copy 0: entries left in the queues of chunks that should have ended (k_check_drained).
constexpr int kStatUndrained = 3;

// Work items of the current chunk (64-bit words, 256 bytes apart): segment s hands out
// items[item_next(s)] .. items[item_end(s)] - 1, and items[item_next(s) + 1] is where its range
// began (the start fill's items run from there to the counter's first value);
// items[kItemChunkBegin] is the first item of the chunk (index 0 of the sample buffer).
constexpr int kItemStride = 32;
SOL_HD constexpr int item_next(int seg) { return seg * kItemStride; }
SOL_HD constexpr int item_end(int seg) { return (kMaxSegments + seg) * kItemStride; }
enum : int {
    kItemChunkBegin = 2 * kMaxSegments * kItemStride,
    kItemWords = kItemChunkBegin + 1
};

// One bounce of all paths in flight is trace (closest hit of the extension rays and any hit of
// the previous bounce's shadow rays), sort (material queues from the hit triangle's kind byte;
// paths that missed end there) and shade (one kernel per material queue in use: hit decoding,
// emission, NEE, Russian roulette, BSDF sampling); trace and sort also clear the counters the
// kernels after them append to.
// A path's state travels with it in its queue entries, written at compacted positions
// (coalesced, field-major): extension rays (64 B: the ray and the path state), hits (16 B), shade
// entries (4 B: the index of the extension entry and its hit), shadow rays (48 B). The ray, hit,
// throughput and PCG state of a path stay where the trace read and wrote them until the shading
// kernel has read them (nothing writes that extension queue or the hits during the bounce), so the
// sort copies nothing. Only what other kernels add to out of order stays in a 16-byte record per
// slot (random access): the radiance (emission added by the shading, unoccluded light samples by
// the trace) and the scene. Appends are compacted per workgroup (one atomic per round of
// kWaveBlock entries, queue and segment).
// A path that ends writes its radiance to its sample's entry of the sample buffer and the slot
// starts the next work item at once (in the kernel that ended it); a shadow ray still in flight
// for the ended path adds its contribution to that sample entry instead of the slot record.
//
// slot record (float4):       radiance xyz, scene index (uint bits; kept by every update)
// extension entry (4 x float4): [0] org xyz, depth (uint bits; 0: a new path, traced from tmin 0,
//                             else from the configured ray offset)  [1] dir xyz, slot
//                             [2] throughput xyz, pdf of the previous bounce (not stored at depth
//                             0: throughput 1, pdf 0)
//                             [3] PCG state lo, hi, work item lo, hi (uint bits)
// shade entry (uint):         position i of the extension entry in its queue (its hit is hits[i])
// shadow entry (3 x float4):  [0] org xyz, target  [1] dir xyz, item hi  [2] contribution xyz, item lo
//                             target = slot, or kSampleTarget | 0 when the path has ended (the
//                             contribution then goes to the sample entry of the item)
//
// Segment capacities (entries), for a launch of n input entries handed out in rounds of at most
// R = kWaveRound entries by static round robin over a grid that is a multiple of kMaxSegments:
// segment s of nseg gets the rounds r with r % nseg == s, at most n / nseg + 2 R entries per
// launch and output queue. Shade queues and the shadow queue receive from launches whose inputs sum to at
// most `slots`; an extension queue receives from sort and the three shade launches, whose inputs
// sum to at most 2 slots.
struct WaveBuf {
    uint32_t slots;
    uint32_t nseg;                    // segments per queue (1, 2, 4 or 8)
    uint32_t merged;                  // 1: one shading queue (queue 0) for every BSDF
    uint32_t in_place;                // 1: every item of the chunk has a path in the start fill, and
                                      // a path's record is its item's sample entry (path = samples,
                                      // slot = item - chunk begin): an ended path needs no copy
    uint32_t cap_e, cap_m, cap_s;     // segment capacities: extension, shade, shadow queues
    uint32_t stride_e, stride_s;      // nseg * cap_e, cap_s (the queues' field strides)
    float4* path;                     // [slots]: slot records
    float4* hits;                     // [slots]: closest hit of extension entry i (in queue order):
                                      // leaf-order triangle (int bits, -1 = miss), t, u, v
    float4* eq[2];                    // [nseg * cap_e * 4] each
    uint32_t* mq[kNumShadeQueues];    // [nseg * cap_m] each (only mq[0] when merged)
    float4* sq;                       // [nseg * cap_s * 3]
    unsigned* counts;                 // kCounterWords
    unsigned long long* items;        // kItemWords
    unsigned long long* stats;        // kStatSlots lines of: [0] extension rays [1] shadow rays [2] samples
    float4* samples;                  // [sample_capacity]: radiance of each item of the chunk
    unsigned long long sample_capacity;

    // Queue entries are stored field-major: 16-byte word k of entry a (its storage index, segment *
    // capacity + position) follows every entry's word k - 1, so a wave reading or writing one word
    // of consecutive entries touches consecutive cache lines, a quarter as many as with whole
    // entries side by side (the texture path's cost is per line). Slot records are reached by
    // slot, which consecutive entries do not share in order. Word
    // indices fit 32 bits: at 2^27 slots an extension queue has 4 words of 2^28 entries.
    SOL_DEV float4* rec(uint32_t slot) const { return path + slot; }
    SOL_DEV float4* ext(int q, uint32_t a, uint32_t k) const { return eq[q] + (k * stride_e + a); }
    SOL_DEV uint32_t* shade_entry(int q, uint32_t a) const { return mq[q] + a; }
    SOL_DEV float4* shadow(uint32_t a, uint32_t k) const { return sq + (k * stride_s + a); }
};

// Jobs are images; a job's work items are ordered band by band (tile_h rows), tile by tile
// (tile_w columns) inside a band, sample by sample inside a tile, pixel by pixel (row-major)
// inside a tile. Edge bands and tiles are smaller. Items of a job are numbered from job_item0.
struct WaveTables {
    const SceneView* scenes;          // per scene
    uint32_t num_jobs;
    const unsigned long long* job_item0;    // num_jobs + 1 entries (last = total items)
    const unsigned long long* job_sample0;  // absolute sample index of the job's first sample
    const uint32_t* job_scene;
    float* const* job_accum;          // device image of each job
    // Optional per-sample radiance of each job (null = not kept): sample s of pixel p (row-major)
    // goes to job_sout[job][p * job_sstride[job] + job_soff[job] + s].
    float4* const* job_sout;
    const uint32_t* job_sstride;
    const uint32_t* job_soff;
    // Reduction layer of each job: jobs whose images overlap get increasing layers in job order,
    // and the layers of a chunk are reduced one launch after the other, so that overlapping jobs
    // add into the image in job order and never race.
    const uint32_t* job_layer;
    const uint32_t* job_w;
    const uint32_t* job_h;
    const uint32_t* job_spp;
    const uint32_t* job_tw;
    const uint32_t* job_th;
    // tiles, for the per-pixel reduction: the same order as the items
    const uint32_t* tile_job;
    const uint32_t* tile_x0;
    const uint32_t* tile_y0;
    const uint32_t* tile_w;
    const uint32_t* tile_h;
    const unsigned long long* tile_item0;   // first global item, plus one entry = total items
    uint32_t num_tiles;
};

} // namespace solhip
