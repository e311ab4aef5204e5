// Work fetch of the persistent traversal kernels (the wavefront renderer's k_trace and the batched
// intersection kernel): guided self-scheduling of a launch's entries over one counter, eight part
// counters or ranks; and k_trace's per-lane refill and leaf-wait thresholds.
#pragma once

#include "common.h"
#include "wave.h"

namespace solhip {

// Lanes below this one in its wave.
SOL_DEV unsigned long long lanes_below() {
    int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Launches of at least this many entries fetch from one counter, smaller ones from kTraceParts
// counters. Device atomics on one address serialise: a wave takes 64 entries at a time towards the
// end of a launch, so a launch of a million rays (the 256 x 256, 16-spp Cornell box) makes some
// 16,000 fetches in a few hundred microseconds, and with one counter they were its bound. A large
// launch fetches 1024 entries at a time, and with one counter all XCDs work on one narrow window
// of the queue, which the shared MALL serves better than eight windows. Measured (ms per frame,
// trace ms in brackets): Cornell box one counter 3.11 (2.50), eight parts always 2.14 (1.55); 1080p
// sphere field one counter 54.5 (42.4), eight parts always 54.8 (43.0), eight parts below 4M
// entries 53.7 (41.7), below 16M 53.7 (41.5).
constexpr unsigned kTracePartsBelow = 16u << 20;

// Launches of at most this many batches of 64 entries per wave are handed out by rank, with no
// atomic. A device atomic on one address waits behind the others on it: when the waves of a launch
// fetch faster than they trace, as in a launch of a million primary rays into the Cornell box,
// the part counters' atomics took 105 of its 119 µs with the traversal switched off (118.6 -> 13.0
// µs by rank). Launches of up to 4 / 16 / 64 / 512 batches a wave by rank: Cornell box 830 ->
// 993 / 998 / 994 / 992 Msamples/s, 1080p frame 51.8 -> 51.85 / 52.0 / 52.2 / 64.4 ms (ray costs
// vary, and a large launch needs the dynamic hand-out).
constexpr unsigned kRankedBatches = 4;

// A wave's place in the traversal kernel's work, the entries [0, total). A launch with no more
// entries than kRankedBatches x 64 per wave is handed out by rank: wave g of W takes the batches of
// 64 entries g, g + W, g + 2W, ... and no atomic is made. Larger launches are handed out from
// `parts` consecutive parts by their counters; the waves of workgroup b take from part b mod parts
// first (with eight parts, the XCD that runs them, as workgroups are dealt to the XCDs round
// robin), then from the following parts.
struct TraceWork {
    unsigned total, parts;
    unsigned part;          // the part it takes from now
    unsigned parts_left;    // parts not yet found used up
    unsigned taken;         // the part's counter as the wave last saw it
    unsigned by_rank;       // the wave's next batch when handed out by rank: [by_rank, by_rank + 64)
    bool ranked;            // handed out by rank

    SOL_DEV TraceWork(unsigned total_, unsigned block_waves) : total(total_) {
        parts = total >= kTracePartsBelow ? 1u : static_cast<unsigned>(kTraceParts);
        part = blockIdx.x % parts;
        parts_left = parts;
        taken = 0;
        const unsigned long long waves = static_cast<unsigned long long>(gridDim.x) * block_waves;
        ranked = total <= waves * 64 * kRankedBatches;
        // the wave's index in the workgroup, read as a scalar (the same in every lane of the wave)
        const unsigned wave = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(threadIdx.x)) / 64u;
        by_rank = ranked ? (blockIdx.x * block_waves + wave) * 64u : 0u;
    }
    SOL_DEV unsigned begin(unsigned p) const {
        return static_cast<unsigned>(static_cast<unsigned long long>(total) * p / parts);
    }
};

// Smallest chunk of the guided self-scheduling below: two entries per lane. One per lane made the
// atomics of the part counters the bound of launches of a few million entries (Cornell box
// 256 x 256, 16 spp: 758 -> 775 Msamples/s at 128; 1080p sphere field frame 52.07 -> 51.78 ms;
// batched intersection of 4M rays on the sphere field, closest hits 5574 -> 5751 Mrays/s, any hits
// 6333 -> 6674). At 256 the tails of the launches grow (Cornell 769, 1M-triangle soup closest hits
// 2903 -> 2796 Mrays/s).
constexpr unsigned kMinChunk = 128;

// Takes the wave's next chunk of entries [base, end); false once there is none left. Within a
// part, guided self-scheduling: a chunk is (entries left in the part) / (4 x waves per part), a
// multiple of 64 in [kMinChunk, 1024], so that a launch starts with few atomics and ends with small
// chunks.
SOL_DEV bool trace_fetch(unsigned* counters, unsigned waves, TraceWork& w, unsigned& base, unsigned& end) {
    if (w.ranked) {
        if (w.by_rank >= w.total) return false;
        base = w.by_rank;
        end = min(base + 64u, w.total);
        const unsigned stride = waves * 64u;   // a persistent grid holds a few thousand waves
        w.by_rank = w.total - base > stride ? base + stride : w.total;
        return true;
    }
    const unsigned per_part = waves / w.parts > 0 ? waves / w.parts : 1u;
    while (w.parts_left > 0) {
        const unsigned lo = w.begin(w.part), hi = w.begin(w.part + 1);
        const unsigned size = hi - lo;
        const unsigned left = size > w.taken ? size - w.taken : 0u;
        unsigned chunk = (left / (4 * per_part)) & ~63u;
        chunk = chunk < kMinChunk ? kMinChunk : (chunk > 1024 ? 1024 : chunk);
        // one atomic for the wave, by its first active lane; its result is read back as a scalar
        unsigned off = 0;
        if (__lane_id() == __builtin_amdgcn_readfirstlane(__lane_id())) off = atomicAdd(&counters[w.part * kCounterStride], chunk);
        off = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(off));
        if (off < size) {
            w.taken = off;
            base = lo + off;
            end = min(base + chunk, hi);
            return true;
        }
        w.part = w.part + 1 == w.parts ? 0u : w.part + 1;
        --w.parts_left;
        w.taken = 0;
    }
    return false;
}

// Lanes whose ray has ended take new entries as soon as this many lanes of the wave are idle, so
// that a wave does not wait for its slowest ray before fetching more work. Measured on the 1080p
// sphere field (trace ms per 64-spp frame): 8 lanes 63.7, 16 62.1, 24 60.7, 32 59.7, 40 59.9,
// 48 60.4, 56 62.2, 64 69.3; refills cost a divergent detour, leaf postponing keeps waiting lanes
// busy, so refilling half a wave at a time pays best.
constexpr int kRefillLanes = 32;

// A traversal step stops descending once this many lanes or fewer still look for their first
// leaf (traverse4_step's LEAF_WAIT): in a third of the node iterations of a 1080p frame fewer than
// 8 of the 64 lanes were still searching. Measured with instrumented counters on the 1080p sphere
// field (ms per frame, instrumented build): 0 lanes 72.5, 2 70.2, 4 70.0, 8 70.9, 12 72.1,
// 16 73.5, 32 80.9.
constexpr int kLeafWaitLanes = 4;

} // namespace solhip
