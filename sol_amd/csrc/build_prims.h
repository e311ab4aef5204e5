// Device-wide primitives of the GPU BVH builder, written for 64-lane waves:
//
//   exclusive_scan     exclusive prefix sum of 32- or 64-bit words (reduce, scan of the tile
//                      sums, downsweep). The element count may be a device word, so a caller can
//                      scan an array whose length an earlier kernel wrote without reading it back.
//   radix_sort_pairs   stable least-significant-digit radix sort of 64-bit keys with 32-bit
//                      values, 8 bits a pass: a per-tile digit histogram, an exclusive scan of
//                      the histograms laid out digit-major (so equal digits keep tile order), and
//                      a scatter that ranks each key among the tile's keys of its digit with wave
//                      ballots, stages the tile in LDS in digit order and writes it out in runs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "common.h"

namespace solhip {
namespace prims {

constexpr int kThreads = 256;                 // threads per workgroup of every primitive
constexpr int kWavesPerBlock = kThreads / 64;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kThreads * kScanItems;    // elements per scan workgroup
constexpr int kSortItems = 16;
constexpr uint32_t kSortTile = kThreads * kSortItems;    // keys per sort workgroup
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

static_assert(kRadix == kThreads, "one thread per digit in the sort's per-digit steps");

inline uint32_t ceil_div(uint64_t a, uint32_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// Workspace words (of T) a scan of up to `bound` elements needs for its tile sums.
inline uint32_t scan_sums_words(uint32_t bound) { return ceil_div(bound, kScanTile) + 1; }
// Tiles of a sort of n keys, and the 32-bit workspace words it needs (digit histograms and the
// tile sums of their scan).
inline uint32_t sort_tiles(uint32_t n) { return ceil_div(n, kSortTile); }
inline uint64_t sort_hist_words(uint32_t n) { return static_cast<uint64_t>(kRadix) * sort_tiles(n); }
inline uint64_t sort_workspace_words(uint32_t n) {
    const uint64_t h = sort_hist_words(n);
    return h + scan_sums_words(static_cast<uint32_t>(h));
}

SOL_DEV unsigned long long lanes_below_mask() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Element count of a launch: the device word when given (never above the host bound the grid
// was sized for), else the bound itself.
SOL_DEV uint32_t count_of(const uint32_t* dcount, uint32_t bound) { return dcount ? min(*dcount, bound) : bound; }

template <typename T>
SOL_DEV T wave_inclusive_sum(T v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive prefix of v over the workgroup's threads in thread order; `total` receives the sum of
// all. Every thread of the workgroup must call it (two barriers).
template <typename T>
SOL_DEV T block_exclusive_sum(T v, T* lds_waves, T& total) {
    const T inc = wave_inclusive_sum(v);
    const int w = static_cast<int>(threadIdx.x) >> 6;
    if (__lane_id() == 63) lds_waves[w] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; ++k) {
        const T s = lds_waves[k];
        base += k < w ? s : T(0);
        tot += s;
    }
    total = tot;
    __syncthreads();
    return base + inc - v;
}

// ---- scan ------------------------------------------------------------------------------------------

// sums[b] = sum of tile b (kScanTile elements, read striped so a wave's loads are contiguous).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const T* in, const uint32_t* dcount, uint32_t bound,
                                                          T* sums) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t n = count_of(dcount, bound);
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;   // uniform over the workgroup
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = base + j * kThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    T total;
    (void)block_exclusive_sum(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan, in place, of the tile sums (one workgroup, kScanTile sums a step).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_sums(T* sums, const uint32_t* dcount, uint32_t bound) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t tiles = (count_of(dcount, bound) + kScanTile - 1) / kScanTile;
    T carry = 0;
    for (uint32_t c = 0; c < tiles; c += kScanTile) {
        T v[kScanItems];
        T s = 0;
        const uint32_t first = c + threadIdx.x * kScanItems;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            v[j] = first + j < tiles ? sums[first + j] : T(0);
            s += v[j];
        }
        T total;
        T run = carry + block_exclusive_sum(s, lds, total);
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (first + j < tiles) sums[first + j] = run;
            run += v[j];
        }
        carry += total;
    }
}

// out[i] = sums[tile] + exclusive prefix of in[] within the tile. Each thread takes kScanItems
// consecutive elements. out may be in.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_down(const T* in, T* out, const T* sums, const uint32_t* dcount,
                                                        uint32_t bound) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t n = count_of(dcount, bound);
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;   // uniform over the workgroup
    const uint32_t first = base + threadIdx.x * kScanItems;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = first + j < n ? in[first + j] : T(0);
        s += v[j];
    }
    T total;
    T run = sums[blockIdx.x] + block_exclusive_sum(s, lds, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (first + j < n) out[first + j] = run;
        run += v[j];
    }
}

// Exclusive prefix sum of in[0, n) into out (out may be in). n = *dcount if dcount is given (it
// must not exceed bound), else bound. sums: scan_sums_words(bound) words of workspace.
template <typename T>
hipError_t exclusive_scan(const T* in, T* out, const uint32_t* dcount, uint32_t bound, T* sums, hipStream_t st) {
    const uint32_t tiles = ceil_div(bound, kScanTile);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(tiles), dim3(kThreads), 0, st, in, dcount, bound, sums);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(kThreads), 0, st, sums, dcount, bound);
    hipLaunchKernelGGL(k_scan_down<T>, dim3(tiles), dim3(kThreads), 0, st, in, out, sums, dcount, bound);
    return hipGetLastError();
}

// ---- radix sort ------------------------------------------------------------------------------------

SOL_DEV uint32_t digit_of(unsigned long long key, int shift) { return static_cast<uint32_t>(key >> shift) & (kRadix - 1); }

// hist[d * tiles + t] = number of keys of tile t whose digit at `shift` is d.
__global__ __launch_bounds__(kThreads) void k_radix_hist(const unsigned long long* keys, uint32_t n, int shift,
                                                         uint32_t tiles, uint32_t* hist) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t i = base + j * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of(keys[i], shift)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter of tile t: wave w holds keys [t * kSortTile + w * kSortTile / 4, + kSortTile / 4) in 16
// rounds of 64 consecutive keys. In each round the lanes of equal digit find each other with eight
// ballots (one per digit bit); a key's rank among the wave's keys of its digit is the wave's
// running count of that digit (an LDS row per wave, advanced by the lowest lane of each digit
// group) plus the lanes below it in its group. After all rounds, per digit: the waves' counts in
// wave order give each wave's offset within the tile, the tile's counts in digit order each
// digit's start within the tile; keys are placed in LDS at digit start + wave offset + rank (the
// tile in stable digit order) and written out in that order to offs[d * tiles + t] + their place
// among the tile's keys of digit d.
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const unsigned long long* keys_in,
                                                            const uint32_t* vals_in, uint32_t n, int shift,
                                                            uint32_t tiles, const uint32_t* offs,
                                                            unsigned long long* keys_out, uint32_t* vals_out) {
    __shared__ uint32_t wrun[kWavesPerBlock][kRadix];   // per wave: running, then prefix counts
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t lds_waves[kWavesPerBlock];
    __shared__ unsigned long long skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    const int w = static_cast<int>(threadIdx.x) >> 6;
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < kRadix / 64; ++k) wrun[w][lane + 64 * k] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint32_t base = blockIdx.x * kSortTile + static_cast<uint32_t>(w) * (kSortTile / kWavesPerBlock);
    const unsigned long long below = lanes_below_mask();
    unsigned long long key[kSortItems];
    uint32_t val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t i = base + j * 64 + lane;
        const bool valid = i < n;
        key[j] = valid ? keys_in[i] : 0ull;
        val[j] = valid ? vals_in[i] : 0u;
        const uint32_t d = digit_of(key[j], shift);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const unsigned long long lower = peers & below;
        const uint32_t before = valid ? wrun[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();   // every lane has read before the group leaders write
        if (valid && lower == 0) wrun[w][d] = before + static_cast<uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        rank[j] = before + static_cast<uint32_t>(__popcll(lower));
    }
    __syncthreads();
    {
        const int d = static_cast<int>(threadIdx.x);
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kWavesPerBlock; ++k) {
            const uint32_t c = wrun[k][d];
            wrun[k][d] = run;
            run += c;
        }
        uint32_t total;
        dstart[d] = block_exclusive_sum(run, lds_waves, total);
        gbase[d] = offs[static_cast<uint32_t>(d) * tiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t i = base + j * 64 + lane;
        if (i < n) {
            const uint32_t d = digit_of(key[j], shift);
            const uint32_t p = dstart[d] + wrun[w][d] + rank[j];
            skeys[p] = key[j];
            svals[p] = val[j];
        }
    }
    __syncthreads();
    const uint32_t tile_base = blockIdx.x * kSortTile;
    const uint32_t count = min(kSortTile, n - tile_base);
    for (uint32_t p = threadIdx.x; p < count; p += kThreads) {
        const unsigned long long k = skeys[p];
        const uint32_t d = digit_of(k, shift);
        const uint32_t o = gbase[d] + (p - dstart[d]);
        if (o < n) {   // always: the histograms' scan places every key below n
            keys_out[o] = k;
            vals_out[o] = svals[p];
        }
    }
}

// Sorts (keys, vals)[0, n) by the low `bits` bits of the keys, stably, in passes of kRadixBits
// bits that alternate between the arrays and the tmp arrays; returns in `in_tmp` whether the
// result ended in the tmp arrays. work: sort_workspace_words(n) 32-bit words.
inline hipError_t radix_sort_pairs(unsigned long long* keys, uint32_t* vals, unsigned long long* keys_tmp,
                                   uint32_t* vals_tmp, uint32_t n, int bits, uint32_t* work, hipStream_t st,
                                   bool& in_tmp) {
    in_tmp = false;
    if (n <= 1) return hipSuccess;
    const uint32_t tiles = sort_tiles(n);
    const uint32_t hist_words = static_cast<uint32_t>(sort_hist_words(n));
    uint32_t* hist = work;
    uint32_t* sums = work + hist_words;
    unsigned long long *ki = keys, *ko = keys_tmp;
    uint32_t *vi = vals, *vo = vals_tmp;
    for (int shift = 0; shift < bits; shift += kRadixBits) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kThreads), 0, st, ki, n, shift, tiles, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = exclusive_scan<uint32_t>(hist, hist, nullptr, hist_words, sums, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kThreads), 0, st, ki, vi, n, shift, tiles, hist, ko, vo);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(ki, ko);
        std::swap(vi, vo);
        in_tmp = !in_tmp;
    }
    return hipSuccess;
}

} // namespace prims
} // namespace solhip
