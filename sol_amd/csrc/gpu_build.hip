// BVH construction on the GPU (solhip_scene_build). sol builds its BVH on the host
// (src/triangle_mesh.cpp:119-146); this builder produces the same kind of tree — binary, leaves of
// up to 8 triangles, SAH-driven — in the layouts the traversal kernels read, without leaving HBM:
//
//   1. triangle boxes, centroid bounds, 63-bit Morton codes of the centroids, stable radix sort
//      (build_prims.h);
//   2. PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
//      Hierarchy Construction", TVCG 2018): every cluster finds the cluster within `radius`
//      positions of the Morton order whose merged box has the smallest surface area; mutual
//      nearest neighbours merge; survivors are compacted in order; repeat until one cluster.
//      Ties go to the smaller index, so the globally closest pair is always mutual and every pass
//      merges at least once; the result is deterministic;
//   3. the SAH collapse, computed as each node is made (both its children were made by earlier
//      passes): a subtree of at most max_leaf_size triangles becomes a leaf when its leaf cost
//      (intersection cost x triangles x area) is no higher than its split cost, and the costs of
//      the 4-wide collapse are prepared;
//   4. breadth-first layouts, one pass per level with a prefix sum: the binary nodes
//      (sol_gpu_node), the leaf ranges of the triangles (leaf order), then the 4-wide nodes
//      (sol_gpu_node4: a binary node's children, opening the largest-area inner child until four,
//      as the host layout does); the top levels come first, as the LDS staging of the traversal
//      kernels needs;
//   5. triangles and shading records gathered into leaf order.
//
// The cluster count of a clustering pass and the node count of a layout level stay in device
// memory: each launch reads them there, with a grid sized for a bound the host knows (the last
// count it read; a level has at most two, or four, times the nodes of the one above it). The host
// reads the counts back every kPassesPerSync passes and kLevelsPerSync levels; launches past the
// end find nothing to do. Temporaries live in one workspace held by the context (grown, never
// shrunk), so a build allocates only the scene's own arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "build_prims.h"
#include "solhip.h"
#include "state.h"
#include "traverse.h"

namespace solhip {

namespace {
#define BUILD_TRY(expr)                                                                          \
    do {                                                                                         \
        hipError_t e__ = (expr);                                                                 \
        if (e__ != hipSuccess) {                                                                 \
            (void)hipGetLastError();                                                             \
            return set_error(SOLHIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));   \
        }                                                                                        \
    } while (0)

constexpr int kBB = 256;          // threads per block of the build kernels
constexpr int kMaxRadius = 32;
constexpr uint32_t kTopLevels = 6;    // binary levels laid out first (63 nodes at most)
constexpr uint32_t kTopLevels4 = 4;   // 4-wide levels laid out first (85 nodes at most)
constexpr int kPassesPerSync = 8;     // clustering passes launched between count read-backs
constexpr int kLevelsPerSync = 4;     // layout levels launched between count read-backs

unsigned grid_for(size_t n) { return static_cast<unsigned>((n + kBB - 1) / kBB); }

// Order-preserving map of floats to unsigned integers (for atomic min / max).
SOL_DEV unsigned f2o(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
SOL_HD float o2f(unsigned o) {
    unsigned u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

SOL_HD float half_area(float4 lo, float4 hi) {
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

SOL_DEV V3 vertex(const float* P, uint32_t i) { return V3{P[3 * i], P[3 * i + 1], P[3 * i + 2]}; }

// Node boxes: nodes [0, n) are the triangles (original index), [n, 2n - 1) the clusters made by
// the merges, in the order they are made. A triangle is a leaf of one triangle for the SAH
// collapse (count 1, cost C_i x area, no 4-wide cost).
__global__ void k_tri_boxes(const float* P, const uint32_t* I, uint32_t n, float4* lo, float4* hi, unsigned* cb,
                            float ci, unsigned* count, float* cost, unsigned* collapse, float* wcost,
                            unsigned* wsplit) {
    __shared__ unsigned s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = (threadIdx.x < 3) ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    uint32_t t = blockIdx.x * kBB + threadIdx.x;
    if (t < n) {
        V3 a = vertex(P, I[3 * t]), b = vertex(P, I[3 * t + 1]), c = vertex(P, I[3 * t + 2]);
        float4 l = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z), 0.0f);
        float4 h = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z), 0.0f);
        lo[t] = l;
        hi[t] = h;
        count[t] = 1;
        cost[t] = ci * half_area(l, h);
        collapse[t] = 1;
        reinterpret_cast<float4*>(wcost)[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        wsplit[t] = 0;
        float cx = (l.x + h.x) * 0.5f, cy = (l.y + h.y) * 0.5f, cz = (l.z + h.z) * 0.5f;
        atomicMin(&s[0], f2o(cx));
        atomicMin(&s[1], f2o(cy));
        atomicMin(&s[2], f2o(cz));
        atomicMax(&s[3], f2o(cx));
        atomicMax(&s[4], f2o(cy));
        atomicMax(&s[5], f2o(cz));
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&cb[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&cb[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_bounds_init(unsigned* cb) {
    if (threadIdx.x < 6) cb[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}

SOL_DEV unsigned long long spread21(unsigned v) {
    unsigned long long x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float4* lo, const float4* hi, uint32_t n, const unsigned* cb, unsigned long long* keys,
                         unsigned* ids) {
    uint32_t t = blockIdx.x * kBB + threadIdx.x;
    if (t >= n) return;
    float mn[3] = {o2f(cb[0]), o2f(cb[1]), o2f(cb[2])};
    float mx[3] = {o2f(cb[3]), o2f(cb[4]), o2f(cb[5])};
    float4 l = lo[t], h = hi[t];
    float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    unsigned q[3];
    for (int k = 0; k < 3; ++k) {
        float ext = mx[k] - mn[k];
        float u = ext > 0.0f ? (c[k] - mn[k]) / ext : 0.0f;
        float s = u * 2097152.0f;
        q[k] = s <= 0.0f ? 0u : (s >= 2097151.0f ? 2097151u : static_cast<unsigned>(s));
    }
    keys[t] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    ids[t] = t;
}

// State of the clustering, double-buffered by pass parity: a pass reads one copy and its last
// cluster writes the other.
struct PlocState {
    uint32_t m;           // clusters
    uint32_t next_node;   // index of the next node to make
    uint32_t passes;      // passes that merged
    uint32_t stalled;     // a pass merged nothing (cannot happen with a consistent order)
};

__global__ void k_ploc_init(PlocState* st, uint32_t n) {
    if (threadIdx.x == 0) st[0] = PlocState{n, n, 0u, 0u};
}

// Nearest neighbour of every cluster within `r` positions (merged half area; ties to the smaller
// index). The boxes of the block's window are staged in LDS.
__global__ __launch_bounds__(kBB) void k_nearest(const int* C, const PlocState* st, uint32_t bound, const float4* nlo,
                                                 const float4* nhi, int r, int* nn) {
    __shared__ float4 slo[kBB + 2 * kMaxRadius], shi[kBB + 2 * kMaxRadius];
    const uint32_t m = min(st->m, bound);
    if (m <= 1 || blockIdx.x * kBB >= m) return;   // uniform over the workgroup
    const int base = static_cast<int>(blockIdx.x) * kBB - r;
    for (int k = threadIdx.x; k < kBB + 2 * r; k += kBB) {
        int j = base + k;
        if (j >= 0 && j < static_cast<int>(m)) {
            int node = C[j];
            slo[k] = nlo[node];
            shi[k] = nhi[node];
        }
    }
    __syncthreads();
    const int i = static_cast<int>(blockIdx.x) * kBB + static_cast<int>(threadIdx.x);
    if (i >= static_cast<int>(m)) return;
    const int li = static_cast<int>(threadIdx.x) + r;
    const float4 al = slo[li], ah = shi[li];
    float best = __builtin_huge_valf();
    int best_j = -1;
    const int j0 = max(0, i - r), j1 = min(static_cast<int>(m) - 1, i + r);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float4 bl = slo[j - base], bh = shi[j - base];
        float4 l = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        float4 h = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
        float a = half_area(l, h);
        if (best_j < 0 || a < best) {   // j ascends: equal areas keep the smaller index
            best = a;
            best_j = j;
        }
    }
    nn[i] = best_j;
}

// Flags packed for one 64-bit prefix sum: high word 1 if cluster i makes a node (lower index of a
// mutual pair), low word 1 if cluster i survives (it is not the higher index of a mutual pair).
__global__ void k_merge_flags(const int* nn, const PlocState* st, uint32_t bound, unsigned long long* flags) {
    const uint32_t m = min(st->m, bound);
    uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (m <= 1 || i >= m) return;
    int j = nn[i];
    bool mutual = j >= 0 && nn[j] == static_cast<int>(i);
    bool make = mutual && static_cast<int>(i) < j;
    bool keep = !(mutual && j < static_cast<int>(i));
    flags[i] = (static_cast<unsigned long long>(make) << 32) | static_cast<unsigned long long>(keep);
}

// SAH collapse of a node made from children a and b (both final): count and cost of the subtree,
// collapse[node] = 1 if it becomes a leaf.
// The same step prepares the 4-wide collapse (the host layout's rule): each inner node becomes a
// 4-wide node or is opened into its parent's so that the summed half area of the 4-wide nodes is
// least. wcost[4 x + k - 1] is the least such sum for subtree x in at most k slots of its parent
// (0 for a leaf; for k = 1 x is a 4-wide node, area(x) + fill(x, 4)); byte k - 2 of wsplit[x] is
// how many of k slots the left subtree takes in the best fill(x, k).
SOL_DEV void finish_node(int node, int a, int b, float area, unsigned* count, float* cost, unsigned* collapse,
                         float* wcost, unsigned* wsplit, float ct, float ci, uint32_t max_leaf) {
    const unsigned cnt = count[a] + count[b];
    const float split = ct * area + cost[a] + cost[b];
    const float as_leaf = ci * static_cast<float>(cnt) * area;
    const bool leaf = cnt <= max_leaf && as_leaf <= split;
    count[node] = cnt;
    cost[node] = leaf ? as_leaf : split;
    collapse[node] = leaf ? 1u : 0u;
    float4* wc4 = reinterpret_cast<float4*>(wcost);
    if (leaf) {
        wc4[node] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        wsplit[node] = 0;
        return;
    }
    const float4 wa = wc4[a], wb = wc4[b];
    const float A[4] = {wa.x, wa.y, wa.z, wa.w}, B[4] = {wb.x, wb.y, wb.z, wb.w};
    float fill[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    unsigned sp = 0;
    for (int k = 2; k <= 4; ++k) {
        float best = HUGE_VALF;
        int bj = 1;
        for (int j = 1; j < k; ++j) {
            const float c = A[j - 1] + B[k - j - 1];
            if (c < best) {
                best = c;
                bj = j;
            }
        }
        fill[k] = best;
        sp |= static_cast<unsigned>(bj) << (8 * (k - 2));
    }
    const float wide = area + fill[4];
    wc4[node] = make_float4(wide, fminf(wide, fill[2]), fminf(wide, fill[3]), fminf(wide, fill[4]));
    wsplit[node] = sp;
}

// Makes the nodes of the mutual pairs, finishes their SAH collapse and compacts the surviving
// clusters in order into C_out. The last cluster writes the next state; a pass that finds one
// cluster or none passes the state and the cluster array on unchanged.
__global__ void k_merge(const int* C, const int* nn, const unsigned long long* flags, const unsigned long long* offs,
                        const PlocState* st, PlocState* st_out, uint32_t bound, float4* nlo, float4* nhi, int* left,
                        int* right, int* C_out, unsigned* count, float* cost, unsigned* collapse,
                        float* wcost, unsigned* wsplit, float ct, float ci, uint32_t max_leaf) {
    const PlocState s = *st;
    const uint32_t m = min(s.m, bound);
    uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (m <= 1) {
        if (i == 0) {
            *st_out = s;
            C_out[0] = C[0];
        }
        return;
    }
    if (i >= m) return;
    const unsigned long long f = flags[i], o = offs[i];
    const bool make = (f >> 32) != 0, keep = (f & 0xFFFFFFFFull) != 0;
    const uint32_t make_off = static_cast<uint32_t>(o >> 32), keep_off = static_cast<uint32_t>(o);
    if (i == m - 1) {
        const uint32_t made = make_off + (make ? 1u : 0u), kept = keep_off + (keep ? 1u : 0u);
        st_out->m = kept;
        st_out->next_node = s.next_node + made;
        st_out->passes = s.passes + 1;
        st_out->stalled = s.stalled | ((made == 0 || kept != m - made) ? 1u : 0u);
    }
    if (!keep) return;
    int out = C[i];
    if (make) {
        const int a = C[i], b = C[nn[i]];
        const int node = static_cast<int>(s.next_node + make_off);
        float4 al = nlo[a], ah = nhi[a], bl = nlo[b], bh = nhi[b];
        const float4 lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        const float4 hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
        nlo[node] = lo;
        nhi[node] = hi;
        left[node] = a;
        right[node] = b;
        finish_node(node, a, b, half_area(lo, hi), count, cost, collapse, wcost, wsplit, ct, ci, max_leaf);
        out = node;
    }
    C_out[keep_off] = out;
}

// Leaf reference of a leaf-like node and the triangles of its subtree written at `first`.
SOL_DEV int32_t emit_leaf(int node, uint32_t first, uint32_t n, const int* left, const int* right,
                          const unsigned* count, unsigned* order) {
    int stack[8];
    int sp = 0;
    uint32_t k = first;
    int cur = node;
    while (true) {
        if (cur < static_cast<int>(n)) {
            order[k++] = static_cast<unsigned>(cur);
            if (sp == 0) break;
            cur = stack[--sp];
        } else {
            if (sp < 8) stack[sp++] = right[cur];
            cur = left[cur];
        }
    }
    return static_cast<int32_t>(~((first << 3) | (count[node] - 1u)));
}

SOL_DEV bool inner_node(int node, uint32_t n, const unsigned* collapse) {
    return node >= static_cast<int>(n) && collapse[node] == 0;
}

// State of a breadth-first layout, double-buffered by level parity like PlocState.
struct LevelState {
    uint32_t size;         // nodes of the current level (0: the layout is done)
    uint32_t level_base;   // index of its first node
    uint32_t tri_base;     // binary layout: triangles placed in leaves so far
    uint32_t level;        // levels laid out
    uint32_t top_nodes;    // nodes of the first top levels
    uint32_t pad[3];
};

// Level 0 is the root (frontier F[0]).
__global__ void k_level_init(LevelState* st, int* F, int root) {
    if (threadIdx.x == 0) {
        st[0] = LevelState{1u, 0u, 0u, 0u, 0u, {0u, 0u, 0u}};
        F[0] = root;
    }
}

// Per node of a binary level: inner children (high word) and leaf triangles (low word).
__global__ void k_level_count(const int* F, const LevelState* st, uint32_t bound, uint32_t n, const int* left,
                              const int* right, const unsigned* count, const unsigned* collapse,
                              unsigned long long* packed) {
    const uint32_t size = min(st->size, bound);
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= size) return;
    const int node = F[k];
    unsigned long long inner = 0, tris = 0;
    for (int c = 0; c < 2; ++c) {
        int ch = c == 0 ? left[node] : right[node];
        if (inner_node(ch, n, collapse)) ++inner;
        else tris += count[ch];
    }
    packed[k] = (inner << 32) | tris;
}

__global__ void k_level_write(const int* F, const LevelState* st, LevelState* st_out, uint32_t bound, uint32_t n,
                              const unsigned long long* offs, const float4* nlo, const float4* nhi, const int* left,
                              const int* right, const unsigned* count, const unsigned* collapse, sol_gpu_node* out,
                              int* F_next, int32_t* leafref, unsigned* order, const unsigned long long* packed) {
    const LevelState s = *st;
    const uint32_t size = min(s.size, bound);
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (size == 0) {
        if (k == 0) *st_out = s;
        return;
    }
    if (k >= size) return;
    const uint32_t next_base = s.level_base + size;
    const unsigned long long o = offs[k];
    uint32_t inner_off = static_cast<uint32_t>(o >> 32), tri_off = static_cast<uint32_t>(o);
    if (k == size - 1) {
        LevelState t = s;
        t.size = inner_off + static_cast<uint32_t>(packed[k] >> 32);
        t.level_base = next_base;
        t.tri_base = s.tri_base + tri_off + static_cast<uint32_t>(packed[k]);
        t.level = s.level + 1;
        t.top_nodes = s.level < kTopLevels ? next_base : s.top_nodes;
        *st_out = t;
    }
    const int node = F[k];
    sol_gpu_node g;
    for (int c = 0; c < 2; ++c) {
        const int ch = c == 0 ? left[node] : right[node];
        int32_t ref;
        if (inner_node(ch, n, collapse)) {
            ref = static_cast<int32_t>(next_base + inner_off);
            F_next[inner_off] = ch;
            ++inner_off;
        } else {
            ref = emit_leaf(ch, s.tri_base + tri_off, n, left, right, count, order);
            leafref[ch] = ref;
            tri_off += count[ch];
        }
        const float4 l = nlo[ch], h = nhi[ch];
        float* b = c == 0 ? g.b0 : g.b1;
        b[0] = l.x; b[1] = h.x; b[2] = l.y; b[3] = h.y;
        g.bz[2 * c] = l.z;
        g.bz[2 * c + 1] = h.z;
        g.child[c] = ref;
    }
    g.pad[0] = g.pad[1] = 0;
    out[s.level_base + k] = g;
}

// The up-to-four children of the 4-wide node made from binary inner node `node`, left to right,
// as the collapse costs choose them: subtree x placed in k slots is opened into its two subtrees
// (split by wsplit) when that costs less than making x a 4-wide node. The slots of the pending
// subtrees and the children placed always sum to four, so at most four are pending.
SOL_DEV int wide_children(int node, uint32_t n, const int* left, const int* right, const unsigned* collapse,
                          const float* wcost, const unsigned* wsplit, int kids[4]) {
    int px[4], pk[4];
    int sp = 0, cnt = 0;
    const int j4 = static_cast<int>((wsplit[node] >> 16) & 0xFFu);
    px[sp] = right[node];
    pk[sp++] = 4 - j4;
    px[sp] = left[node];
    pk[sp++] = j4;
    while (sp > 0) {
        --sp;
        const int x = px[sp], k = pk[sp];
        if (!inner_node(x, n, collapse) || k == 1 || !(wcost[4 * size_t(x) + k - 1] < wcost[4 * size_t(x)])) {
            kids[cnt++] = x;
            continue;
        }
        const int j = static_cast<int>((wsplit[x] >> (8 * (k - 2))) & 0xFFu);
        px[sp] = right[x];
        pk[sp++] = k - j;
        px[sp] = left[x];
        pk[sp++] = j;
    }
    return cnt;
}

__global__ void k_wide_count(const int* F, const LevelState* st, uint32_t bound, uint32_t n, const int* left,
                             const int* right, const unsigned* collapse, const float* wcost, const unsigned* wsplit,
                             unsigned* cnt) {
    const uint32_t size = min(st->size, bound);
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= size) return;
    int kids[4];
    int c = wide_children(F[k], n, left, right, collapse, wcost, wsplit, kids);
    unsigned inner = 0;
    for (int i = 0; i < c; ++i) inner += inner_node(kids[i], n, collapse) ? 1u : 0u;
    cnt[k] = inner;
}

__global__ void k_wide_write(const int* F, const LevelState* st, LevelState* st_out, uint32_t bound, uint32_t n,
                             const unsigned* offs, const unsigned* cnt, const float4* nlo, const float4* nhi,
                             const int* left, const int* right, const unsigned* collapse, const float* wcost,
                             const unsigned* wsplit, const int32_t* leafref, sol_gpu_node4* out, int* F_next) {
    const LevelState s = *st;
    const uint32_t size = min(s.size, bound);
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (size == 0) {
        if (k == 0) *st_out = s;
        return;
    }
    if (k >= size) return;
    const uint32_t next_base = s.level_base + size;
    uint32_t off = offs[k];
    if (k == size - 1) {
        LevelState t = s;
        t.size = off + cnt[k];
        t.level_base = next_base;
        t.level = s.level + 1;
        t.top_nodes = s.level < kTopLevels4 ? next_base : s.top_nodes;
        *st_out = t;
    }
    int kids[4];
    const int c = wide_children(F[k], n, left, right, collapse, wcost, wsplit, kids);
    sol_gpu_node4 g;
    for (int i = 0; i < 4; ++i) {
        if (i >= c) {
            g.lo_x[i] = g.lo_y[i] = g.lo_z[i] = __builtin_huge_valf();   // the inverted box: never hit
            g.hi_x[i] = g.hi_y[i] = g.hi_z[i] = -__builtin_huge_valf();
            g.child[i] = SOL_GPU_EMPTY;
            g.pad[i] = 0;
            continue;
        }
        const int ch = kids[i];
        const float4 l = nlo[ch], h = nhi[ch];
        g.lo_x[i] = l.x; g.hi_x[i] = h.x;
        g.lo_y[i] = l.y; g.hi_y[i] = h.y;
        g.lo_z[i] = l.z; g.hi_z[i] = h.z;
        g.pad[i] = 0;
        if (inner_node(ch, n, collapse)) {
            g.child[i] = static_cast<int32_t>(next_base + off);
            F_next[off] = ch;
            ++off;
        } else {
            g.child[i] = leafref[ch];
        }
    }
    out[s.level_base + k] = g;
}

// A tree whose root is a leaf: one binary node with that leaf twice, one 4-wide node with it once.
__global__ void k_single_leaf(int root, uint32_t n, const float4* nlo, const float4* nhi, const int* left,
                              const int* right, const unsigned* count, unsigned* order, sol_gpu_node* out,
                              sol_gpu_node4* out4) {
    if (threadIdx.x != 0) return;
    const int32_t ref = emit_leaf(root, 0, n, left, right, count, order);
    const float4 l = nlo[root], h = nhi[root];
    sol_gpu_node g;
    for (int c = 0; c < 2; ++c) {
        float* b = c == 0 ? g.b0 : g.b1;
        b[0] = l.x; b[1] = h.x; b[2] = l.y; b[3] = h.y;
        g.bz[2 * c] = l.z;
        g.bz[2 * c + 1] = h.z;
        g.child[c] = ref;
    }
    g.pad[0] = g.pad[1] = 0;
    out[0] = g;
    sol_gpu_node4 w;
    const float inf = __builtin_huge_valf();   // unused slots: the inverted box, never hit
    for (int i = 0; i < 4; ++i) {
        w.lo_x[i] = i == 0 ? l.x : inf; w.hi_x[i] = i == 0 ? h.x : -inf;
        w.lo_y[i] = i == 0 ? l.y : inf; w.hi_y[i] = i == 0 ? h.y : -inf;
        w.lo_z[i] = i == 0 ? l.z : inf; w.hi_z[i] = i == 0 ? h.z : -inf;
        w.child[i] = i == 0 ? ref : SOL_GPU_EMPTY;
        w.pad[i] = 0;
    }
    out4[0] = w;
}

// Triangles and shading records in leaf order (the geometric normal as the host computes it:
// cross(v1 - v0, v2 - v0) in fp32 without contraction).
__global__ void k_gather(const unsigned* order, uint32_t n, const float* P, const float* N, const uint32_t* I,
                         const int32_t* tmat, const int32_t* tlight, const sol_material* mats, sol_gpu_tri* tris,
                         sol_gpu_shade* shade) {
    uint32_t s = blockIdx.x * kBB + threadIdx.x;
    if (s >= n) return;
    const uint32_t t = order[s];
    const uint32_t i0 = I[3 * t], i1 = I[3 * t + 1], i2 = I[3 * t + 2];
    const V3 a = vertex(P, i0), b = vertex(P, i1), c = vertex(P, i2);
    sol_gpu_tri g;
    g.v0[0] = a.x; g.v0[1] = a.y; g.v0[2] = a.z;
    g.v1[0] = b.x; g.v1[1] = b.y; g.v1[2] = b.z;
    g.v2[0] = c.x; g.v2[1] = c.y; g.v2[2] = c.z;
    g.prim = static_cast<int32_t>(t);
    const int32_t bsdf = tmat[t] >= 0 ? mats[tmat[t]].type : SOL_BSDF_NONE;
    g.pad0 = tri_kind(bsdf, tlight[t]);   // the kind word the traversal kernel packs into its hits
    g.pad1 = 0;
    tris[s] = g;
    sol_gpu_shade h;
    const uint32_t idx[3] = {i0, i1, i2};
    float* dst[3] = {h.n0, h.n1, h.n2};
    for (int k = 0; k < 3; ++k)
        for (int q = 0; q < 3; ++q) dst[k][q] = N[3 * idx[k] + q];
    const V3 fn = cross(b - a, c - a);
    h.face_n[0] = fn.x; h.face_n[1] = fn.y; h.face_n[2] = fn.z;
    h.material = tmat[t];
    h.light = tlight[t];
    h.bsdf = h.material >= 0 ? mats[h.material].type : SOL_BSDF_NONE;
    h.pad = 0;
    shade[s] = h;
}

// SAH cost of the binary layout: C_t x area of every inner node + C_i x triangles x area of every
// leaf, in double, divided by the root area on the host.
__global__ void k_sah(const sol_gpu_node* nodes, uint32_t num, float ct, float ci, double* acc) {
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= num) return;
    const sol_gpu_node g = nodes[k];
    double sum = 0.0;
    float4 l0 = make_float4(g.b0[0], g.b0[2], g.bz[0], 0), h0 = make_float4(g.b0[1], g.b0[3], g.bz[1], 0);
    float4 l1 = make_float4(g.b1[0], g.b1[2], g.bz[2], 0), h1 = make_float4(g.b1[1], g.b1[3], g.bz[3], 0);
    float4 lu = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0);
    float4 hu = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0);
    sum += static_cast<double>(ct) * half_area(lu, hu);
    for (int c = 0; c < 2; ++c) {
        if (g.child[c] < 0 && !(num == 1 && c == 1)) {
            uint32_t code = static_cast<uint32_t>(~g.child[c]);
            sum += static_cast<double>(ci) * ((code & 7u) + 1u) * half_area(c == 0 ? l0 : l1, c == 0 ? h0 : h1);
        }
    }
    atomicAdd(acc, sum);
}

template <typename T>
struct Buf {
    T* p = nullptr;
    Buf() = default;
    Buf(const Buf&) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1)); }
    T* release() {
        T* q = p;
        p = nullptr;
        return q;
    }
};

// Temporaries of a build, carved from the context's build workspace.
struct Workspace {
    float *P, *N;
    uint32_t* I;
    int32_t *tmat, *tlight;
    float4 *nlo, *nhi;
    int *left, *right, *C0, *C1, *nn, *F0, *F1;
    unsigned *cb, *count, *collapse, *wsplit, *order;
    float *cost, *wcost;
    int32_t* leafref;
    unsigned long long *keys0, *keys1, *flags, *offs, *sums64;
    uint32_t *sort_work, *sums32;
    PlocState* ploc;
    LevelState* lev;
    sol_gpu_node* nodes;
    sol_gpu_node4* nodes4;
    double* acc;

    // Lays the arrays out from `base` (null: only measures); returns the bytes taken.
    size_t carve(char* base, uint32_t n, uint32_t nv) {
        size_t off = 0;
        auto take = [&](auto*& p, size_t count) {
            using T = std::remove_reference_t<decltype(*p)>;
            off = (off + 255) & ~size_t(255);
            p = base ? reinterpret_cast<T*>(base + off) : nullptr;
            off += sizeof(T) * std::max<size_t>(count, 1);
        };
        const size_t nodes_cap = 2 * size_t(n);
        take(P, 3 * size_t(nv));
        take(N, 3 * size_t(nv));
        take(I, 3 * size_t(n));
        take(tmat, n);
        take(tlight, n);
        take(nlo, nodes_cap);
        take(nhi, nodes_cap);
        take(left, nodes_cap);
        take(right, nodes_cap);
        take(C0, n);
        take(C1, n);
        take(nn, n);
        take(F0, n);
        take(F1, n);
        take(cb, 6);
        take(count, nodes_cap);
        take(collapse, nodes_cap);
        take(wsplit, nodes_cap);
        take(order, n);
        take(cost, nodes_cap);
        take(wcost, 4 * nodes_cap);
        take(leafref, nodes_cap);
        take(keys0, n);
        take(keys1, n);
        take(flags, n);
        take(offs, n);
        take(sums64, prims::scan_sums_words(n));
        take(sort_work, prims::sort_workspace_words(n));
        take(sums32, prims::scan_sums_words(n));
        take(ploc, 2);
        take(lev, 2);
        take(nodes, n);
        take(nodes4, n);
        take(acc, 1);
        return off;
    }
};

} // namespace

int build_scene(Context& ctx, const solhip_mesh_desc& m, const solhip_build_config& cfg, solhip_scene** out,
                solhip_build_stats* stats) {
    const uint32_t n = m.num_tris, nv = m.num_vertices;
    hipStream_t st = ctx.stream;
    auto t0 = std::chrono::steady_clock::now();
    BUILD_TRY(hipSetDevice(ctx.device));

    Workspace w{};
    BUILD_TRY(ctx.build_ws.reserve(w.carve(nullptr, n, nv)));
    w.carve(ctx.build_ws.as<char>(), n, nv);

    // inputs
    BUILD_TRY(hipMemcpyAsync(w.P, m.positions, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(w.N, m.normals, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(w.I, m.indices, sizeof(uint32_t) * 3 * n, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(w.tmat, m.tri_material, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(w.tlight, m.tri_light, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));

    // 1. boxes (and the triangles' SAH leaf values), Morton codes, sort
    const float ct = cfg.traversal_cost, ci = cfg.intersection_cost;
    const uint32_t max_leaf = cfg.max_leaf_size;
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, w.cb);
    hipLaunchKernelGGL(k_tri_boxes, dim3(grid_for(n)), dim3(kBB), 0, st, w.P, w.I, n, w.nlo, w.nhi, w.cb, ci, w.count,
                       w.cost, w.collapse, w.wcost, w.wsplit);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBB), 0, st, w.nlo, w.nhi, n, w.cb, w.keys0,
                       reinterpret_cast<unsigned*>(w.C0));
    BUILD_TRY(hipGetLastError());
    bool in_tmp = false;
    BUILD_TRY(prims::radix_sort_pairs(w.keys0, reinterpret_cast<uint32_t*>(w.C0), w.keys1,
                                      reinterpret_cast<uint32_t*>(w.C1), n, 63, w.sort_work, st, in_tmp));
    // clusters start as the triangles in Morton order
    int* Cbuf[2] = {in_tmp ? w.C1 : w.C0, in_tmp ? w.C0 : w.C1};

    // 2. PLOC, with the SAH collapse of each node as it is made
    const int r = static_cast<int>(cfg.search_radius);
    uint32_t iterations = 0;
    int root = 0;
    if (n > 1) {
        hipLaunchKernelGGL(k_ploc_init, dim3(1), dim3(64), 0, st, w.ploc, n);
        uint32_t bound = n;
        int p = 0;
        while (true) {
            const unsigned g = grid_for(bound);
            for (int k = 0; k < kPassesPerSync; ++k, ++p) {
                const PlocState* si = w.ploc + (p & 1);
                PlocState* so = w.ploc + ((p + 1) & 1);
                int* Cc = Cbuf[p & 1];
                int* Cn = Cbuf[(p + 1) & 1];
                hipLaunchKernelGGL(k_nearest, dim3(g), dim3(kBB), 0, st, Cc, si, bound, w.nlo, w.nhi, r, w.nn);
                hipLaunchKernelGGL(k_merge_flags, dim3(g), dim3(kBB), 0, st, w.nn, si, bound, w.flags);
                BUILD_TRY(hipGetLastError());
                BUILD_TRY(prims::exclusive_scan<unsigned long long>(w.flags, w.offs, &si->m, bound, w.sums64, st));
                hipLaunchKernelGGL(k_merge, dim3(g), dim3(kBB), 0, st, Cc, w.nn, w.flags, w.offs, si, so, bound, w.nlo,
                                   w.nhi, w.left, w.right, Cn, w.count, w.cost, w.collapse, w.wcost,
                                   w.wsplit, ct, ci, max_leaf);
                BUILD_TRY(hipGetLastError());
            }
            PlocState hs{};
            BUILD_TRY(hipMemcpyAsync(&hs, w.ploc + (p & 1), sizeof(hs), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipMemcpyAsync(&root, Cbuf[p & 1], sizeof(int), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipStreamSynchronize(st));
            if (hs.stalled || hs.m == 0 || hs.m > bound)
                return set_error(SOLHIP_E_HIP, "GPU build: clustering made no progress");
            iterations = hs.passes;
            if (hs.m == 1) break;
            bound = hs.m;
        }
    }
    unsigned root_collapsed = 0;
    if (root >= static_cast<int>(n)) {
        BUILD_TRY(hipMemcpyAsync(&root_collapsed, w.collapse + root, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
    }
    // a root of more triangles than a leaf holds is never collapsed; a smaller one may be
    const bool single_leaf = root < static_cast<int>(n) || root_collapsed;

    // 3. layouts
    Buf<sol_gpu_node> gnodes;
    Buf<sol_gpu_node4> gnodes4;
    uint32_t num_nodes = 0, num_nodes4 = 0, depth = 0, depth4 = 0, top_nodes = 0, top_nodes4 = 0, leaves = 0;
    if (single_leaf) {
        BUILD_TRY(gnodes.alloc(1));
        BUILD_TRY(gnodes4.alloc(1));
        hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, st, root, n, w.nlo, w.nhi, w.left, w.right, w.count,
                           w.order, gnodes.p, gnodes4.p);
        BUILD_TRY(hipGetLastError());
        num_nodes = num_nodes4 = 1;
        depth = depth4 = 2;
        top_nodes = top_nodes4 = 1;
        leaves = 1;
    } else {
        int* Fb[2] = {w.F0, w.F1};
        // binary levels: a level has at most twice the inner nodes of the one above it
        hipLaunchKernelGGL(k_level_init, dim3(1), dim3(64), 0, st, w.lev, w.F0, root);
        LevelState hs{};
        uint32_t bound = 1;
        int p = 0;
        while (true) {
            for (int k = 0; k < kLevelsPerSync; ++k, ++p) {
                const LevelState* si = w.lev + (p & 1);
                LevelState* so = w.lev + ((p + 1) & 1);
                const unsigned g = grid_for(bound);
                hipLaunchKernelGGL(k_level_count, dim3(g), dim3(kBB), 0, st, Fb[p & 1], si, bound, n, w.left, w.right,
                                   w.count, w.collapse, w.flags);
                BUILD_TRY(hipGetLastError());
                BUILD_TRY(prims::exclusive_scan<unsigned long long>(w.flags, w.offs, &si->size, bound, w.sums64, st));
                hipLaunchKernelGGL(k_level_write, dim3(g), dim3(kBB), 0, st, Fb[p & 1], si, so, bound, n, w.offs, w.nlo,
                                   w.nhi, w.left, w.right, w.count, w.collapse, w.nodes, Fb[(p + 1) & 1], w.leafref,
                                   w.order, w.flags);
                BUILD_TRY(hipGetLastError());
                bound = std::min<uint32_t>(2 * bound, n);
            }
            BUILD_TRY(hipMemcpyAsync(&hs, w.lev + (p & 1), sizeof(hs), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipStreamSynchronize(st));
            if (hs.size == 0) break;
            bound = hs.size;
        }
        if (hs.tri_base != n) return set_error(SOLHIP_E_HIP, "GPU build: leaves do not cover every triangle");
        num_nodes = hs.level_base;
        depth = hs.level + 1;
        top_nodes = hs.top_nodes;
        leaves = num_nodes + 1;
        // 4-wide levels: at most four times the nodes of the level above
        hipLaunchKernelGGL(k_level_init, dim3(1), dim3(64), 0, st, w.lev, w.F0, root);
        bound = 1;
        p = 0;
        unsigned* cnt = reinterpret_cast<unsigned*>(w.flags);
        unsigned* cnt_off = reinterpret_cast<unsigned*>(w.offs);
        while (true) {
            for (int k = 0; k < kLevelsPerSync; ++k, ++p) {
                const LevelState* si = w.lev + (p & 1);
                LevelState* so = w.lev + ((p + 1) & 1);
                const unsigned g = grid_for(bound);
                hipLaunchKernelGGL(k_wide_count, dim3(g), dim3(kBB), 0, st, Fb[p & 1], si, bound, n, w.left, w.right,
                                   w.collapse, w.wcost, w.wsplit, cnt);
                BUILD_TRY(hipGetLastError());
                BUILD_TRY(prims::exclusive_scan<uint32_t>(cnt, cnt_off, &si->size, bound, w.sums32, st));
                hipLaunchKernelGGL(k_wide_write, dim3(g), dim3(kBB), 0, st, Fb[p & 1], si, so, bound, n, cnt_off, cnt,
                                   w.nlo, w.nhi, w.left, w.right, w.collapse, w.wcost, w.wsplit, w.leafref, w.nodes4,
                                   Fb[(p + 1) & 1]);
                BUILD_TRY(hipGetLastError());
                bound = std::min<uint32_t>(4 * bound, n);
            }
            BUILD_TRY(hipMemcpyAsync(&hs, w.lev + (p & 1), sizeof(hs), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipStreamSynchronize(st));
            if (hs.size == 0) break;
            bound = hs.size;
        }
        num_nodes4 = hs.level_base;
        depth4 = hs.level + 1;
        top_nodes4 = hs.top_nodes;
        // the scene keeps arrays of the exact sizes
        BUILD_TRY(gnodes.alloc(num_nodes));
        BUILD_TRY(gnodes4.alloc(num_nodes4));
        BUILD_TRY(hipMemcpyAsync(gnodes.p, w.nodes, sizeof(sol_gpu_node) * num_nodes, hipMemcpyDeviceToDevice, st));
        BUILD_TRY(hipMemcpyAsync(gnodes4.p, w.nodes4, sizeof(sol_gpu_node4) * num_nodes4, hipMemcpyDeviceToDevice, st));
    }
    if (depth > static_cast<uint32_t>(8 + kSpillEntries) || 3 * (depth4 - 1) > static_cast<uint32_t>(8 + kSpillEntries))
        return set_error(SOLHIP_E_TOO_DEEP, "GPU-built BVH depth " + std::to_string(depth) +
                                                " exceeds the traversal stack");

    // 4. leaf-order triangles and shading records
    Buf<sol_gpu_tri> tris;
    Buf<sol_gpu_shade> shade;
    Buf<sol_material> mats;
    Buf<sol_light> lights;
    BUILD_TRY(tris.alloc(n));
    BUILD_TRY(shade.alloc(n));
    BUILD_TRY(mats.alloc(m.num_materials));
    BUILD_TRY(lights.alloc(m.num_lights));
    if (m.num_materials)
        BUILD_TRY(hipMemcpyAsync(mats.p, m.materials, sizeof(sol_material) * m.num_materials, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(lights.p, m.lights, sizeof(sol_light) * m.num_lights, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n)), dim3(kBB), 0, st, w.order, n, w.P, w.N, w.I, w.tmat, w.tlight,
                       mats.p, tris.p, shade.p);
    BUILD_TRY(hipGetLastError());
    float sah = 0.0f;
    if (stats) {
        BUILD_TRY(hipMemsetAsync(w.acc, 0, sizeof(double), st));
        hipLaunchKernelGGL(k_sah, dim3(grid_for(num_nodes)), dim3(kBB), 0, st, gnodes.p, num_nodes, ct, ci, w.acc);
        BUILD_TRY(hipGetLastError());
        double total = 0.0;
        float4 rl, rh;
        BUILD_TRY(hipMemcpyAsync(&total, w.acc, sizeof(double), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipMemcpyAsync(&rl, w.nlo + root, sizeof(float4), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipMemcpyAsync(&rh, w.nhi + root, sizeof(float4), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
        float ra = half_area(rl, rh);
        sah = ra > 0.0f ? static_cast<float>(total / ra) : 0.0f;
    }
    BUILD_TRY(hipStreamSynchronize(st));

    auto s = std::make_unique<solhip_scene>();
    s->ctx = &ctx;
    s->num_nodes = num_nodes;
    s->num_tris = n;
    s->num_materials = m.num_materials;
    s->num_lights = m.num_lights;
    s->bvh_depth = depth;
    s->bvh4_depth = depth4;
    s->materials_host.assign(m.materials, m.materials + m.num_materials);
    s->lights_host.assign(m.lights, m.lights + m.num_lights);
    s->nodes = gnodes.release();
    s->nodes4 = gnodes4.release();
    s->tris = tris.release();
    s->shade = shade.release();
    s->materials = mats.release();
    s->lights = lights.release();
    SceneView& v = s->view;
    v.nodes = reinterpret_cast<const float4*>(s->nodes);
    v.nodes4 = reinterpret_cast<const float4*>(s->nodes4);
    v.num_nodes4 = num_nodes4;
    v.top_nodes4 = top_nodes4;
    v.tris = reinterpret_cast<const float4*>(s->tris);
    v.shade = reinterpret_cast<const float4*>(s->shade);
    v.materials = s->materials;
    v.lights = s->lights;
    v.num_lights = m.num_lights;
    v.top_nodes = top_nodes;
    v.num_nodes = num_nodes;
    v.num_tris = n;
    v.camera = m.camera;
    ctx.live_scenes += 1;
    *out = s.release();
    if (stats) {
        stats->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->iterations = iterations;
        stats->num_nodes = num_nodes;
        stats->num_nodes4 = num_nodes4;
        stats->num_leaves = leaves;
        stats->depth = depth;
        stats->depth4 = depth4;
        stats->sah_cost = sah;
    }
    return SOLHIP_OK;
}

} // namespace solhip
