// BVH construction on the GPU (solhip_scene_build). sol builds its BVH on the host
// (src/triangle_mesh.cpp:119-146); this builder produces the same kind of tree — binary, leaves of
// up to 8 triangles, SAH-driven — in the layouts the traversal kernels read, without leaving HBM:
//
//   1. triangle boxes, centroid bounds, 63-bit Morton codes of the centroids, radix sort;
//   2. PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
//      Hierarchy Construction", TVCG 2018): every cluster finds the cluster within `radius`
//      positions of the Morton order whose merged box has the smallest surface area; mutual
//      nearest neighbours merge; survivors are compacted in order; repeat until one cluster.
//      Ties go to the smaller index, so the globally closest pair is always mutual and every pass
//      merges at least once; the result is deterministic;
//   3. bottom-up SAH pass (one thread per triangle climbing to the root, the second thread to
//      reach a node finishes it): a subtree of at most max_leaf_size triangles becomes a leaf when
//      its leaf cost (intersection cost x triangles x area) is no higher than its split cost;
//   4. breadth-first layouts, one pass per level with a prefix sum: the binary nodes
//      (sol_gpu_node), the leaf ranges of the triangles (leaf order), then the 4-wide nodes
//      (sol_gpu_node4: a binary node's children, opening the largest-area inner child until four,
//      as the host layout does); the top levels come first, as the LDS staging of the traversal
//      kernels needs;
//   5. triangles and shading records gathered into leaf order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

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
// the merges, in the order they are made.
__global__ void k_tri_boxes(const float* P, const uint32_t* I, uint32_t n, float4* lo, float4* hi, unsigned* cb) {
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

// Nearest neighbour of every cluster within `r` positions (merged half area; ties to the smaller
// index). The boxes of the block's window are staged in LDS.
__global__ __launch_bounds__(kBB) void k_nearest(const int* C, uint32_t m, const float4* nlo, const float4* nhi, int r,
                                                 int* nn) {
    __shared__ float4 slo[kBB + 2 * kMaxRadius], shi[kBB + 2 * kMaxRadius];
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
__global__ void k_merge_flags(const int* nn, uint32_t m, unsigned long long* flags) {
    uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= m) return;
    int j = nn[i];
    bool mutual = j >= 0 && nn[j] == static_cast<int>(i);
    bool make = mutual && static_cast<int>(i) < j;
    bool keep = !(mutual && j < static_cast<int>(i));
    flags[i] = (static_cast<unsigned long long>(make) << 32) | static_cast<unsigned long long>(keep);
}

__global__ void k_merge(const int* C, const int* nn, const unsigned long long* flags, const unsigned long long* offs,
                        uint32_t m, uint32_t next_node, float4* nlo, float4* nhi, int* left, int* right, int* parent,
                        int* C_out, unsigned* totals) {
    uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= m) return;
    const unsigned long long f = flags[i], o = offs[i];
    const bool make = (f >> 32) != 0, keep = (f & 0xFFFFFFFFull) != 0;
    const uint32_t make_off = static_cast<uint32_t>(o >> 32), keep_off = static_cast<uint32_t>(o);
    if (i == m - 1) {
        totals[0] = make_off + (make ? 1u : 0u);
        totals[1] = keep_off + (keep ? 1u : 0u);
    }
    if (!keep) return;
    int out = C[i];
    if (make) {
        const int a = C[i], b = C[nn[i]];
        const int node = static_cast<int>(next_node + make_off);
        float4 al = nlo[a], ah = nhi[a], bl = nlo[b], bh = nhi[b];
        nlo[node] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        nhi[node] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
        left[node] = a;
        right[node] = b;
        parent[a] = node;
        parent[b] = node;
        out = node;
    }
    C_out[keep_off] = out;
}

// Bottom-up SAH pass: count and cost of every node; collapse[node] = 1 if the subtree becomes a
// leaf. Triangles (nodes < n) are leaves of one triangle.
// The same pass prepares the 4-wide collapse (the host layout's rule): each inner node becomes a
// 4-wide node or is opened into its parent's so that the summed half area of the 4-wide nodes is
// least. wcost[4 x + k - 1] is the least such sum for subtree x in at most k slots of its parent
// (0 for a leaf; for k = 1 x is a 4-wide node, area(x) + fill(x, 4)); byte k - 2 of wsplit[x] is
// how many of k slots the left subtree takes in the best fill(x, k).
__global__ void k_collapse(uint32_t n, const float4* nlo, const float4* nhi, const int* left, const int* right,
                           const int* parent, unsigned* visits, unsigned* count, float* cost, unsigned* collapse,
                           float* wcost, unsigned* wsplit, float ct, float ci, uint32_t max_leaf) {
    uint32_t t = blockIdx.x * kBB + threadIdx.x;
    if (t >= n) return;
    count[t] = 1;
    cost[t] = ci * half_area(nlo[t], nhi[t]);
    collapse[t] = 1;
    for (int k = 0; k < 4; ++k) wcost[4 * size_t(t) + k] = 0.0f;
    wsplit[t] = 0;
    __threadfence();
    int node = parent[t];
    while (node >= 0) {
        if (atomicAdd(&visits[node], 1u) == 0) return;   // the other child finishes this node
        __threadfence();
        const int a = left[node], b = right[node];
        const unsigned cnt = __hip_atomic_load(&count[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                             __hip_atomic_load(&count[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float ca = __hip_atomic_load(&cost[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float cb = __hip_atomic_load(&cost[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float area = half_area(nlo[node], nhi[node]);
        const float split = ct * area + ca + cb;
        const float as_leaf = ci * static_cast<float>(cnt) * area;
        const bool leaf = cnt <= max_leaf && as_leaf <= split;
        count[node] = cnt;
        cost[node] = leaf ? as_leaf : split;
        collapse[node] = leaf ? 1u : 0u;
        float* wc = wcost + 4 * size_t(node);
        if (leaf) {
            wc[0] = wc[1] = wc[2] = wc[3] = 0.0f;
            wsplit[node] = 0;
        } else {
            float A[4], B[4];
            for (int k = 0; k < 4; ++k) {
                A[k] = __hip_atomic_load(&wcost[4 * size_t(a) + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                B[k] = __hip_atomic_load(&wcost[4 * size_t(b) + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
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
            wc[0] = wide;
            wc[1] = fminf(wide, fill[2]);
            wc[2] = fminf(wide, fill[3]);
            wc[3] = fminf(wide, fill[4]);
            wsplit[node] = sp;
        }
        __threadfence();
        node = parent[node];
    }
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

// Per node of a binary level: inner children (high word) and leaf triangles (low word).
__global__ void k_level_count(const int* F, uint32_t size, uint32_t n, const int* left, const int* right,
                              const unsigned* count, const unsigned* collapse, unsigned long long* packed) {
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

__global__ void k_level_write(const int* F, uint32_t size, uint32_t n, uint32_t level_base, uint32_t next_base,
                              uint32_t tri_base, const unsigned long long* offs, const float4* nlo, const float4* nhi,
                              const int* left, const int* right, const unsigned* count, const unsigned* collapse,
                              sol_gpu_node* out, int* F_next, int32_t* leafref, unsigned* order,
                              unsigned* totals, const unsigned long long* packed) {
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= size) return;
    const unsigned long long o = offs[k];
    uint32_t inner_off = static_cast<uint32_t>(o >> 32), tri_off = static_cast<uint32_t>(o);
    if (k == size - 1) {
        totals[0] = inner_off + static_cast<uint32_t>(packed[k] >> 32);
        totals[1] = tri_off + static_cast<uint32_t>(packed[k]);
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
            ref = emit_leaf(ch, tri_base + tri_off, n, left, right, count, order);
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
    out[level_base + k] = g;
}

// The up-to-four children of the 4-wide node made from binary inner node `node`, left to right,
// as k_collapse's costs choose them: subtree x placed in k slots is opened into its two subtrees
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

__global__ void k_wide_count(const int* F, uint32_t size, uint32_t n, const int* left, const int* right,
                             const unsigned* collapse, const float* wcost, const unsigned* wsplit, unsigned* cnt) {
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= size) return;
    int kids[4];
    int c = wide_children(F[k], n, left, right, collapse, wcost, wsplit, kids);
    unsigned inner = 0;
    for (int i = 0; i < c; ++i) inner += inner_node(kids[i], n, collapse) ? 1u : 0u;
    cnt[k] = inner;
}

__global__ void k_wide_write(const int* F, uint32_t size, uint32_t n, uint32_t level_base, uint32_t next_base,
                             const unsigned* offs, const unsigned* cnt, const float4* nlo, const float4* nhi,
                             const int* left, const int* right, const unsigned* collapse, const float* wcost,
                             const unsigned* wsplit, const int32_t* leafref, sol_gpu_node4* out, int* F_next,
                             unsigned* totals) {
    uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= size) return;
    uint32_t off = offs[k];
    if (k == size - 1) totals[0] = off + cnt[k];
    int kids[4];
    const int c = wide_children(F[k], n, left, right, collapse, wcost, wsplit, kids);
    sol_gpu_node4 g;
    for (int i = 0; i < 4; ++i) {
        if (i >= c) {
            g.lo_x[i] = g.hi_x[i] = g.lo_y[i] = g.hi_y[i] = g.lo_z[i] = g.hi_z[i] = 0.0f;
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
    out[level_base + k] = g;
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
    for (int i = 0; i < 4; ++i) {
        w.lo_x[i] = i == 0 ? l.x : 0.0f; w.hi_x[i] = i == 0 ? h.x : 0.0f;
        w.lo_y[i] = i == 0 ? l.y : 0.0f; w.hi_y[i] = i == 0 ? h.y : 0.0f;
        w.lo_z[i] = i == 0 ? l.z : 0.0f; w.hi_z[i] = i == 0 ? h.z : 0.0f;
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
    g.pad0 = g.pad1 = 0;
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

} // namespace

int build_scene(Context& ctx, const solhip_mesh_desc& m, const solhip_build_config& cfg, solhip_scene** out,
                solhip_build_stats* stats) {
    const uint32_t n = m.num_tris, nv = m.num_vertices;
    hipStream_t st = ctx.stream;
    auto t0 = std::chrono::steady_clock::now();
    BUILD_TRY(hipSetDevice(ctx.device));

    // inputs
    Buf<float> P, N;
    Buf<uint32_t> I;
    Buf<int32_t> tmat, tlight;
    BUILD_TRY(P.alloc(3 * size_t(nv)));
    BUILD_TRY(N.alloc(3 * size_t(nv)));
    BUILD_TRY(I.alloc(3 * size_t(n)));
    BUILD_TRY(tmat.alloc(n));
    BUILD_TRY(tlight.alloc(n));
    BUILD_TRY(hipMemcpyAsync(P.p, m.positions, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(N.p, m.normals, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(I.p, m.indices, sizeof(uint32_t) * 3 * n, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(tmat.p, m.tri_material, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    BUILD_TRY(hipMemcpyAsync(tlight.p, m.tri_light, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));

    // node arrays: [0, n) triangles, [n, 2n - 1) clusters
    const size_t nodes_cap = 2 * size_t(n);
    Buf<float4> nlo, nhi;
    Buf<int> left, right, parent, C0, C1, nn;
    Buf<unsigned> cb, totals, visits, count, collapse, wsplit, ids0, ids1;
    Buf<float> cost, wcost;
    Buf<unsigned long long> keys0, keys1, flags, offs;
    BUILD_TRY(nlo.alloc(nodes_cap));
    BUILD_TRY(nhi.alloc(nodes_cap));
    BUILD_TRY(left.alloc(nodes_cap));
    BUILD_TRY(right.alloc(nodes_cap));
    BUILD_TRY(parent.alloc(nodes_cap));
    BUILD_TRY(C0.alloc(n));
    BUILD_TRY(C1.alloc(n));
    BUILD_TRY(nn.alloc(n));
    BUILD_TRY(cb.alloc(6));
    BUILD_TRY(totals.alloc(4));
    BUILD_TRY(visits.alloc(nodes_cap));
    BUILD_TRY(count.alloc(nodes_cap));
    BUILD_TRY(collapse.alloc(nodes_cap));
    BUILD_TRY(cost.alloc(nodes_cap));
    BUILD_TRY(wcost.alloc(4 * nodes_cap));
    BUILD_TRY(wsplit.alloc(nodes_cap));
    BUILD_TRY(ids0.alloc(n));
    BUILD_TRY(ids1.alloc(n));
    BUILD_TRY(keys0.alloc(n));
    BUILD_TRY(keys1.alloc(n));
    BUILD_TRY(flags.alloc(n));
    BUILD_TRY(offs.alloc(n));

    // 1. boxes, Morton codes, sort
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, cb.p);
    hipLaunchKernelGGL(k_tri_boxes, dim3(grid_for(n)), dim3(kBB), 0, st, P.p, I.p, n, nlo.p, nhi.p, cb.p);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBB), 0, st, nlo.p, nhi.p, n, cb.p, keys0.p, ids0.p);
    BUILD_TRY(hipGetLastError());
    size_t sort_bytes = 0, scan_bytes = 0, scan32_bytes = 0;
    BUILD_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys0.p, keys1.p, ids0.p, ids1.p,
                                                 static_cast<int>(n), 0, 63, st));
    BUILD_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags.p, offs.p, static_cast<int>(n), st));
    BUILD_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan32_bytes, ids0.p, ids1.p, static_cast<int>(n), st));
    Buf<unsigned char> tmp;
    BUILD_TRY(tmp.alloc(std::max(sort_bytes, std::max(scan_bytes, scan32_bytes))));
    BUILD_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, sort_bytes, keys0.p, keys1.p, ids0.p, ids1.p,
                                                 static_cast<int>(n), 0, 63, st));
    // clusters start as the triangles in Morton order
    BUILD_TRY(hipMemcpyAsync(C0.p, ids1.p, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    BUILD_TRY(hipMemsetAsync(parent.p, 0xFF, sizeof(int) * nodes_cap, st));

    // 2. PLOC
    const int r = static_cast<int>(cfg.search_radius);
    uint32_t mcount = n, next_node = n, iterations = 0;
    int* Ccur = C0.p;
    int* Cnext = C1.p;
    while (mcount > 1) {
        hipLaunchKernelGGL(k_nearest, dim3(grid_for(mcount)), dim3(kBB), 0, st, Ccur, mcount, nlo.p, nhi.p, r, nn.p);
        hipLaunchKernelGGL(k_merge_flags, dim3(grid_for(mcount)), dim3(kBB), 0, st, nn.p, mcount, flags.p);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_bytes, flags.p, offs.p, static_cast<int>(mcount), st));
        hipLaunchKernelGGL(k_merge, dim3(grid_for(mcount)), dim3(kBB), 0, st, Ccur, nn.p, flags.p, offs.p, mcount,
                           next_node, nlo.p, nhi.p, left.p, right.p, parent.p, Cnext, totals.p);
        BUILD_TRY(hipGetLastError());
        unsigned tot[2];
        BUILD_TRY(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
        if (tot[0] == 0 || tot[1] != mcount - tot[0]) return set_error(SOLHIP_E_HIP, "GPU build: clustering made no progress");
        next_node += tot[0];
        mcount = tot[1];
        std::swap(Ccur, Cnext);
        ++iterations;
    }
    int root = 0;
    BUILD_TRY(hipMemcpyAsync(&root, Ccur, sizeof(int), hipMemcpyDeviceToHost, st));
    BUILD_TRY(hipStreamSynchronize(st));

    // 3. SAH collapse
    const uint32_t max_leaf = cfg.max_leaf_size;
    BUILD_TRY(hipMemsetAsync(visits.p, 0, sizeof(unsigned) * nodes_cap, st));
    hipLaunchKernelGGL(k_collapse, dim3(grid_for(n)), dim3(kBB), 0, st, n, nlo.p, nhi.p, left.p, right.p, parent.p,
                       visits.p, count.p, cost.p, collapse.p, wcost.p, wsplit.p, cfg.traversal_cost,
                       cfg.intersection_cost, max_leaf);
    BUILD_TRY(hipGetLastError());
    unsigned root_collapsed = 0;
    BUILD_TRY(hipMemcpyAsync(&root_collapsed, collapse.p + root, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    BUILD_TRY(hipStreamSynchronize(st));
    // a root of more triangles than a leaf holds is never collapsed; a smaller one may be
    const bool single_leaf = root < static_cast<int>(n) || root_collapsed;

    // 4. layouts
    Buf<sol_gpu_node> gnodes;
    Buf<sol_gpu_node4> gnodes4;
    Buf<int32_t> leafref;
    Buf<unsigned> order;
    Buf<int> F0, F1;
    BUILD_TRY(order.alloc(n));
    BUILD_TRY(leafref.alloc(nodes_cap));
    uint32_t num_nodes = 0, num_nodes4 = 0, depth = 0, depth4 = 0, top_nodes = 0, top_nodes4 = 0, leaves = 0;
    if (single_leaf) {
        BUILD_TRY(gnodes.alloc(1));
        BUILD_TRY(gnodes4.alloc(1));
        hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, st, root, n, nlo.p, nhi.p, left.p, right.p, count.p,
                           order.p, gnodes.p, gnodes4.p);
        BUILD_TRY(hipGetLastError());
        num_nodes = num_nodes4 = 1;
        depth = depth4 = 2;
        top_nodes = top_nodes4 = 1;
        leaves = 1;
    } else {
        // inner binary nodes: at most n - 1
        BUILD_TRY(gnodes.alloc(n));
        BUILD_TRY(F0.alloc(n));
        BUILD_TRY(F1.alloc(n));
        BUILD_TRY(hipMemcpyAsync(F0.p, &root, sizeof(int), hipMemcpyHostToDevice, st));
        uint32_t size = 1, level_base = 0, tri_base = 0, level = 0;
        int* Fc = F0.p;
        int* Fn = F1.p;
        while (size > 0) {
            hipLaunchKernelGGL(k_level_count, dim3(grid_for(size)), dim3(kBB), 0, st, Fc, size, n, left.p, right.p,
                               count.p, collapse.p, flags.p);
            BUILD_TRY(hipGetLastError());
            BUILD_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_bytes, flags.p, offs.p, static_cast<int>(size), st));
            hipLaunchKernelGGL(k_level_write, dim3(grid_for(size)), dim3(kBB), 0, st, Fc, size, n, level_base,
                               level_base + size, tri_base, offs.p, nlo.p, nhi.p, left.p, right.p, count.p,
                               collapse.p, gnodes.p, Fn, leafref.p, order.p, totals.p, flags.p);
            BUILD_TRY(hipGetLastError());
            unsigned tot[2];
            BUILD_TRY(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipStreamSynchronize(st));
            level_base += size;
            if (level < kTopLevels) top_nodes = level_base;
            tri_base += tot[1];
            size = tot[0];
            std::swap(Fc, Fn);
            ++level;
        }
        if (tri_base != n) return set_error(SOLHIP_E_HIP, "GPU build: leaves do not cover every triangle");
        num_nodes = level_base;
        depth = level + 1;
        leaves = num_nodes + 1;
        // 4-wide: at most as many nodes as binary inner nodes
        BUILD_TRY(gnodes4.alloc(num_nodes));
        BUILD_TRY(hipMemcpyAsync(F0.p, &root, sizeof(int), hipMemcpyHostToDevice, st));
        Fc = F0.p;
        Fn = F1.p;
        size = 1;
        level_base = 0;
        level = 0;
        unsigned* cnt = reinterpret_cast<unsigned*>(flags.p);
        unsigned* cnt_off = reinterpret_cast<unsigned*>(offs.p);
        while (size > 0) {
            hipLaunchKernelGGL(k_wide_count, dim3(grid_for(size)), dim3(kBB), 0, st, Fc, size, n, left.p, right.p,
                               collapse.p, wcost.p, wsplit.p, cnt);
            BUILD_TRY(hipGetLastError());
            BUILD_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan32_bytes, cnt, cnt_off, static_cast<int>(size), st));
            hipLaunchKernelGGL(k_wide_write, dim3(grid_for(size)), dim3(kBB), 0, st, Fc, size, n, level_base,
                               level_base + size, cnt_off, cnt, nlo.p, nhi.p, left.p, right.p, collapse.p, wcost.p,
                               wsplit.p, leafref.p, gnodes4.p, Fn, totals.p);
            BUILD_TRY(hipGetLastError());
            unsigned tot = 0;
            BUILD_TRY(hipMemcpyAsync(&tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, st));
            BUILD_TRY(hipStreamSynchronize(st));
            level_base += size;
            if (level < kTopLevels4) top_nodes4 = level_base;
            size = tot;
            std::swap(Fc, Fn);
            ++level;
        }
        num_nodes4 = level_base;
        depth4 = level + 1;
    }
    if (depth > static_cast<uint32_t>(8 + kSpillEntries) || 3 * (depth4 - 1) > static_cast<uint32_t>(8 + kSpillEntries))
        return set_error(SOLHIP_E_TOO_DEEP, "GPU-built BVH depth " + std::to_string(depth) +
                                                " exceeds the traversal stack");

    // 5. leaf-order triangles and shading records
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
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n)), dim3(kBB), 0, st, order.p, n, P.p, N.p, I.p, tmat.p, tlight.p,
                       mats.p, tris.p, shade.p);
    BUILD_TRY(hipGetLastError());
    float sah = 0.0f;
    if (stats) {
        Buf<double> acc;
        BUILD_TRY(acc.alloc(1));
        BUILD_TRY(hipMemsetAsync(acc.p, 0, sizeof(double), st));
        hipLaunchKernelGGL(k_sah, dim3(grid_for(num_nodes)), dim3(kBB), 0, st, gnodes.p, num_nodes,
                           cfg.traversal_cost, cfg.intersection_cost, acc.p);
        BUILD_TRY(hipGetLastError());
        double total = 0.0;
        float4 rl, rh;
        BUILD_TRY(hipMemcpyAsync(&total, acc.p, sizeof(double), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipMemcpyAsync(&rl, nlo.p + root, sizeof(float4), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipMemcpyAsync(&rh, nhi.p + root, sizeof(float4), hipMemcpyDeviceToHost, st));
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
