// BVH traversal for one ray per lane: closest hit and any hit over the binary GPU layout
// (sol_gpu_node: both children's boxes in the parent, leaves referenced from the parent).
// Replaces bvh::SingleRayTraverser<Bvh>::traverse<IsAnyHit> as called by
// TriangleMesh::intersect_closest / intersect_any (src/triangle_mesh.cpp:64-117 of sol).
//
// Order: both children are tested; if both are hit the nearer one (the first on a tie) is
// visited and the other pushed. The stack is a per-lane short stack in LDS (entry e of lane l
// at lds[e * BLOCK + l], bank-conflict free) that spills to a per-lane global area
// (coalesced, entry e at spill[(e - LDS_STACK) * spill_stride]). The top levels of the tree are
// read from an LDS copy made at kernel start.
//
// Closest hit is the lexicographic minimum of (t, original triangle index) over the triangles
// with tmin < t < tmax, so the result does not depend on the traversal order.
#pragma once

#include "common.h"

namespace solhip {

// Maximum number of top nodes copied to LDS (4 KB: six levels).
constexpr int kTopNodesMax = 63;

// Word 3 of a raw buffer descriptor on gfx9-family parts (32-bit data format, no swizzle).
constexpr int kBufferFlags = 0x00020000;

struct HitRecord {
    int slot;   // leaf-order triangle index, -1 = miss
    int prim;   // original triangle index
    float t, u, v;
};

typedef __attribute__((address_space(3))) int lds_int;     // an int in LDS
typedef __attribute__((address_space(1))) int global_int;  // an int in device memory

struct Stack {
    lds_int* lds;        // this lane's first entry
    global_int* spill;   // this lane's first spill entry
    int spill_stride;    // distance between consecutive spill entries
    int capacity;        // LDS + spill entries
    global_int* overflow;  // set to 1 if a push would exceed the capacity

    // Entries live in LDS (lds_stack, entry e of lane l at lds_stack[e * block + l]) and spill to
    // the lane's column of the global area. Explicit address spaces keep the two paths apart
    // (ds_read / global_load rather than flat loads).
    SOL_DEV static Stack make(int* lds_stack, int lane_in_block, int* spill_base, int lane_global, int spill_stride,
                              int capacity, int* overflow) {
        Stack st;
        st.lds = (lds_int*)(lds_stack + lane_in_block);
        st.spill = (global_int*)(spill_base + lane_global);
        st.spill_stride = spill_stride;
        st.capacity = capacity;
        st.overflow = (global_int*)(overflow);
        return st;
    }
};

template <int BLOCK>
SOL_DEV void load_top_nodes(const SceneView& s, float4* lds_nodes) {
    int n = static_cast<int>(s.top_nodes < static_cast<uint32_t>(kTopNodesMax) ? s.top_nodes : kTopNodesMax) * 4;
    for (int i = threadIdx.x; i < n; i += BLOCK) lds_nodes[i] = s.nodes[i];
    __syncthreads();
}

// Node and triangle fetch of a scene that is the same for the whole wave: the top levels from the
// workgroup's LDS copy (ds_read), the rest through buffer loads with 32-bit byte offsets from a
// scalar buffer descriptor (the hardware returns zeros beyond num_records, never faults).
struct FetchUniform {
    __amdgpu_buffer_rsrc_t nodes, tris;
    const float4* lds;
    int top;

    SOL_DEV static __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, static_cast<int>(bytes), kBufferFlags);
    }
    SOL_DEV FetchUniform(const SceneView& s, const float4* lds_nodes)
        : nodes(rsrc(s.nodes, s.num_nodes * 64u)), tris(rsrc(s.tris, s.num_tris * 48u)), lds(lds_nodes),
          top(static_cast<int>(s.top_nodes < static_cast<uint32_t>(kTopNodesMax) ? s.top_nodes : kTopNodesMax)) {}

    SOL_DEV static float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(off), 0, 0);
        return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
    }
    SOL_DEV static float4 ld3(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        auto v = __builtin_amdgcn_raw_buffer_load_b96(r, static_cast<int>(off), 0, 0);
        return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.0f);
    }
    SOL_DEV void node(int cur, float4& a, float4& b, float4& c, float4& d) const {
        if (cur < top) {
            a = lds[cur * 4 + 0];
            b = lds[cur * 4 + 1];
            c = lds[cur * 4 + 2];
            d = lds[cur * 4 + 3];
        } else {
            uint32_t o = static_cast<uint32_t>(cur) * 64u;
            a = ld4(nodes, o);
            b = ld4(nodes, o + 16);
            c = ld4(nodes, o + 32);
            d = ld4(nodes, o + 48);
        }
    }
    // v0 and the original triangle index (w), v1, v2
    SOL_DEV void tri(uint32_t i, float4& p0, float4& p1, float4& p2) const {
        uint32_t o = i * 48u;
        p0 = ld4(tris, o);
        p1 = ld4(tris, o + 16);   // w: the triangle's kind word
        p2 = ld3(tris, o + 32);
    }
};

// Node and triangle fetch of a scene that may differ between the lanes of a wave (batches):
// plain global loads (pointers in the global address space, so no flat loads).
struct FetchGlobal {
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(1))) v4f* gptr;
    gptr nodes, tris;

    SOL_DEV explicit FetchGlobal(const SceneView& s) : nodes((gptr)(s.nodes)), tris((gptr)(s.tris)) {}
    SOL_DEV static float4 f4(v4f v) { return make_float4(v.x, v.y, v.z, v.w); }
    SOL_DEV void node(int cur, float4& a, float4& b, float4& c, float4& d) const {
        gptr p = nodes + cur * 4;
        a = f4(p[0]);
        b = f4(p[1]);
        c = f4(p[2]);
        d = f4(p[3]);
    }
    SOL_DEV void tri(uint32_t i, float4& p0, float4& p1, float4& p2) const {
        gptr p = tris + i * 3;
        p0 = f4(p[0]);
        p1 = f4(p[1]);
        p2 = f4(p[2]);
    }
};

// State of one lane's traversal between steps.
struct TraversalState {
    int node;        // next node (>= 0 inner, < 0 leaf reference)
    int sp;          // stack depth
    float tmax;      // shrinks with closest hits
    HitRecord hit;

    SOL_DEV void start(float t_max) {
        node = 0;
        sp = 0;
        tmax = t_max;
        hit.slot = -1;
        hit.prim = -1;
    }
};

enum : int { kStepContinue = 0, kStepDone = 1, kStepOccluded = 2 };

// The hit's leaf-order triangle index, with the triangle's kind (the w word of its second vertex
// row) packed above it when PACK_KIND (the wavefront traversal kernel: its hit sorting then needs
// no per-triangle table).
template <bool PACK_KIND>
SOL_DEV int hit_slot(uint32_t index, const float4& p1) {
    return static_cast<int>(PACK_KIND ? (index | (__float_as_uint(p1.w) << kHitKindShift)) : index);
}

template <int BLOCK, int LDS_STACK>
SOL_DEV bool stack_pop(const Stack& st, TraversalState& ts) {
    if (ts.sp == 0) return false;
    --ts.sp;
    if (ts.sp < LDS_STACK) ts.node = st.lds[ts.sp * BLOCK];
    else ts.node = st.spill[(ts.sp - LDS_STACK) * st.spill_stride];
    return true;
}

// One round of a lane's traversal: descend through inner nodes (both child boxes; the nearer hit
// child next, the farther pushed; pop when neither is hit) until a leaf, test the leaf's
// triangles, pop. Returns kStepDone when the stack is empty (closest hit in ts.hit, if any),
// kStepOccluded when an any-hit query found a hit, else kStepContinue. Descending first and
// testing leaves after keeps the lanes of a wave in the same code (inner nodes, then leaves).
template <int BLOCK, int LDS_STACK, bool PACK_KIND = false, typename Fetch>
SOL_DEV int traverse_step(const Fetch& f, const Stack& st, const RayPre& r, float tmin, bool any, TraversalState& ts) {
    while (ts.node >= 0) {
        float4 a, b, c, d;
        f.node(ts.node, a, b, c, d);
        float t0, t1;
        bool h0 = intersect_box(r, a.x, a.y, a.z, a.w, c.x, c.y, tmin, ts.tmax, t0);
        bool h1 = intersect_box(r, b.x, b.y, b.z, b.w, c.z, c.w, tmin, ts.tmax, t1);
        int c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
        if (h0 && h1) {
            int near = c0, far = c1;
            if (t1 < t0) { near = c1; far = c0; }
            if (ts.sp < LDS_STACK) {
                st.lds[ts.sp * BLOCK] = far;
            } else if (ts.sp < st.capacity) {
                st.spill[(ts.sp - LDS_STACK) * st.spill_stride] = far;
            } else {
                *st.overflow = 1;
                --ts.sp;
            }
            ++ts.sp;
            ts.node = near;
        } else if (h0) {
            ts.node = c0;
        } else if (h1) {
            ts.node = c1;
        } else if (!stack_pop<BLOCK, LDS_STACK>(st, ts)) {
            return kStepDone;
        }
    }
    uint32_t first, count;
    decode_leaf(ts.node, first, count);
    for (uint32_t i = 0; i < count; ++i) {
        float4 p0, p1, p2;
        f.tri(first + i, p0, p1, p2);
        float t, u, v;
        if (intersect_triangle(r, V3{p0.x, p0.y, p0.z}, V3{p1.x, p1.y, p1.z}, V3{p2.x, p2.y, p2.z}, t, u, v)) {
            int prim = __float_as_int(p0.w);
            if (any) {
                if (t > tmin && t < ts.tmax) return kStepOccluded;
            } else if (t > tmin && (t < ts.tmax || (t == ts.tmax && ts.hit.prim >= 0 && prim < ts.hit.prim))) {
                ts.tmax = t;
                ts.hit.slot = hit_slot<PACK_KIND>(first + i, p1);
                ts.hit.prim = prim;
                ts.hit.t = t;
                ts.hit.u = u;
                ts.hit.v = v;
            }
        }
    }
    return stack_pop<BLOCK, LDS_STACK>(st, ts) ? kStepContinue : kStepDone;
}

// A whole traversal in one loop (one node or one leaf per iteration). The kernels that run whole
// paths per lane use this form: it compiles to fewer registers there than looping over
// traverse_step.
template <bool ANY, int BLOCK, int LDS_STACK, typename Fetch>
SOL_DEV bool traverse_with(const Fetch& f, const Stack& st, const RayPre& r, float tmin, float tmax, HitRecord& hit) {
    int sp = 0;
    int cur = 0;
    hit.slot = -1;
    hit.prim = -1;
    while (true) {
        if (cur >= 0) {
            float4 a, b, c, d;
            f.node(cur, a, b, c, d);
            float t0, t1;
            bool h0 = intersect_box(r, a.x, a.y, a.z, a.w, c.x, c.y, tmin, tmax, t0);
            bool h1 = intersect_box(r, b.x, b.y, b.z, b.w, c.z, c.w, tmin, tmax, t1);
            int c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
            if (h0 && h1) {
                int near = c0, far = c1;
                if (t1 < t0) { near = c1; far = c0; }
                if (sp < LDS_STACK) {
                    st.lds[sp * BLOCK] = far;
                } else if (sp < st.capacity) {
                    st.spill[(sp - LDS_STACK) * st.spill_stride] = far;
                } else {
                    *st.overflow = 1;
                    --sp;
                }
                ++sp;
                cur = near;
                continue;
            }
            if (h0) { cur = c0; continue; }
            if (h1) { cur = c1; continue; }
        } else {
            uint32_t first, count;
            decode_leaf(cur, first, count);
            for (uint32_t i = 0; i < count; ++i) {
                float4 p0, p1, p2;
                f.tri(first + i, p0, p1, p2);
                float t, u, v;
                if (intersect_triangle(r, V3{p0.x, p0.y, p0.z}, V3{p1.x, p1.y, p1.z}, V3{p2.x, p2.y, p2.z}, t, u, v)) {
                    int prim = __float_as_int(p0.w);
                    if (ANY) {
                        if (t > tmin && t < tmax) return true;
                    } else if (t > tmin && (t < tmax || (t == tmax && hit.prim >= 0 && prim < hit.prim))) {
                        tmax = t;
                        hit.slot = static_cast<int>(first + i);
                        hit.prim = prim;
                        hit.t = t;
                        hit.u = u;
                        hit.v = v;
                    }
                }
            }
        }
        if (sp == 0) break;
        --sp;
        if (sp < LDS_STACK) cur = st.lds[sp * BLOCK];
        else cur = st.spill[(sp - LDS_STACK) * st.spill_stride];
    }
    return ANY ? false : hit.prim >= 0;
}

// ---- 4-wide BVH (sol_gpu_node4: 8 float4 per node) -------------------------------------------------

// Maximum number of top 4-wide nodes copied to LDS (10 KB; the layouts put four levels, at most
// 85 nodes, first). With a 16-entry stack a 256-lane workgroup of the traversal kernel then takes
// 26.6 KB of LDS, so six such workgroups - six waves per SIMD - still fit a CU's 160 KB.
constexpr int kTopNodes4Max = 80;

// float4 entries of a kernel's LDS copy of the top levels of its layout.
template <bool WIDE>
constexpr int kTopLdsFloat4 = WIDE ? kTopNodes4Max * 8 : kTopNodesMax * 4;

template <int BLOCK>
SOL_DEV void load_top_nodes4(const SceneView& s, float4* lds_nodes) {
    int n = static_cast<int>(s.top_nodes4 < static_cast<uint32_t>(kTopNodes4Max) ? s.top_nodes4 : kTopNodes4Max) * 8;
    for (int i = threadIdx.x; i < n; i += BLOCK) lds_nodes[i] = s.nodes4[i];
    __syncthreads();
}

// A 4-wide node as one ray sees it: per axis the row of near planes and the row of far planes of
// the four children (which of lo / hi is near depends on the ray's direction), and the children.
struct Node4 {
    float4 near_x, far_x, near_y, far_y, near_z, far_z;
    int4 child;
};

SOL_DEV int4 as_int4(float4 v) {
    return make_int4(__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w));
}

// Wave-uniform scene: top levels from LDS, the rest through buffer loads.
struct FetchUniform4 {
    __amdgpu_buffer_rsrc_t nodes, tris;
    const float4* lds;
    int top;

    SOL_DEV FetchUniform4(const SceneView& s, const float4* lds_nodes)
        : nodes(FetchUniform::rsrc(s.nodes4, s.num_nodes4 * 128u)), tris(FetchUniform::rsrc(s.tris, s.num_tris * 48u)),
          lds(lds_nodes),
          top(static_cast<int>(s.top_nodes4 < static_cast<uint32_t>(kTopNodes4Max) ? s.top_nodes4 : kTopNodes4Max)) {}

    // The rows are picked by address: a node is 128 bytes at a multiple of 128, so the near row
    // of an axis is at o | near and the far row at (o | near) ^ 16.
    SOL_DEV void node(int cur, const RayPre& r, Node4& n) const {
        const uint32_t o = static_cast<uint32_t>(cur) * 128u;
        const uint32_t ax = near_row(r.inv.x, o), ay = near_row(r.inv.y, o | 32u), az = near_row(r.inv.z, o | 64u);
        if (cur < top) {
            const char* p = reinterpret_cast<const char*>(lds);
            n.near_x = *reinterpret_cast<const float4*>(p + ax);
            n.far_x = *reinterpret_cast<const float4*>(p + (ax ^ 16u));
            n.near_y = *reinterpret_cast<const float4*>(p + ay);
            n.far_y = *reinterpret_cast<const float4*>(p + (ay ^ 16u));
            n.near_z = *reinterpret_cast<const float4*>(p + az);
            n.far_z = *reinterpret_cast<const float4*>(p + (az ^ 16u));
            n.child = as_int4(*reinterpret_cast<const float4*>(p + o + 96));
        } else {
            n.near_x = FetchUniform::ld4(nodes, ax);
            n.far_x = FetchUniform::ld4(nodes, ax ^ 16u);
            n.near_y = FetchUniform::ld4(nodes, ay);
            n.far_y = FetchUniform::ld4(nodes, ay ^ 16u);
            n.near_z = FetchUniform::ld4(nodes, az);
            n.far_z = FetchUniform::ld4(nodes, az ^ 16u);
            n.child = as_int4(FetchUniform::ld4(nodes, o + 96));
        }
    }
    SOL_DEV void tri(uint32_t i, float4& p0, float4& p1, float4& p2) const {
        uint32_t o = i * 48u;
        p0 = FetchUniform::ld4(tris, o);
        p1 = FetchUniform::ld4(tris, o + 16);   // w: the triangle's kind word
        p2 = FetchUniform::ld3(tris, o + 32);
    }
};

// Per-lane scene: global loads.
struct FetchGlobal4 {
    typedef FetchGlobal::v4f v4f;
    typedef FetchGlobal::gptr gptr;
    gptr nodes, tris;

    SOL_DEV explicit FetchGlobal4(const SceneView& s) : nodes((gptr)(s.nodes4)), tris((gptr)(s.tris)) {}
    SOL_DEV void node(int cur, const RayPre& r, Node4& n) const {
        gptr p = nodes + cur * 8;
        const uint32_t ix = near_row(r.inv.x, 0u) >> 4, iy = near_row(r.inv.y, 32u) >> 4, iz = near_row(r.inv.z, 64u) >> 4;
        n.near_x = FetchGlobal::f4(p[ix]); n.far_x = FetchGlobal::f4(p[ix ^ 1u]);
        n.near_y = FetchGlobal::f4(p[iy]); n.far_y = FetchGlobal::f4(p[iy ^ 1u]);
        n.near_z = FetchGlobal::f4(p[iz]); n.far_z = FetchGlobal::f4(p[iz ^ 1u]);
        n.child = as_int4(FetchGlobal::f4(p[6]));
    }
    SOL_DEV void tri(uint32_t i, float4& p0, float4& p1, float4& p2) const {
        gptr p = tris + i * 3;
        p0 = FetchGlobal::f4(p[0]);
        p1 = FetchGlobal::f4(p[1]);
        p2 = FetchGlobal::f4(p[2]);
    }
};

// Per-lane scenes (batches) with the wave's LDS copy of the first nodes of one scene: a lane of
// that scene reads its nodes below `top` from the copy (rows picked by address as in
// FetchUniform4), every other node and every triangle through global loads.
struct FetchMulti4 {
    FetchGlobal4 g;
    const float4* lds;   // the wave's copy
    int top;             // nodes [0, top) of this lane's scene are in the copy (0: not this scene)

    SOL_DEV void node(int cur, const RayPre& r, Node4& n) const {
        if (cur < top) {
            const uint32_t o = static_cast<uint32_t>(cur) * 128u;
            const uint32_t ax = near_row(r.inv.x, o), ay = near_row(r.inv.y, o | 32u), az = near_row(r.inv.z, o | 64u);
            const char* p = reinterpret_cast<const char*>(lds);
            n.near_x = *reinterpret_cast<const float4*>(p + ax);
            n.far_x = *reinterpret_cast<const float4*>(p + (ax ^ 16u));
            n.near_y = *reinterpret_cast<const float4*>(p + ay);
            n.far_y = *reinterpret_cast<const float4*>(p + (ay ^ 16u));
            n.near_z = *reinterpret_cast<const float4*>(p + az);
            n.far_z = *reinterpret_cast<const float4*>(p + (az ^ 16u));
            n.child = as_int4(*reinterpret_cast<const float4*>(p + o + 96));
        } else {
            g.node(cur, r, n);
        }
    }
    SOL_DEV void tri(uint32_t i, float4& p0, float4& p1, float4& p2) const { g.tri(i, p0, p1, p2); }
};

template <int BLOCK, int LDS_STACK>
SOL_DEV void stack_push(const Stack& st, TraversalState& ts, int v) {
    if (ts.sp < LDS_STACK) {
        st.lds[ts.sp * BLOCK] = v;
    } else if (ts.sp < st.capacity) {
        st.spill[(ts.sp - LDS_STACK) * st.spill_stride] = v;
    } else {
        *st.overflow = 1;
        --ts.sp;
    }
    ++ts.sp;
}

// The four child boxes of a 4-wide node against the ray: entry distances, hits (an empty slot never
// hits) and the child references.
SOL_DEV void test_node4(const RayPre& r, const Node4& n, float tmax, float& t0, float& t1, float& t2, float& t3,
                        bool& h0, bool& h1, bool& h2, bool& h3, int& c0, int& c1, int& c2, int& c3) {
    h0 = intersect_box_nf(r, n.near_x.x, n.far_x.x, n.near_y.x, n.far_y.x, n.near_z.x, n.far_z.x, tmax, t0);
    h1 = intersect_box_nf(r, n.near_x.y, n.far_x.y, n.near_y.y, n.far_y.y, n.near_z.y, n.far_z.y, tmax, t1);
    h2 = intersect_box_nf(r, n.near_x.z, n.far_x.z, n.near_y.z, n.far_y.z, n.near_z.z, n.far_z.z, tmax, t2);
    h3 = intersect_box_nf(r, n.near_x.w, n.far_x.w, n.near_y.w, n.far_y.w, n.near_z.w, n.far_z.w, tmax, t3);
    // an unused slot (SOL_GPU_EMPTY) holds the inverted box: every slab has near +inf, far -inf
    c0 = n.child.x;
    c1 = n.child.y;
    c2 = n.child.z;
    c3 = n.child.w;
}

// Pushes the k (0..3) farther hit children of a 4-wide node, farthest first: c3 (k == 3), c2
// (k >= 2), c1 (k >= 1), so that c1 ends on top. While the three fit in the LDS part of the stack
// the writes are unconditional: child j goes to sp + k - j if j <= k, else to sp + k, a slot at or
// above the new top that no pop reads before it is written again. Otherwise one push at a time.
template <int BLOCK, int LDS_STACK>
SOL_DEV void stack_push_children(const Stack& st, TraversalState& ts, int k, int c1, int c2, int c3) {
    if (ts.sp + 3 <= LDS_STACK) {
        const int top = ts.sp + k;
        st.lds[(top - (k >= 3 ? 3 : 0)) * BLOCK] = c3;
        st.lds[(top - (k >= 2 ? 2 : 0)) * BLOCK] = c2;
        st.lds[(top - (k >= 1 ? 1 : 0)) * BLOCK] = c1;
        ts.sp = top;
    } else {
        if (k > 2) stack_push<BLOCK, LDS_STACK>(st, ts, c3);
        if (k > 1) stack_push<BLOCK, LDS_STACK>(st, ts, c2);
        if (k > 0) stack_push<BLOCK, LDS_STACK>(st, ts, c1);
    }
}

// Any-hit order: the first hit child (in slot order) is visited next and the other hit children
// are pushed in slot order - an any-hit query stops at the first occluder it meets, so the order
// does not change its answer. Each child's stack slot is its rank among the hit children after the
// first (a prefix count); a child that is not pushed writes to the slot at the new top, which no
// pop reads before it is written again. Needs sp + 3 <= LDS_STACK.
template <int BLOCK>
SOL_DEV int push_hit_children_any(const Stack& st, TraversalState& ts, int hits, bool h0, bool h1, bool h2, bool h3,
                                  int c0, int c1, int c2, int c3) {
    const int s1 = int(h0), s2 = s1 + int(h1), s3 = s2 + int(h2);
    const int top = ts.sp + hits - 1;
    st.lds[(h1 && s1 >= 1 ? ts.sp + s1 - 1 : top) * BLOCK] = c1;
    st.lds[(h2 && s2 >= 1 ? ts.sp + s2 - 1 : top) * BLOCK] = c2;
    st.lds[(h3 && s3 >= 1 ? ts.sp + s3 - 1 : top) * BLOCK] = c3;
    ts.sp = top;
    return h0 ? c0 : (h1 ? c1 : (h2 ? c2 : c3));
}

// Orders two (distance, child) pairs by distance.
SOL_DEV void order_pair(float& ta, int& ca, float& tb, int& cb) {
    bool s = tb < ta;
    float t = s ? tb : ta;
    int c = s ? cb : ca;
    tb = s ? ta : tb;
    cb = s ? ca : cb;
    ta = t;
    ca = c;
}

// traverse_step over 4-wide nodes: the four child boxes of a node are tested, the hit children
// are ordered by entry distance (a five-comparison sorting network), the nearest is visited next
// and the others are pushed farthest first. Leaves are postponed (Aila and Laine 2009,
// speculative traversal): a lane that reaches a leaf keeps it and goes on visiting inner nodes
// while other lanes of the wave are still looking for theirs, so that fewer lanes idle in the
// node loop; then every lane tests the triangles of its leaf. The visiting order does not change
// the result: the closest hit is unique by (t, original triangle index).
// With LEAF_WAIT > 0 the node loop also ends once no more than LEAF_WAIT lanes are still looking
// for their first leaf while some lane holds one: those lanes keep their node and stack and go on
// in the next step, instead of the whole wave running node iterations for a handful of them.
// PACK_KIND packs the hit triangle's kind into the hit (hit_slot). An any-hit query visits the hit
// children in slot order (push_hit_children_any, no sorting) unless ANY_NEAR_FIRST.
template <int BLOCK, int LDS_STACK, int LEAF_WAIT = 0, bool PACK_KIND = false, bool ANY_NEAR_FIRST = false,
          typename Fetch>
SOL_DEV int traverse4_step(const Fetch& f, const Stack& st, const RayPre& r, float tmin, bool any, TraversalState& ts) {
    constexpr float kMiss = __builtin_huge_valf();
    int leaf = 0;          // postponed leaf (leaf references are negative), 0 = none
    int leaf2 = 0;         // a second postponed leaf
    int leaf3 = 0;         // a third
    bool more = true;      // ts.node still has to be visited
    while (true) {
        const unsigned long long searching = __ballot(more && leaf == 0);
        if (searching == 0) break;
        if (LEAF_WAIT > 0 && __popcll(searching) <= LEAF_WAIT && __ballot(leaf != 0) != 0) break;
        if (more && ts.node >= 0) {
            Node4 n;
            f.node(ts.node, r, n);
            float t0, t1, t2, t3;
            bool h0, h1, h2, h3;
            int c0, c1, c2, c3;
            test_node4(r, n, ts.tmax, t0, t1, t2, t3, h0, h1, h2, h3, c0, c1, c2, c3);
            const int hits = int(h0) + int(h1) + int(h2) + int(h3);
            if (hits == 0) {
                more = stack_pop<BLOCK, LDS_STACK>(st, ts);
            } else if (!ANY_NEAR_FIRST && any && ts.sp + 3 <= LDS_STACK) {
                ts.node = push_hit_children_any<BLOCK>(st, ts, hits, h0, h1, h2, h3, c0, c1, c2, c3);
            } else {
                float k0 = h0 ? t0 : kMiss, k1 = h1 ? t1 : kMiss, k2 = h2 ? t2 : kMiss, k3 = h3 ? t3 : kMiss;
                order_pair(k0, c0, k1, c1);
                order_pair(k2, c2, k3, c3);
                order_pair(k0, c0, k2, c2);
                order_pair(k1, c1, k3, c3);
                order_pair(k1, c1, k2, c2);
                stack_push_children<BLOCK, LDS_STACK>(st, ts, hits - 1, c1, c2, c3);
                ts.node = c0;
            }
        }
        if (more && ts.node < 0 && leaf3 == 0) {
            if (leaf == 0) leaf = ts.node;
            else if (leaf2 == 0) leaf2 = ts.node;
            else leaf3 = ts.node;
            more = stack_pop<BLOCK, LDS_STACK>(st, ts);
        }
    }
    if (leaf != 0) {
        uint32_t first, count, first2 = 0, count2 = 0, first3 = 0, count3 = 0;
        decode_leaf(leaf, first, count);
        if (leaf2 != 0) decode_leaf(leaf2, first2, count2);
        if (leaf3 != 0) decode_leaf(leaf3, first3, count3);
        const uint32_t total = count + count2 + count3;
        for (uint32_t j = 0; j < total; ++j) {
            const uint32_t i = j < count ? first + j : (j < count + count2 ? first2 + (j - count) : first3 + (j - count - count2));
            float4 p0, p1, p2;
            f.tri(i, p0, p1, p2);
            float t, u, v;
            if (intersect_triangle(r, V3{p0.x, p0.y, p0.z}, V3{p1.x, p1.y, p1.z}, V3{p2.x, p2.y, p2.z}, t, u, v)) {
                int prim = __float_as_int(p0.w);
                if (any) {
                    if (t > tmin && t < ts.tmax) return kStepOccluded;
                } else if (t > tmin && (t < ts.tmax || (t == ts.tmax && ts.hit.prim >= 0 && prim < ts.hit.prim))) {
                    ts.tmax = t;
                    ts.hit.slot = hit_slot<PACK_KIND>(i, p1);
                    ts.hit.prim = prim;
                    ts.hit.t = t;
                    ts.hit.u = u;
                    ts.hit.v = v;
                }
            }
        }
    }
    return more ? kStepContinue : kStepDone;
}

// A whole traversal over the 4-wide layout, looping traverse4_step.
template <bool ANY, int BLOCK, int LDS_STACK, bool ANY_NEAR_FIRST = false, typename Fetch>
SOL_DEV bool traverse4_with(const Fetch& f, const Stack& st, const RayPre& r, float tmin, float tmax, HitRecord& hit) {
    TraversalState ts;
    ts.start(tmax);
    int status;
    do {
        status = traverse4_step<BLOCK, LDS_STACK, 0, false, ANY_NEAR_FIRST>(f, st, r, tmin, ANY, ts);
    } while (status == kStepContinue);
    hit = ts.hit;
    if (ANY) return status == kStepOccluded;
    return ts.hit.prim >= 0;
}

// Traversal of a wave-uniform scene whose top levels are staged in lds_nodes.
template <bool ANY, int BLOCK, int LDS_STACK>
SOL_DEV bool traverse(const SceneView& s, const float4* lds_nodes, const Stack& st, const RayPre& r, float tmin,
                      float tmax, HitRecord& hit) {
    return traverse_with<ANY, BLOCK, LDS_STACK>(FetchUniform(s, lds_nodes), st, r, tmin, tmax, hit);
}

// Traversal of a per-lane scene (no LDS staging).
template <bool ANY, int BLOCK, int LDS_STACK>
SOL_DEV bool traverse_lane(const SceneView& s, const Stack& st, const RayPre& r, float tmin, float tmax,
                           HitRecord& hit) {
    return traverse_with<ANY, BLOCK, LDS_STACK>(FetchGlobal(s), st, r, tmin, tmax, hit);
}

} // namespace solhip
