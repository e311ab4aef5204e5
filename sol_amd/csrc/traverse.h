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

struct HitRecord {
    int slot;   // leaf-order triangle index, -1 = miss
    int prim;   // original triangle index
    float t, u, v;
};

struct Stack {
    int* lds;          // this lane's first entry
    int* spill;        // this lane's first spill entry
    int spill_stride;  // distance between consecutive spill entries
    int capacity;      // LDS + spill entries
    int* overflow;     // set to 1 if a push would exceed the capacity
};

template <int BLOCK>
SOL_DEV void load_top_nodes(const SceneView& s, float4* lds_nodes) {
    int n = static_cast<int>(s.top_nodes < static_cast<uint32_t>(kTopNodesMax) ? s.top_nodes : kTopNodesMax) * 4;
    for (int i = threadIdx.x; i < n; i += BLOCK) lds_nodes[i] = s.nodes[i];
    __syncthreads();
}

template <bool ANY, int BLOCK, int LDS_STACK>
SOL_DEV bool traverse(const SceneView& s, const float4* lds_nodes, const Stack& st, const RayPre& r,
                      float tmin, float tmax, HitRecord& hit) {
    const int top = static_cast<int>(s.top_nodes < static_cast<uint32_t>(kTopNodesMax) ? s.top_nodes : kTopNodesMax);
    int sp = 0;
    int cur = 0;
    hit.slot = -1;
    hit.prim = -1;
    while (true) {
        if (cur >= 0) {
            float4 a, b, c, d;
            if (cur < top) {
                a = lds_nodes[cur * 4 + 0];
                b = lds_nodes[cur * 4 + 1];
                c = lds_nodes[cur * 4 + 2];
                d = lds_nodes[cur * 4 + 3];
            } else {
                const float4* p = s.nodes + cur * 4;
                a = p[0];
                b = p[1];
                c = p[2];
                d = p[3];
            }
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
                const float4* tp = s.tris + (first + i) * 3;
                float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
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
        cur = sp < LDS_STACK ? st.lds[sp * BLOCK] : st.spill[(sp - LDS_STACK) * st.spill_stride];
    }
    return ANY ? false : hit.prim >= 0;
}

} // namespace solhip
