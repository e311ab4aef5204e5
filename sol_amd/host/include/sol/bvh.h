// CPU BVH builder in the style of bvh v2 (the library sol's TriangleMesh uses,
// src/triangle_mesh.cpp:119-146 of the reference): a top-down sweep-SAH builder (primitives
// pre-sorted on each axis, every split position of every axis evaluated), then the reinsertion
// optimiser of Meister and Bittner ("Parallel Reinsertion for Bounding Volume Hierarchy
// Optimization", 2018) in its sequential form: take the largest-area nodes, remove each from
// the tree and reinsert it where the summed node area grows least.
#ifndef SOL_BVH_H
#define SOL_BVH_H

#include <cstdint>
#include <vector>

#include "sol/math.h"
#include "sol_types.h"

namespace sol::bvh {

struct Node {
    BBox box;
    uint32_t index = 0;       // first child (inner) or first entry of prim_indices (leaf)
    uint32_t prim_count = 0;  // 0 = inner node
    bool is_leaf() const { return prim_count != 0; }
};

struct Bvh {
    std::vector<Node> nodes;             // nodes[0] is the root; siblings are adjacent
    std::vector<uint32_t> prim_indices;  // leaf ranges index this array

    uint32_t depth() const;              // nodes on the longest root-to-leaf path
    uint32_t leaf_count() const;
    // SAH cost normalised by the root area.
    float sah_cost(float traversal_cost, float intersection_cost) const;
};

struct BuildConfig {
    uint32_t min_leaf_size = 1;
    uint32_t max_leaf_size = 8;
    float traversal_cost = 1.0f;
    float intersection_cost = 1.0f;
    uint32_t reinsertion_iters = 3;
    float reinsertion_batch = 0.05f;
};

Bvh build_sweep_sah(const BBox* bboxes, const Vec3f* centers, size_t count, const BuildConfig& config);
// Returns the number of nodes moved.
size_t optimize_reinsertion(Bvh& bvh, const BuildConfig& config);

// GPU layout: both children's boxes in the parent (sol_gpu_node). The first nodes are the
// top `top_levels` levels in breadth-first order, the rest follow in depth-first order.
struct GpuLayout {
    std::vector<sol_gpu_node> nodes;
    uint32_t top_nodes = 0;
    uint32_t depth = 0;
};
GpuLayout to_gpu_layout(const Bvh& bvh, uint32_t top_levels);

// The binary tree collapsed to 4-wide nodes: each inner binary node becomes a 4-wide node or is
// opened into its parent's, chosen so that the summed surface area of the 4-wide nodes is least
// (the leaves stay the binary tree's). Top levels breadth-first, the rest depth-first.
struct GpuLayout4 {
    std::vector<sol_gpu_node4> nodes;
    uint32_t top_nodes = 0;
    uint32_t depth = 0;
};
GpuLayout4 to_gpu_layout4(const Bvh& bvh, uint32_t top_levels);

} // namespace sol::bvh

#endif
