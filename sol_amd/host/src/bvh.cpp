// CPU BVH construction: sweep SAH, reinsertion optimisation, GPU layout.
// Behaviour follows bvh v2 as sol uses it (src/triangle_mesh.cpp:119-146 of the reference):
// SweepSahBuilder, then TopologyModifier + SequentialReinsertionOptimizer. The bvh library is
// not vendored; the algorithms are written here from their published descriptions.
#include "sol/bvh.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <deque>
#include <limits>
#include <unordered_map>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <vector>

namespace sol::bvh {

// ---------------------------------------------------------------------------------------------
// Sweep SAH builder
// ---------------------------------------------------------------------------------------------

namespace {

constexpr size_t kTaskThreshold = 4096;

struct SweepBuilder {
    const BBox* bboxes;
    const Vec3f* centers;
    size_t n;
    const BuildConfig& cfg;

    std::vector<uint32_t> refs[3];   // primitive ids sorted by centre on each axis
    std::vector<uint8_t> goes_left;  // per primitive, during one partition
    std::vector<float> right_cost;   // per position, right-to-left sweep
    std::vector<uint32_t> scratch;   // per position, stable partition buffer
    std::vector<Node> nodes;
    std::atomic<uint32_t> node_count{1};

    SweepBuilder(const BBox* b, const Vec3f* c, size_t count, const BuildConfig& config)
        : bboxes(b), centers(c), n(count), cfg(config) {}

    Bvh run() {
        nodes.resize(2 * n - 1);
        #pragma omp parallel for schedule(static)
        for (int axis = 0; axis < 3; ++axis) {
            auto& r = refs[axis];
            r.resize(n);
            std::iota(r.begin(), r.end(), 0u);
            std::sort(r.begin(), r.end(), [&](uint32_t a, uint32_t b) {
                float ca = centers[a][axis], cb = centers[b][axis];
                return ca < cb || (ca == cb && a < b);
            });
        }
        goes_left.assign(n, 0);
        right_cost.assign(n, 0.0f);
        scratch.assign(n, 0);
        #pragma omp parallel
        #pragma omp single
        build(0, 0, n);
        nodes.resize(node_count.load());
        Bvh bvh;
        bvh.nodes = std::move(nodes);
        bvh.prim_indices = std::move(refs[0]);
        return bvh;
    }

    BBox range_bbox(size_t begin, size_t end) const {
        BBox box = BBox::empty();
        for (size_t i = begin; i < end; ++i) box.extend(bboxes[refs[0][i]]);
        return box;
    }

    void build(uint32_t node_id, size_t begin, size_t end) {
        Node& node = nodes[node_id];
        node.box = range_bbox(begin, end);
        size_t count = end - begin;
        auto make_leaf = [&] {
            node.index = static_cast<uint32_t>(begin);
            node.prim_count = static_cast<uint32_t>(count);
        };
        if (count <= cfg.min_leaf_size) return make_leaf();

        float best_cost = std::numeric_limits<float>::infinity();
        int best_axis = -1;
        size_t best_split = 0;
        for (int axis = 0; axis < 3; ++axis) {
            const auto& r = refs[axis];
            BBox acc = BBox::empty();
            for (size_t i = end - 1; i > begin; --i) {
                acc.extend(bboxes[r[i]]);
                right_cost[i] = acc.half_area() * static_cast<float>(end - i);
            }
            acc = BBox::empty();
            for (size_t i = begin; i + 1 < end; ++i) {
                acc.extend(bboxes[r[i]]);
                float cost = acc.half_area() * static_cast<float>(i + 1 - begin) + right_cost[i + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = i + 1;
                }
            }
        }

        float area = node.box.half_area();
        float leaf_cost = area * static_cast<float>(count) * cfg.intersection_cost;
        float split_cost = area * cfg.traversal_cost + best_cost * cfg.intersection_cost;
        if (best_axis < 0 || (count <= cfg.max_leaf_size && leaf_cost <= split_cost))
            return make_leaf();

        for (size_t i = begin; i < end; ++i)
            goes_left[refs[best_axis][i]] = i < best_split ? 1 : 0;
        for (int axis = 0; axis < 3; ++axis) {
            if (axis == best_axis) continue;
            auto& r = refs[axis];
            size_t left = begin, right = best_split;
            for (size_t i = begin; i < end; ++i) {
                uint32_t p = r[i];
                if (goes_left[p]) scratch[left++] = p; else scratch[right++] = p;
            }
            std::copy(scratch.begin() + begin, scratch.begin() + end, r.begin() + begin);
        }

        uint32_t first = node_count.fetch_add(2);
        node.index = first;
        node.prim_count = 0;
        if (count > kTaskThreshold) {
            #pragma omp task default(shared) firstprivate(first, begin, best_split)
            build(first, begin, best_split);
            build(first + 1, best_split, end);
            #pragma omp taskwait
        } else {
            build(first, begin, best_split);
            build(first + 1, best_split, end);
        }
    }
};

} // namespace

Bvh build_sweep_sah(const BBox* bboxes, const Vec3f* centers, size_t count, const BuildConfig& config) {
    if (count == 0) throw std::invalid_argument("cannot build a BVH over zero primitives");
    if (config.max_leaf_size < 1 || config.max_leaf_size > 8)
        throw std::invalid_argument("max_leaf_size must be in [1, 8]");
    SweepBuilder builder(bboxes, centers, count, config);
    return builder.run();
}

uint32_t Bvh::depth() const {
    if (nodes.empty()) return 0;
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 1u}};
    while (!stack.empty()) {
        auto [id, d] = stack.back();
        stack.pop_back();
        best = std::max(best, d);
        if (!nodes[id].is_leaf()) {
            stack.push_back({nodes[id].index, d + 1});
            stack.push_back({nodes[id].index + 1, d + 1});
        }
    }
    return best;
}

uint32_t Bvh::leaf_count() const {
    uint32_t c = 0;
    for (auto& nd : nodes) c += nd.is_leaf() ? 1 : 0;
    return c;
}

float Bvh::sah_cost(float traversal_cost, float intersection_cost) const {
    double cost = 0;
    for (auto& nd : nodes)
        cost += nd.box.half_area() * (nd.is_leaf() ? intersection_cost * nd.prim_count : traversal_cost);
    return static_cast<float>(cost / nodes[0].box.half_area());
}

// ---------------------------------------------------------------------------------------------
// Reinsertion optimiser
// ---------------------------------------------------------------------------------------------

namespace {

struct WNode {
    BBox box;
    int32_t parent = -1, left = -1, right = -1;
    uint32_t first = 0, count = 0;
    bool is_leaf() const { return count != 0; }
};

struct Reinserter {
    std::vector<WNode> w;
    int32_t root = 0;

    explicit Reinserter(const Bvh& bvh) {
        w.resize(bvh.nodes.size());
        for (size_t i = 0; i < bvh.nodes.size(); ++i) {
            const Node& nd = bvh.nodes[i];
            w[i].box = nd.box;
            if (nd.is_leaf()) {
                w[i].first = nd.index;
                w[i].count = nd.prim_count;
            } else {
                w[i].left = static_cast<int32_t>(nd.index);
                w[i].right = static_cast<int32_t>(nd.index + 1);
                w[nd.index].parent = static_cast<int32_t>(i);
                w[nd.index + 1].parent = static_cast<int32_t>(i);
            }
        }
    }

    void refit_up(int32_t a) {
        while (a >= 0) {
            w[a].box = merge(w[w[a].left].box, w[w[a].right].box);
            a = w[a].parent;
        }
    }

    void replace_child(int32_t parent, int32_t old_child, int32_t new_child) {
        if (parent < 0) { root = new_child; w[new_child].parent = -1; return; }
        if (w[parent].left == old_child) w[parent].left = new_child; else w[parent].right = new_child;
        w[new_child].parent = parent;
    }

    // Best sibling for the detached subtree c: minimises the area of the new parent plus the
    // growth of every ancestor (branch and bound, best first).
    int32_t search(int32_t c) const {
        const BBox& cb = w[c].box;
        float area_c = cb.half_area();
        float best = std::numeric_limits<float>::infinity();
        int32_t best_x = root;
        using Entry = std::pair<float, int32_t>;
        std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> heap;
        heap.push({0.0f, root});
        while (!heap.empty()) {
            auto [induced, x] = heap.top();
            heap.pop();
            if (induced + area_c >= best) break;
            float merged = merge(w[x].box, cb).half_area();
            float total = induced + merged;
            if (total < best) { best = total; best_x = x; }
            if (!w[x].is_leaf()) {
                float child_induced = total - w[x].box.half_area();
                if (child_induced + area_c < best) {
                    heap.push({child_induced, w[x].left});
                    heap.push({child_induced, w[x].right});
                }
            }
        }
        return best_x;
    }

    // Removes c (and its parent p), then reinserts c next to the best sibling, reusing p.
    // Returns true if c ends up somewhere else.
    bool reinsert(int32_t c) {
        int32_t p = w[c].parent;
        if (p < 0) return false;
        int32_t s = w[p].left == c ? w[p].right : w[p].left;
        int32_t g = w[p].parent;
        replace_child(g, p, s);
        if (g >= 0) refit_up(g);

        int32_t x = search(c);
        int32_t gx = w[x].parent;
        replace_child(gx, x, p);
        w[p].left = x;
        w[p].right = c;
        w[x].parent = p;
        w[c].parent = p;
        refit_up(p);
        return x != s;
    }

    double area_sum() const {
        double a = 0;
        for (auto& nd : w) if (!nd.is_leaf()) a += nd.box.half_area();
        return a;
    }

    Bvh to_bvh(const std::vector<uint32_t>& prim_indices) const {
        Bvh out;
        out.prim_indices = prim_indices;
        out.nodes.reserve(w.size());
        std::vector<int32_t> src;
        src.reserve(w.size());
        src.push_back(root);
        out.nodes.push_back({});
        for (size_t i = 0; i < src.size(); ++i) {
            const WNode& nd = w[src[i]];
            Node& o = out.nodes[i];
            o.box = nd.box;
            if (nd.is_leaf()) {
                o.index = nd.first;
                o.prim_count = nd.count;
            } else {
                o.index = static_cast<uint32_t>(src.size());
                o.prim_count = 0;
                src.push_back(nd.left);
                src.push_back(nd.right);
                out.nodes.push_back({});
                out.nodes.push_back({});
            }
        }
        return out;
    }
};

} // namespace

size_t optimize_reinsertion(Bvh& bvh, const BuildConfig& config) {
    if (config.reinsertion_iters == 0 || bvh.nodes.size() < 5) return 0;
    Reinserter r(bvh);
    size_t total_moved = 0;
    double cost = r.area_sum();
    for (uint32_t iter = 0; iter < config.reinsertion_iters; ++iter) {
        std::vector<std::pair<float, int32_t>> cand;
        cand.reserve(r.w.size());
        for (int32_t i = 0; i < static_cast<int32_t>(r.w.size()); ++i)
            if (i != r.root) cand.push_back({r.w[i].box.half_area(), i});
        size_t k = std::max<size_t>(1, static_cast<size_t>(cand.size() * config.reinsertion_batch));
        k = std::min(k, cand.size());
        std::partial_sort(cand.begin(), cand.begin() + k, cand.end(),
                          [](auto& a, auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
        size_t moved = 0;
        for (size_t i = 0; i < k; ++i) moved += r.reinsert(cand[i].second) ? 1 : 0;
        total_moved += moved;
        double new_cost = r.area_sum();
        bool small_gain = new_cost > cost * (1.0 - 1e-3);
        cost = new_cost;
        if (moved == 0 || small_gain) break;
    }
    bvh = r.to_bvh(bvh.prim_indices);
    return total_moved;
}

// ---------------------------------------------------------------------------------------------
// GPU layout
// ---------------------------------------------------------------------------------------------

namespace {

int32_t encode_child(const Bvh& bvh, uint32_t id, const std::vector<int32_t>& gpu_index) {
    const Node& nd = bvh.nodes[id];
    if (!nd.is_leaf()) return gpu_index[id];
    if (nd.prim_count > 8 || nd.index >= (1u << 28))
        throw std::invalid_argument("leaf too large for the GPU node encoding");
    uint32_t code = (nd.index << 3) | (nd.prim_count - 1);
    return ~static_cast<int32_t>(code);
}

void write_box(sol_gpu_node& g, int slot, const BBox& b) {
    float* xy = slot == 0 ? g.b0 : g.b1;
    xy[0] = b.lo.x; xy[1] = b.hi.x; xy[2] = b.lo.y; xy[3] = b.hi.y;
    g.bz[slot * 2 + 0] = b.lo.z;
    g.bz[slot * 2 + 1] = b.hi.z;
}

} // namespace

GpuLayout to_gpu_layout(const Bvh& bvh, uint32_t top_levels) {
    GpuLayout out;
    out.depth = bvh.depth();
    const Node& root = bvh.nodes[0];
    if (root.is_leaf()) {
        // A single leaf: a root whose two children are that leaf.
        sol_gpu_node g{};
        std::vector<int32_t> none;
        int32_t c = encode_child(bvh, 0, none);
        write_box(g, 0, root.box);
        write_box(g, 1, root.box);
        g.child[0] = g.child[1] = c;
        out.nodes.push_back(g);
        out.top_nodes = 1;
        out.depth = 2;
        return out;
    }
    std::vector<uint32_t> order;
    order.reserve(bvh.nodes.size() / 2 + 1);
    std::vector<uint32_t> rest;
    std::deque<std::pair<uint32_t, uint32_t>> queue{{0u, 0u}};
    while (!queue.empty()) {
        auto [id, level] = queue.front();
        queue.pop_front();
        if (level < top_levels) {
            order.push_back(id);
            for (uint32_t c = 0; c < 2; ++c) {
                uint32_t ch = bvh.nodes[id].index + c;
                if (!bvh.nodes[ch].is_leaf()) queue.push_back({ch, level + 1});
            }
        } else {
            rest.push_back(id);
        }
    }
    out.top_nodes = static_cast<uint32_t>(order.size());
    std::vector<uint32_t> stack;
    for (uint32_t r : rest) {
        stack.push_back(r);
        while (!stack.empty()) {
            uint32_t id = stack.back();
            stack.pop_back();
            order.push_back(id);
            const Node& nd = bvh.nodes[id];
            if (!bvh.nodes[nd.index + 1].is_leaf()) stack.push_back(nd.index + 1);
            if (!bvh.nodes[nd.index].is_leaf()) stack.push_back(nd.index);
        }
    }
    std::vector<int32_t> gpu_index(bvh.nodes.size(), -1);
    for (size_t i = 0; i < order.size(); ++i) gpu_index[order[i]] = static_cast<int32_t>(i);
    out.nodes.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        const Node& nd = bvh.nodes[order[i]];
        sol_gpu_node& g = out.nodes[i];
        g = sol_gpu_node{};
        for (int c = 0; c < 2; ++c) {
            uint32_t ch = nd.index + c;
            write_box(g, c, bvh.nodes[ch].box);
            g.child[c] = encode_child(bvh, ch, gpu_index);
        }
    }
    return out;
}

GpuLayout4 to_gpu_layout4(const Bvh& bvh, uint32_t top_levels) {
    GpuLayout4 out;
    std::vector<int32_t> none;
    auto set_child = [&](sol_gpu_node4& g, int slot, uint32_t id, int32_t ref) {
        const BBox& b = bvh.nodes[id].box;
        g.lo_x[slot] = b.lo.x; g.hi_x[slot] = b.hi.x;
        g.lo_y[slot] = b.lo.y; g.hi_y[slot] = b.hi.y;
        g.lo_z[slot] = b.lo.z; g.hi_z[slot] = b.hi.z;
        g.child[slot] = ref;
    };
    auto empty_node = [] {   // every slot unused: the inverted box, which no ray hits
        sol_gpu_node4 g{};
        const float inf = std::numeric_limits<float>::infinity();
        for (int k = 0; k < 4; ++k) {
            g.lo_x[k] = g.lo_y[k] = g.lo_z[k] = inf;
            g.hi_x[k] = g.hi_y[k] = g.hi_z[k] = -inf;
            g.child[k] = SOL_GPU_EMPTY;
        }
        return g;
    };
    const Node& root = bvh.nodes[0];
    if (root.is_leaf()) {
        sol_gpu_node4 g = empty_node();
        set_child(g, 0, 0, encode_child(bvh, 0, none));
        out.nodes.push_back(g);
        out.top_nodes = 1;
        out.depth = 2;
        return out;
    }
    // Every inner binary node either becomes a 4-wide node or is opened into the 4-wide node of
    // its parent. The leaves stay those of the binary tree, so the collapse only decides the
    // summed area of the 4-wide nodes - the expected number of node visits of a ray, up to the
    // root area - and it is chosen least (dynamic programming over the tree, bottom up):
    //   fill(x, k) = min over j of cost(l, j) + cost(r, k - j)   x's two subtrees in k slots
    //   wide(x)    = area(x) + fill(x, 4)                         x a 4-wide node
    //   cost(x, k) = min(wide(x), fill(x, k)), 0 for a leaf       subtree x in at most k slots
    const size_t n_nodes = bvh.nodes.size();
    std::vector<std::array<double, 5>> cost(n_nodes), fill(n_nodes);
    std::vector<std::array<uint8_t, 5>> split(n_nodes);   // best j of fill(x, k)
    std::vector<double> wide(n_nodes, 0.0);
    {
        std::vector<std::pair<uint32_t, bool>> stack{{0u, false}};
        while (!stack.empty()) {
            auto [x, expanded] = stack.back();
            stack.pop_back();
            const Node& nd = bvh.nodes[x];
            if (nd.is_leaf()) {
                cost[x].fill(0.0);
                continue;
            }
            if (!expanded) {
                stack.push_back({x, true});
                stack.push_back({nd.index, false});
                stack.push_back({nd.index + 1, false});
                continue;
            }
            const uint32_t l = nd.index, r = nd.index + 1;
            for (int k = 2; k <= 4; ++k) {
                fill[x][k] = HUGE_VAL;
                for (int j = 1; j < k; ++j) {
                    double c = cost[l][j] + cost[r][k - j];
                    if (c < fill[x][k]) {
                        fill[x][k] = c;
                        split[x][k] = static_cast<uint8_t>(j);
                    }
                }
            }
            wide[x] = static_cast<double>(nd.box.half_area()) + fill[x][4];
            cost[x][0] = HUGE_VAL;
            cost[x][1] = wide[x];
            for (int k = 2; k <= 4; ++k) cost[x][k] = std::min(wide[x], fill[x][k]);
        }
    }
    // The children (binary node ids) of the 4-wide node made from binary inner node id.
    auto collapse = [&](uint32_t id) {
        std::vector<uint32_t> kids;
        auto place = [&](auto&& self, uint32_t x, int k) -> void {
            const Node& nd = bvh.nodes[x];
            if (nd.is_leaf() || k == 1 || !(fill[x][k] < wide[x])) {
                kids.push_back(x);
                return;
            }
            const int j = split[x][k];
            self(self, nd.index, j);
            self(self, nd.index + 1, k - j);
        };
        const int j = split[id][4];
        place(place, bvh.nodes[id].index, j);
        place(place, bvh.nodes[id].index + 1, 4 - j);
        return kids;
    };
    // Order the 4-wide nodes (identified by their binary node id): top levels breadth-first, then
    // each remaining subtree depth-first.
    std::vector<uint32_t> order, rest;
    std::unordered_map<uint32_t, std::vector<uint32_t>> kids;
    std::deque<std::pair<uint32_t, uint32_t>> queue{{0u, 1u}};
    uint32_t depth = 1;
    while (!queue.empty()) {
        auto [id, level] = queue.front();
        queue.pop_front();
        if (level <= top_levels) {
            order.push_back(id);
            auto k = collapse(id);
            for (uint32_t c : k)
                if (!bvh.nodes[c].is_leaf()) queue.push_back({c, level + 1});
            kids.emplace(id, std::move(k));
        } else {
            rest.push_back(id);
        }
    }
    out.top_nodes = static_cast<uint32_t>(order.size());
    std::vector<uint32_t> stack;
    for (uint32_t r : rest) {
        stack.push_back(r);
        while (!stack.empty()) {
            uint32_t id = stack.back();
            stack.pop_back();
            order.push_back(id);
            auto k = collapse(id);
            for (size_t j = k.size(); j-- > 0;)
                if (!bvh.nodes[k[j]].is_leaf()) stack.push_back(k[j]);
            kids.emplace(id, std::move(k));
        }
    }
    std::vector<int32_t> gpu_index(bvh.nodes.size(), -1);
    for (size_t i = 0; i < order.size(); ++i) gpu_index[order[i]] = static_cast<int32_t>(i);
    out.nodes.resize(order.size());
    std::vector<uint32_t> level_of(order.size(), 0);
    level_of[0] = 1;
    for (size_t i = 0; i < order.size(); ++i) {
        sol_gpu_node4 g = empty_node();
        const auto& k = kids.at(order[i]);
        for (size_t j = 0; j < k.size(); ++j) {
            set_child(g, static_cast<int>(j), k[j], encode_child(bvh, k[j], gpu_index));
            if (!bvh.nodes[k[j]].is_leaf()) level_of[gpu_index[k[j]]] = level_of[i] + 1;
            else depth = std::max(depth, level_of[i] + 1);
        }
        out.nodes[i] = g;
    }
    out.depth = depth;
    return out;
}

} // namespace sol::bvh
