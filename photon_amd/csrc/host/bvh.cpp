// Binned SAH BVH builder (Wald 2007), restating the reference builder's decisions exactly
// (src/bvh/bvh.cpp:51-147 for the split search, :196-255 for the flattening) because the tree shape
// fixes the BVH order of the triangles, which is the triangle ID every hit reports:
//   * 32 centroid bins per axis, cost C_trav + (A_l N_l + A_r N_r) / A_node, first minimum wins
//     (axis 0..2, then bin 0..nb-2);
//   * a leaf when the best cost is not below N (unless N > leaf limit), when no axis has extent,
//     at N <= 1, or below depth 60;
//   * std::partition on the BVH order with the winning bin predicate (its permutation sets the order
//     inside leaves), falling back to the middle when one side would be empty;
//   * flattening with an explicit LIFO stack: a node's two children are numbered consecutively and the
//     right child's subtree is emitted before the left child's.
// The tree is built as linked nodes; subtrees above 20000 triangles may be built on other threads,
// which cannot change the result (each task owns a disjoint range of the order array).
#include "bvh.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <future>
#include <memory>
#include <thread>

namespace phh {

namespace {

struct TreeNode {
    Bounds box;
    int first = 0, count = 0;
    std::unique_ptr<TreeNode> kid[2];
    bool leaf() const { return !kid[0]; }
};

class SahBuilder {
public:
    SahBuilder(const BvhSettings& o, std::vector<int>& order, std::vector<Bounds>& boxes, std::vector<vec3>& centres)
        : opt_(o), order_(order), boxes_(boxes), centres_(centres) {}

    int thread_budget = 1;

    std::unique_ptr<TreeNode> split(int first, int count, int depth) {
        auto node = std::make_unique<TreeNode>();
        Bounds cbox;
        for (int i = first; i < first + count; i++) {
            node->box.include(boxes_[order_[i]]);
            cbox.include(centres_[order_[i]]);
        }
        node->first = first;
        node->count = count;
        if (count <= 1 || depth > 60) return node;

        const int nb = opt_.bin_count;
        float best = kInf;
        int best_axis = -1, best_bin = -1;
        const float leaf_cost = float(count);
        const float node_area = node->box.surface_area();
        for (int axis = 0; axis < 3; axis++) {
            float lo = cbox.low[axis], hi = cbox.high[axis];
            if (hi - lo < 1e-12f) continue;
            float scale = float(nb) / (hi - lo);
            Bounds bin_box[64];
            int bin_n[64] = {};
            for (int i = first; i < first + count; i++) {
                int t = order_[i];
                int b = std::min(nb - 1, int((centres_[t][axis] - lo) * scale));
                bin_n[b]++;
                bin_box[b].include(boxes_[t]);
            }
            // Suffix sweep: area and count right of each plane.
            float suffix_area[64];
            int suffix_n[64];
            Bounds acc;
            int n = 0;
            for (int b = nb - 1; b > 0; b--) {
                acc.include(bin_box[b]);
                n += bin_n[b];
                suffix_area[b] = acc.surface_area();
                suffix_n[b] = n;
            }
            // Prefix sweep: evaluate each plane.
            acc = Bounds();
            n = 0;
            for (int b = 0; b < nb - 1; b++) {
                acc.include(bin_box[b]);
                n += bin_n[b];
                int nr = suffix_n[b + 1];
                if (n == 0 || nr == 0) continue;
                float cost = opt_.visit_cost + (acc.surface_area() * float(n) + suffix_area[b + 1] * float(nr)) / node_area;
                if (cost < best) {
                    best = cost;
                    best_axis = axis;
                    best_bin = b;
                }
            }
        }
        const bool forced = count > opt_.leaf_limit;
        if (best_axis < 0) {
            if (!forced) return node;
        } else if (best >= leaf_cost && !forced) {
            return node;
        }

        int mid = first + count / 2;
        if (best_axis >= 0) {
            float lo = cbox.low[best_axis], hi = cbox.high[best_axis];
            float scale = float(nb) / (hi - lo);
            const int axis = best_axis, cut = best_bin;
            int* p = std::partition(order_.data() + first, order_.data() + first + count, [&](int t) {
                return std::min(nb - 1, int((centres_[t][axis] - lo) * scale)) <= cut;
            });
            mid = int(p - order_.data());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        int nl = mid - first, nr = count - nl;
        if (count > 20000 && busy_.fetch_add(1) < thread_budget - 1) {
            auto left = std::async(std::launch::async, [&] { return split(first, nl, depth + 1); });
            node->kid[1] = split(mid, nr, depth + 1);
            node->kid[0] = left.get();
            busy_.fetch_sub(1);
        } else {
            if (count > 20000) busy_.fetch_sub(1);
            node->kid[0] = split(first, nl, depth + 1);
            node->kid[1] = split(mid, nr, depth + 1);
        }
        return node;
    }

private:
    const BvhSettings& opt_;
    std::vector<int>& order_;
    std::vector<Bounds>& boxes_;
    std::vector<vec3>& centres_;
    std::atomic<int> busy_{0};
};

void clear_side(ph_bvh_node& n, int side) {
    float* lo = side == 0 ? n.lmin : n.rmin;
    float* hi = side == 0 ? n.lmax : n.rmax;
    for (int k = 0; k < 3; k++) {
        lo[k] = kInf;
        hi[k] = -kInf;
    }
    if (side == 0) {
        n.left = -1;
        n.lcount = 0;
    } else {
        n.right = -1;
        n.rcount = 0;
    }
}

void write_side(ph_bvh_node& n, int side, const TreeNode& c, int child_index) {
    float* lo = side == 0 ? n.lmin : n.rmin;
    float* hi = side == 0 ? n.lmax : n.rmax;
    lo[0] = c.box.low.x, lo[1] = c.box.low.y, lo[2] = c.box.low.z;
    hi[0] = c.box.high.x, hi[1] = c.box.high.y, hi[2] = c.box.high.z;
    int idx = c.leaf() ? c.first : child_index;
    int cnt = c.leaf() ? c.count : 0;
    if (side == 0) {
        n.left = idx;
        n.lcount = cnt;
    } else {
        n.right = idx;
        n.rcount = cnt;
    }
}

}  // namespace

BvhTree build_tree(const std::vector<vec3>& verts, const BvhSettings& opt) {
    auto t0 = std::chrono::steady_clock::now();
    BvhTree out;
    const int n = int(verts.size() / 3);
    out.order.resize(size_t(n));
    for (int i = 0; i < n; i++) out.order[size_t(i)] = i;
    std::vector<Bounds> boxes(static_cast<size_t>(n));
    std::vector<vec3> centres(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) {
        Bounds b;
        b.include(verts[size_t(3 * i)]);
        b.include(verts[size_t(3 * i + 1)]);
        b.include(verts[size_t(3 * i + 2)]);
        boxes[size_t(i)] = b;
        centres[size_t(i)] = b.midpoint();
    }
    if (n == 0) {
        ph_bvh_node root{};
        clear_side(root, 0);
        clear_side(root, 1);
        out.nodes.push_back(root);
        return out;
    }
    SahBuilder builder(opt, out.order, boxes, centres);
    builder.thread_budget = opt.workers > 0 ? opt.workers : int(std::max(1u, std::thread::hardware_concurrency()));
    std::unique_ptr<TreeNode> root = builder.split(0, n, 0);
    out.bounds = root->box;

    BvhReport& st = out.report;
    const float root_area = std::max(out.bounds.surface_area(), 1e-20f);
    out.nodes.push_back(ph_bvh_node{});
    if (root->leaf()) {
        write_side(out.nodes[0], 0, *root, 0);
        clear_side(out.nodes[0], 1);
        st.nodes = 1;
        st.leaves = 1;
        st.sah = float(root->count);
    } else {
        struct Pending {
            const TreeNode* node;
            int slot;
            int depth;
        };
        std::vector<Pending> stack{{root.get(), 0, 0}};
        st.sah = opt.visit_cost;
        while (!stack.empty()) {
            Pending p = stack.back();
            stack.pop_back();
            st.nodes++;
            st.depth = std::max(st.depth, p.depth);
            for (int side = 0; side < 2; side++) {
                const TreeNode& c = *p.node->kid[side];
                int child_slot = -1;
                if (!c.leaf()) {
                    child_slot = int(out.nodes.size());
                    out.nodes.push_back(ph_bvh_node{});
                    stack.push_back({&c, child_slot, p.depth + 1});
                    st.sah += opt.visit_cost * c.box.surface_area() / root_area;
                } else {
                    st.leaves++;
                    st.sah += float(c.count) * c.box.surface_area() / root_area;
                }
                write_side(out.nodes[size_t(p.slot)], side, c, child_slot);
            }
        }
    }
    st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace phh
