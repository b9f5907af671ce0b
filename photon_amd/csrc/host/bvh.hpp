// Binned-SAH BVH construction over triangles, flattened to 64-byte dual-box nodes.
#pragma once

#include <vector>

#include "hmath.hpp"
#include "photon_types.h"

namespace phh {

struct Box {
    vec3 lo = vec3(kInf);
    vec3 hi = vec3(-kInf);
    void extend(vec3 p) {
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    void extend(const Box& b) {
        lo = vmin(lo, b.lo);
        hi = vmax(hi, b.hi);
    }
    bool empty() const { return lo.x > hi.x; }
    float surface() const {
        if (empty()) return 0.0f;
        vec3 d = hi - lo;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    vec3 centre() const { return (lo + hi) * 0.5f; }
};

struct BvhOptions {
    int bins = 32;                // candidate planes per axis
    int max_leaf_size = 8;        // a leaf never holds more triangles
    float traversal_cost = 1.0f;  // relative to one triangle test
    int threads = 0;              // 0: hardware concurrency
};

struct BvhStats {
    int nodes = 0;
    int leaves = 0;
    int max_depth = 0;
    float sah_cost = 0.0f;
    double build_ms = 0.0;
};

struct BvhResult {
    std::vector<ph_bvh_node> nodes;
    std::vector<int> order;  // BVH position -> input triangle
    Box bounds;
    BvhStats stats;
};

// `verts` holds 3 positions per triangle.
BvhResult build_bvh(const std::vector<vec3>& verts, const BvhOptions& opt);

}  // namespace phh
