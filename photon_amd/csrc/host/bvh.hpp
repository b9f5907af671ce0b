// Binned-SAH bounding volume hierarchy over triangles, flattened into the 64-byte two-box nodes the
// device traverses (ph_bvh_node). The tree it builds decides the BVH order of the triangles, and with it
// the triangle ID every hit reports, so its decisions follow the reference builder's exactly
// (src/bvh/bvh.cpp:51-255); the interface is this library's own.
#pragma once

#include <vector>

#include "hmath.hpp"
#include "photon_types.h"

namespace phh {

// An axis-aligned box; a box that has included nothing is empty (low above high).
class Bounds {
public:
    vec3 low = vec3(kInf);
    vec3 high = vec3(-kInf);

    void include(vec3 p) {
        low = vmin(low, p);
        high = vmax(high, p);
    }
    void include(const Bounds& b) {
        low = vmin(low, b.low);
        high = vmax(high, b.high);
    }
    bool empty() const { return low.x > high.x; }
    // Surface area (0 for an empty box): the SAH's probability weight.
    float surface_area() const {
        if (empty()) return 0.0f;
        const vec3 e = high - low;
        return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
    vec3 midpoint() const { return (low + high) * 0.5f; }
};

struct BvhSettings {
    int bin_count = 32;       // candidate split planes per axis (at most 64)
    int leaf_limit = 8;       // a node with more triangles is always split
    float visit_cost = 1.0f;  // cost of visiting a node, in triangle tests
    int workers = 0;          // threads for large subtrees; 0: every hardware thread
};

struct BvhReport {
    int nodes = 0, leaves = 0, depth = 0;
    float sah = 0.0f;  // expected cost of a ray through the root, in triangle tests
    double ms = 0.0;
};

struct BvhTree {
    std::vector<ph_bvh_node> nodes;
    std::vector<int> order;  // position in the BVH -> input triangle
    Bounds bounds;
    BvhReport report;
};

// `corners` holds the three vertices of each triangle, one triangle after another.
BvhTree build_tree(const std::vector<vec3>& corners, const BvhSettings& settings);

}  // namespace phh
