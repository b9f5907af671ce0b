// Binned-SAH BVH build on the device (photon_hip.h, ph_bvh_build). It produces the host builder's tree,
// node array and triangle order bit for bit (photon_amd/csrc/host/bvh.cpp, which restates the reference
// builder, src/bvh/bvh.cpp:51-255), so scenes built either way trace identically.
//
// What has to match, and how:
//   * node boxes are the host's sequential std::fmin/std::fmax folds, which keep the later operand on
//     ties: over a node's range that is the last occurrence of the extreme value, which only matters
//     for +0/-0. Folds run in position order inside a thread and are combined lower-half-first.
//   * centroid and bin boxes enter only through differences and areas, where the sign of a zero makes
//     no difference, so bins use unordered LDS atomics on the order-preserving integer image of floats.
//   * the split search (first minimum over axis 0..2, then bin 0..nb-2) and every leaf test are the
//     host's own code on the host for large nodes; for small ones a wave scans the bins and reduces
//     (cost, candidate) pairs to the same first minimum.
//   * the partition is std::partition's: the k-th element on the wrong side counted from the front
//     swaps with the k-th counted from the back; ranks come from scans, so it runs in parallel.
// Large nodes (more than kSubtree triangles) are split level by level by multi-block passes with the
// per-node decisions on the host; every subtree of at most kSubtree triangles is then built level by
// level by one block, a wave per node. The node indices of the build tree differ from the host's, but
// the flattening, also on the device, numbers nodes by tree structure only.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <vector>

#include "internal.h"

// Host and device: the large-node decisions run on the host with the same functions.
#define PH_HD __host__ __device__ inline

namespace phk {

namespace {

constexpr float kBoxInf = 1e30f;  // the host builder's empty box bound
constexpr int kMaxBins = 64;
constexpr int kChunk = 2048;    // triangles per block of the large-node passes
constexpr int kSubtree = 2048;  // subtrees of at most this many triangles are built by one block each
static_assert(kSubtree <= kChunk, "the subtree kernel's level lists hold kChunk entries");
constexpr int kThreads = 256;

// The host's std::fmin / std::fmax for finite values: the second operand unless the first is strictly
// smaller (larger).
PH_HD float hmin(float a, float b) { return a < b ? a : b; }
PH_HD float hmax(float a, float b) { return a > b ? a : b; }

struct BBox {
    float lo[3], hi[3];
};

PH_HD BBox empty_box() { return BBox{{kBoxInf, kBoxInf, kBoxInf}, {-kBoxInf, -kBoxInf, -kBoxInf}}; }

// a then b, in that order (b is the later operand).
PH_HD BBox fold(const BBox& a, const BBox& b) {
    BBox r;
    for (int k = 0; k < 3; k++) {
        r.lo[k] = hmin(a.lo[k], b.lo[k]);
        r.hi[k] = hmax(a.hi[k], b.hi[k]);
    }
    return r;
}

PH_HD float surface(const BBox& b) {
    if (b.lo[0] > b.hi[0]) return 0.0f;
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Build-tree node: a leaf while left < 0.
struct TNode {
    BBox box;
    int first, count, left, right;
};

// Order-preserving integer image of a float (for atomic min/max of values).
PH_DEV int ordered(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
PH_DEV float from_ordered(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void prep_kernel(const float* positions, const int* indices, int n, BBox* boxes, float4* centres, int* order) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    BBox b = empty_box();
    for (int k = 0; k < 3; k++) {
        const float* v = positions + 3 * size_t(indices[3 * t + k]);
        BBox p{{v[0], v[1], v[2]}, {v[0], v[1], v[2]}};
        b = fold(b, p);
    }
    boxes[t] = b;
    centres[t] = make_float4((b.lo[0] + b.hi[0]) * 0.5f, (b.lo[1] + b.hi[1]) * 0.5f, (b.lo[2] + b.hi[2]) * 0.5f, 0.0f);
    order[t] = t;
}

// Ordered fold of the block's per-thread boxes; thread t holds the fold of positions before thread t+1's.
PH_DEV BBox block_fold(BBox v, BBox* scratch) {
    for (int off = 1; off < 64; off <<= 1) {
        BBox o;
        for (int k = 0; k < 3; k++) {
            o.lo[k] = __shfl_down(v.lo[k], off);
            o.hi[k] = __shfl_down(v.hi[k], off);
        }
        if ((threadIdx.x & 63) + off < 64) v = fold(v, o);
    }
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
    __syncthreads();
    BBox r = scratch[0];
    for (int w = 1; w < kThreads / 64; w++) r = fold(r, scratch[w]);
    __syncthreads();
    return r;
}

// Node box and centroid box of the positions [a, b) of `order` (all threads call; thread t folds its
// kPer-or-fewer consecutive positions).
PH_DEV void range_boxes(const int* order, const BBox* boxes, const float4* centres, int a, int b, BBox* scratch, BBox& box, BBox& cbox) {
    const int per = (b - a + kThreads - 1) / kThreads;
    const int p0 = min(b, a + int(threadIdx.x) * per), p1 = min(b, p0 + per);
    BBox nb = empty_box(), cb = empty_box();
    for (int p = p0; p < p1; p++) {
        const int t = order[p];
        nb = fold(nb, boxes[t]);
        const float4 c = centres[t];
        cb = fold(cb, BBox{{c.x, c.y, c.z}, {c.x, c.y, c.z}});
    }
    box = block_fold(nb, scratch);
    cbox = block_fold(cb, scratch);
}

PH_DEV int bin_of(float c, float lo, float scale, int nb) { return min(nb - 1, int((c - lo) * scale)); }

// ---- large nodes: per-chunk passes ---------------------------------------------------------------------

struct Chunk {
    int node;   // index into the level's node list
    int a, b;   // positions [a, b)
};

struct Split {
    int axis, bin, nb, mid;
    float lo, scale;
};

__global__ __launch_bounds__(kThreads) void chunk_boxes_kernel(const Chunk* chunks, const int* order, const BBox* boxes, const float4* centres,
                                                                BBox* out) {
    __shared__ BBox scratch[kThreads / 64];
    const Chunk c = chunks[blockIdx.x];
    BBox box, cbox;
    range_boxes(order, boxes, centres, c.a, c.b, scratch, box, cbox);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = box;
        out[2 * blockIdx.x + 1] = cbox;
    }
}

// Per-chunk bins of the three axes: count and value box (ordered ints) per bin.
struct AxisParams {
    float lo[3], scale[3];
    int valid[3], nb;
};

__global__ __launch_bounds__(kThreads) void chunk_bins_kernel(const Chunk* chunks, const AxisParams* params, const int* order, const BBox* boxes,
                                                               const float4* centres, int* out) {
    __shared__ int cnt[3][kMaxBins];
    __shared__ int bx[3][kMaxBins][6];
    const Chunk c = chunks[blockIdx.x];
    const AxisParams ap = params[c.node];
    for (int i = threadIdx.x; i < 3 * kMaxBins; i += kThreads) {
        cnt[i / kMaxBins][i % kMaxBins] = 0;
        for (int k = 0; k < 3; k++) {
            bx[i / kMaxBins][i % kMaxBins][k] = ordered(kBoxInf);
            bx[i / kMaxBins][i % kMaxBins][3 + k] = ordered(-kBoxInf);
        }
    }
    __syncthreads();
    for (int p = c.a + threadIdx.x; p < c.b; p += kThreads) {
        const int t = order[p];
        const float4 ce = centres[t];
        const float cc[3] = {ce.x, ce.y, ce.z};
        const BBox tb = boxes[t];
        for (int axis = 0; axis < 3; axis++) {
            if (!ap.valid[axis]) continue;
            const int b = bin_of(cc[axis], ap.lo[axis], ap.scale[axis], ap.nb);
            atomicAdd(&cnt[axis][b], 1);
            for (int k = 0; k < 3; k++) {
                atomicMin(&bx[axis][b][k], ordered(tb.lo[k]));
                atomicMax(&bx[axis][b][3 + k], ordered(tb.hi[k]));
            }
        }
    }
    __syncthreads();
    int* o = out + size_t(blockIdx.x) * 3 * kMaxBins * 7;
    for (int i = threadIdx.x; i < 3 * kMaxBins; i += kThreads) {
        const int axis = i / kMaxBins, b = i % kMaxBins;
        o[i * 7] = cnt[axis][b];
        for (int k = 0; k < 6; k++) o[i * 7 + 1 + k] = bx[axis][b][k];
    }
}

// Exclusive scan of a block's flags (in thread order); returns the total.
PH_DEV int block_scan(int v, int* red, int& excl) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        if (w < wave) before += red[w];
        total += red[w];
    }
    __syncthreads();
    excl = before + inc - v;
    return total;
}

// Counts per chunk of the elements on the wrong side of the split: pred false left of mid, pred true
// at or right of mid.
__global__ __launch_bounds__(kThreads) void chunk_misplaced_kernel(const Chunk* chunks, const Split* splits, const int* order, const float4* centres,
                                                                    int* counts) {
    __shared__ int red[kThreads / 64];
    const Chunk c = chunks[blockIdx.x];
    const Split s = splits[c.node];
    int l = 0, r = 0;
    for (int p = c.a + threadIdx.x; s.axis >= 0 && p < c.b; p += kThreads) {
        const float4 ce = centres[order[p]];
        const float cc[3] = {ce.x, ce.y, ce.z};
        const bool pred = bin_of(cc[s.axis], s.lo, s.scale, s.nb) <= s.bin;
        l += (p < s.mid && !pred) ? 1 : 0;
        r += (p >= s.mid && pred) ? 1 : 0;
    }
    int e;
    const int tl = block_scan(l, red, e);
    const int tr = block_scan(r, red, e);
    if (threadIdx.x == 0) {
        counts[2 * blockIdx.x] = tl;
        counts[2 * blockIdx.x + 1] = tr;
    }
}

// Writes the position of each misplaced element at its rank: left ones from the front, right ones
// from the back. `base` holds per chunk the ranks of its first left and (counted from the back) its
// last right misplaced element, and the node's offset into the pair arrays.
__global__ __launch_bounds__(kThreads) void chunk_ranks_kernel(const Chunk* chunks, const Split* splits, const int* order, const float4* centres,
                                                                const int* base, int* left_pos, int* right_pos) {
    __shared__ int red[kThreads / 64];
    const Chunk c = chunks[blockIdx.x];
    const Split s = splits[c.node];
    int lbase = base[3 * blockIdx.x], rbase = base[3 * blockIdx.x + 1];
    const int off = base[3 * blockIdx.x + 2];
    for (int p0 = c.a; p0 < c.b; p0 += kThreads) {
        const int p = p0 + int(threadIdx.x);
        bool lm = false, rm = false;
        if (p < c.b) {
            const float4 ce = centres[order[p]];
            const float cc[3] = {ce.x, ce.y, ce.z};
            const bool pred = bin_of(cc[s.axis], s.lo, s.scale, s.nb) <= s.bin;
            lm = p < s.mid && !pred;
            rm = p >= s.mid && pred;
        }
        int le, re;
        const int lt = block_scan(lm ? 1 : 0, red, le);
        const int rt = block_scan(rm ? 1 : 0, red, re);
        // Right ranks run from the back: the chunk's right elements come after those of later chunks.
        if (lm) left_pos[off + lbase + le] = p;
        if (rm) right_pos[off + rbase - 1 - re] = p;
        lbase += lt;
        rbase -= rt;
    }
}

__global__ void swap_kernel(int* order, const int* left_pos, const int* right_pos, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = left_pos[k], b = right_pos[k];
    const int t = order[a];
    order[a] = order[b];
    order[b] = t;
}

// ---- small subtrees: one block each, one wave per node, level by level ----------------------------------
//
// A block builds the subtree of one root (at most kSubtree triangles) breadth first: the nodes of a level are
// shared out to the block's waves, each wave builds its node alone (boxes, bins, split, partition) and
// appends the children to the next level's list; a block barrier separates levels. The nodes of a level
// cover disjoint ranges of `order`, so the waves never touch the same elements. Node indices come from a
// global counter in completion order; the flattening numbers nodes by tree structure only.

constexpr int kSubWaves = 8;  // waves per subtree block

struct Pending {
    int node, first, count, depth;
};

// Orders the calling wave's earlier LDS and global accesses before its later ones, across lanes.
PH_DEV void wave_sync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

// Ordered fold over the wave's lanes (lane 0 first); every lane returns the total.
PH_DEV BBox wave_fold(BBox v) {
    for (int off = 1; off < 64; off <<= 1) {
        BBox o;
        for (int k = 0; k < 3; k++) {
            o.lo[k] = __shfl_down(v.lo[k], off);
            o.hi[k] = __shfl_down(v.hi[k], off);
        }
        if ((threadIdx.x & 63) + off < 64) v = fold(v, o);
    }
    BBox r;
    for (int k = 0; k < 3; k++) {
        r.lo[k] = __shfl(v.lo[k], 0);
        r.hi[k] = __shfl(v.hi[k], 0);
    }
    return r;
}

// Lanes below `lane` whose flag is set.
PH_DEV int lanes_below(uint64_t mask, int lane) { return __popcll(mask & ((uint64_t(1) << lane) - 1)); }

// One node built by one wave: record, and for an inner node the partition of its range and its two
// children appended to `next` (count in *next_count). The decisions are the host builder's; see the
// split-search notes below.
PH_DEV void build_node(const Pending& nd, int nb, int max_leaf, float traversal_cost, int* order, const BBox* boxes, const float4* centres,
                       TNode* tree, int* node_counter, int* tmp, int (*cnt)[kMaxBins], int (*bx)[kMaxBins][6], Pending* next, int* next_count) {
    const int lane = int(threadIdx.x & 63);
    const int a = nd.first, b = nd.first + nd.count;
    // Node box and centroid box: each lane folds a contiguous run of positions, then the lanes in order.
    const int per = (nd.count + 63) / 64;
    const int p0 = min(b, a + lane * per), p1 = min(b, p0 + per);
    BBox nbox = empty_box(), cb = empty_box();
    for (int p = p0; p < p1; p++) {
        const int t = order[p];
        nbox = fold(nbox, boxes[t]);
        const float4 c = centres[t];
        cb = fold(cb, BBox{{c.x, c.y, c.z}, {c.x, c.y, c.z}});
    }
    const BBox box = wave_fold(nbox), cbox = wave_fold(cb);
    TNode rec;
    rec.box = box;
    rec.first = nd.first;
    rec.count = nd.count;
    rec.left = rec.right = -1;
    if (nd.count <= 1 || nd.depth > 60) {
        if (lane == 0) tree[nd.node] = rec;
        return;
    }
    // Bins of the three axes (unordered LDS atomics: only differences and areas of bin boxes are used).
    wave_sync();  // the wave's reads of its previous node's bins come first
    for (int i = lane; i < 3 * kMaxBins; i += 64) {
        cnt[i / kMaxBins][i % kMaxBins] = 0;
        for (int k = 0; k < 3; k++) {
            bx[i / kMaxBins][i % kMaxBins][k] = ordered(kBoxInf);
            bx[i / kMaxBins][i % kMaxBins][3 + k] = ordered(-kBoxInf);
        }
    }
    wave_sync();
    float scale[3];
    bool valid[3];
    for (int axis = 0; axis < 3; axis++) {
        valid[axis] = !(cbox.hi[axis] - cbox.lo[axis] < 1e-12f);
        scale[axis] = valid[axis] ? float(nb) / (cbox.hi[axis] - cbox.lo[axis]) : 0.0f;
    }
    for (int p = a + lane; p < b; p += 64) {
        const int t = order[p];
        const float4 ce = centres[t];
        const float cc[3] = {ce.x, ce.y, ce.z};
        const BBox tb = boxes[t];
        for (int axis = 0; axis < 3; axis++) {
            if (!valid[axis]) continue;
            const int bin = bin_of(cc[axis], cbox.lo[axis], scale[axis], nb);
            atomicAdd(&cnt[axis][bin], 1);
            for (int k = 0; k < 3; k++) {
                atomicMin(&bx[axis][bin][k], ordered(tb.lo[k]));
                atomicMax(&bx[axis][bin][3 + k], ordered(tb.hi[k]));
            }
        }
    }
    wave_sync();
    // Split search. Lane j holds bin j; prefix folds give the left side of the split after bin j, suffix
    // folds the right side from bin j + 1. Min/max folds are exact in any order and the counts are
    // integers, so every candidate's cost is the value the host's sequential sweep computes; the winner
    // is the host's first minimum in (axis, bin) order, found as the smallest (cost, candidate index)
    // among candidates the host would accept (both sides non-empty, cost below its 1e30 start).
    const float node_area = surface(box);
    float best = kBoxInf;
    int best_key = 0x7fffffff, best_left_n = 0;
    for (int axis = 0; axis < 3; axis++) {
        if (!valid[axis]) continue;
        BBox v = empty_box();
        int n = 0;
        if (lane < nb) {
            n = cnt[axis][lane];
            for (int k = 0; k < 3; k++) {
                v.lo[k] = from_ordered(bx[axis][lane][k]);
                v.hi[k] = from_ordered(bx[axis][lane][3 + k]);
            }
        }
        BBox pre = v, suf = v;
        int pre_n = n, suf_n = n;
        for (int off = 1; off < 64; off <<= 1) {
            BBox up, dn;
            for (int k = 0; k < 3; k++) {
                up.lo[k] = __shfl_up(pre.lo[k], off);
                up.hi[k] = __shfl_up(pre.hi[k], off);
                dn.lo[k] = __shfl_down(suf.lo[k], off);
                dn.hi[k] = __shfl_down(suf.hi[k], off);
            }
            const int up_n = __shfl_up(pre_n, off), dn_n = __shfl_down(suf_n, off);
            if (lane >= off) {
                pre = fold(up, pre);
                pre_n += up_n;
            }
            if (lane + off < 64) {
                suf = fold(suf, dn);
                suf_n += dn_n;
            }
        }
        // Right side of the split after bin `lane`: the suffix starting at lane + 1.
        BBox right;
        for (int k = 0; k < 3; k++) {
            right.lo[k] = __shfl_down(suf.lo[k], 1);
            right.hi[k] = __shfl_down(suf.hi[k], 1);
        }
        const int right_n = __shfl_down(suf_n, 1);
        float cost = kBoxInf;
        bool ok = lane < nb - 1 && pre_n != 0 && right_n != 0;
        if (ok) {
            cost = traversal_cost + (surface(pre) * float(pre_n) + surface(right) * float(right_n)) / node_area;
            ok = cost < kBoxInf;
        }
        float c = ok ? cost : kBoxInf;
        int key = ok ? axis * 64 + lane : 0x7fffffff;
        int ln = pre_n;
        for (int off = 32; off > 0; off >>= 1) {
            const float oc = __shfl_xor(c, off);
            const int ok2 = __shfl_xor(key, off), on = __shfl_xor(ln, off);
            if (oc < c || (oc == c && ok2 < key)) {
                c = oc;
                key = ok2;
                ln = on;
            }
        }
        if (key != 0x7fffffff && (c < best || (c == best && key < best_key))) {
            best = c;
            best_key = key;
            best_left_n = ln;
        }
    }
    const int best_axis = best_key == 0x7fffffff ? -1 : best_key / 64;
    const int best_bin = best_key == 0x7fffffff ? -1 : best_key % 64;
    const bool forced = nd.count > max_leaf;
    const bool leaf = best_axis < 0 ? !forced : (best >= float(nd.count) && !forced);
    if (leaf) {
        if (lane == 0) tree[nd.node] = rec;
        return;
    }
    int mid = a + (best_axis >= 0 ? best_left_n : nd.count / 2);
    if (best_axis >= 0) {
        // std::partition: the k-th element on the wrong side from the front swaps with the k-th from the
        // back. Left misplaced positions go to tmp[a + k], right ones (front to back) to tmp[b - 1 - j].
        const float lo = cbox.lo[best_axis], sc = scale[best_axis];
        int nl = 0, nr = 0;
        for (int p0w = a; p0w < b; p0w += 64) {
            const int p = p0w + lane;
            bool lm = false, rm = false;
            if (p < b) {
                const float4 ce = centres[order[p]];
                const float cc[3] = {ce.x, ce.y, ce.z};
                const bool pred = bin_of(cc[best_axis], lo, sc, nb) <= best_bin;
                lm = p < mid && !pred;
                rm = p >= mid && pred;
            }
            const uint64_t lmask = __ballot(lm), rmask = __ballot(rm);
            if (lm) tmp[a + nl + lanes_below(lmask, lane)] = p;
            if (rm) tmp[b - 1 - (nr + lanes_below(rmask, lane))] = p;
            nl += __popcll(lmask);
            nr += __popcll(rmask);
        }
        wave_sync();
        // nl == nr (both count the elements that change sides); the k-th from the back of the right list
        // is tmp[b - nr + k].
        const int pairs = min(nl, nr);
        for (int k = lane; k < pairs; k += 64) {
            const int pa = tmp[a + k], pb = tmp[b - nr + k];
            const int t = order[pa];
            order[pa] = order[pb];
            order[pb] = t;
        }
        wave_sync();
        if (mid == a || mid == b) mid = a + nd.count / 2;
    }
    if (lane == 0) {
        const int kids = atomicAdd(node_counter, 2);
        rec.left = kids;
        rec.right = kids + 1;
        tree[nd.node] = rec;
        const int slot = atomicAdd(next_count, 2);
        next[slot] = Pending{kids, a, mid - a, nd.depth + 1};
        next[slot + 1] = Pending{kids + 1, mid, b - mid, nd.depth + 1};
    }
}

// lists: per block two level lists of kChunk entries. A level of a subtree of at most kSubtree <= kChunk
// triangles has at most that many nodes (their ranges are disjoint and non-empty).
__global__ __launch_bounds__(kSubWaves * 64) void subtree_kernel(const Pending* roots, int bins, int max_leaf, float traversal_cost, int* order,
                                                                  const BBox* boxes, const float4* centres, TNode* tree, int* node_counter,
                                                                  Pending* lists, int* tmp) {
    __shared__ int cnt[kSubWaves][3][kMaxBins];
    __shared__ int bx[kSubWaves][3][kMaxBins][6];
    __shared__ int s_count[2];
    const int wave = int(threadIdx.x >> 6);
    Pending* const lv = lists + size_t(blockIdx.x) * 2 * kChunk;  // level lists at lv and lv + kChunk
    if (threadIdx.x == 0) {
        lv[0] = roots[blockIdx.x];
        s_count[0] = 1;
        s_count[1] = 0;
    }
    __syncthreads();
    int cur = 0;
    while (true) {
        const int n = s_count[cur];
        if (n == 0) break;  // every thread reads the same count: the loop ends for the whole block at once
        for (int i = wave; i < n; i += kSubWaves)
            build_node(lv[cur * kChunk + i], bins, max_leaf, traversal_cost, order, boxes, centres, tree, node_counter, tmp, cnt[wave], bx[wave],
                       lv + (cur ^ 1) * kChunk, &s_count[cur ^ 1]);
        __syncthreads();
        if (threadIdx.x == 0) s_count[cur] = 0;
        cur ^= 1;
        __syncthreads();
    }
}

// Flattening on the device. The host builder numbers the flat nodes by a stack walk (reference
// bvh.cpp:214-255): popping node X gives X's internal children the next free slots, left then right, and
// pushes them in that order, so the right subtree is walked before the left. Let I(n) be the number of
// internal nodes in n's subtree (0 for a leaf), c(X) the walk's slot counter when X is popped and p(X) X's
// pop index; at the root c = 1, p = 0 and slot = 0. With k(P) the internal children of X's parent P and
// R = P's right child:
//   X = right child of P:  p = p(P) + 1,          c = c(P) + k(P),                      slot = c(P) + [P's left child internal]
//   X = left child of P:   p = p(P) + 1 + I(R),   c = c(P) + k(P) + max(I(R) - 1, 0),   slot = c(P)
// and X's internal children get slots c(X) (left) and c(X) + [left internal] (right). Every node sums
// these terms over its path to the root, so all nodes are placed at once with no ordered walk.
// Kernels, in order: parent links, leaf-start flags, their prefix sum (two passes), I(n), placement.

// Parent links; leaf-start flags [0, n] and the depth maximum start at zero. The grid covers
// max(m, n + 1) threads.
__global__ void flat_link_kernel(const TNode* tree, int m, int n, int* parent, int* flag, int* max_depth) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) flag[i] = 0;
    if (i >= m) return;
    if (i == 0) {
        parent[0] = -1;
        *max_depth = 0;
    }
    const TNode t = tree[i];
    if (t.left >= 0) {
        parent[t.left] = i;
        parent[t.right] = i;
    }
}

// Leaf ranges are non-empty and tile [0, n), and a node's range is the union of its subtree's leaves,
// so I(n) is the number of leaf starts in n's range, less one: a prefix sum over these flags.
__global__ void flat_mark_kernel(const TNode* tree, int m, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const TNode t = tree[i];
    if (t.left < 0) flag[t.first] = 1;
}

constexpr int kTile = 4 * kThreads;  // flags per block of the prefix sum

// Exclusive prefix of flag[0, len) within each tile of kTile flags; the tile's total goes to sums.
__global__ __launch_bounds__(kThreads) void tile_scan_kernel(const int* flag, int len, int* pre, int* sums) {
    __shared__ int red[kThreads / 64];
    const int base = blockIdx.x * kTile + threadIdx.x * 4;
    int v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < len ? flag[base + k] : 0;
        s += v[k];
    }
    int excl;
    const int total = block_scan(s, red, excl);
    for (int k = 0; k < 4; k++) {
        if (base + k < len) pre[base + k] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile totals in place, by one block.
__global__ __launch_bounds__(kThreads) void tile_sums_kernel(int* sums, int tiles) {
    __shared__ int red[kThreads / 64];
    int carry = 0;
    for (int a = 0; a < tiles; a += kThreads) {  // the same trip count for every thread
        const int i = a + int(threadIdx.x);
        const int v = i < tiles ? sums[i] : 0;
        int excl;
        const int total = block_scan(v, red, excl);
        if (i < tiles) sums[i] = carry + excl;
        carry += total;
    }
}

__global__ void flat_inner_kernel(const TNode* tree, int m, const int* pre, const int* sums, int* inner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const TNode t = tree[i];
    if (t.left < 0) {
        inner[i] = 0;
        return;
    }
    const int a = t.first, b = t.first + t.count;
    inner[i] = (pre[b] + sums[b / kTile]) - (pre[a] + sums[a / kTile]) - 1;
}

// One flat node per internal build node, at its slot; terms[p] holds the two SAH terms the walk adds when
// it pops the node (pop index p), so the host can sum them in the walk's order.
__global__ void flat_place_kernel(const TNode* tree, int m, int internal, const int* parent, const int* inner, float traversal_cost,
                                  float root_area, ph_bvh_node* nodes, float2* terms, int* max_depth) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const TNode t = tree[x];
    if (t.left < 0) return;
    int p = 0, c = 1, depth = 0, first_dc = 0, slot_add = 0, n = x;
    for (int pp = parent[n]; pp >= 0; pp = parent[n]) {
        const TNode pt = tree[pp];
        const int ir = inner[pt.right], k = (inner[pt.left] > 0) + (ir > 0);
        const bool right = n == pt.right;
        const int dc = right ? k : k + max(ir - 1, 0);
        p += right ? 1 : 1 + ir;
        c += dc;
        if (depth == 0) {
            first_dc = dc;
            slot_add = right ? (inner[pt.left] > 0) : 0;
        }
        depth++;
        n = pp;
    }
    const int slot = x == 0 ? 0 : c - first_dc + slot_add;
    if (slot >= internal || p >= internal) return;  // only a malformed tree; the host's count check reports it
    const TNode l = tree[t.left], r = tree[t.right];
    ph_bvh_node nd;
    for (int k = 0; k < 3; k++) {
        nd.lmin[k] = l.box.lo[k], nd.lmax[k] = l.box.hi[k];
        nd.rmin[k] = r.box.lo[k], nd.rmax[k] = r.box.hi[k];
    }
    const bool li = l.left >= 0, ri = r.left >= 0;
    nd.left = li ? c : l.first;
    nd.lcount = li ? 0 : l.count;
    nd.right = ri ? c + (li ? 1 : 0) : r.first;
    nd.rcount = ri ? 0 : r.count;
    nodes[slot] = nd;
    terms[p] = make_float2((li ? traversal_cost : float(l.count)) * surface(l.box) / root_area,
                           (ri ? traversal_cost : float(r.count)) * surface(r.box) / root_area);
    atomicMax(max_depth, depth);
}

// Device buffers that only grow: per-level arrays keep their allocation across levels (a smaller request
// reuses it), so the build does not free and reallocate, and hipFree's device synchronisation is avoided.
template <class T>
void grow(DevBuf<T>& b, size_t n) {
    if (!b.ptr || b.count < n) b.allocate(std::max<size_t>(n, 1));
}
template <class T>
void put(DevBuf<T>& b, const T* src, size_t n, hipStream_t st) {
    grow(b, n);
    if (n) PH_CHECK(hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, st));
}

// One stream per device for the build, created on first use and kept for the process.
hipStream_t build_stream() {
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    int dev = 0;
    PH_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (size_t(dev) >= streams.size()) streams.resize(size_t(dev) + 1, nullptr);
    if (!streams[size_t(dev)]) PH_CHECK(hipStreamCreateWithFlags(&streams[size_t(dev)], hipStreamNonBlocking));
    return streams[size_t(dev)];
}

}  // namespace

int64_t bvh_build_device(const float* positions, int num_vertices, const int* indices, int n, int bins, int max_leaf, float traversal_cost,
                         ph_bvh_node* nodes_out, int* order_out, ph_bvh_build_stats& stats) {
    const auto t0 = std::chrono::steady_clock::now();
    stats = ph_bvh_build_stats{};
    auto clear_side = [](ph_bvh_node& nd, int side) {
        float* lo = side == 0 ? nd.lmin : nd.rmin;
        float* hi = side == 0 ? nd.lmax : nd.rmax;
        for (int k = 0; k < 3; k++) lo[k] = kBoxInf, hi[k] = -kBoxInf;
        (side == 0 ? nd.left : nd.right) = -1;
        (side == 0 ? nd.lcount : nd.rcount) = 0;
    };
    // A root that is a leaf: its range on the left, an empty right side.
    auto write_leaf_root = [&](ph_bvh_node& nd, const TNode& root) {
        for (int k = 0; k < 3; k++) nd.lmin[k] = root.box.lo[k], nd.lmax[k] = root.box.hi[k];
        nd.left = root.first;
        nd.lcount = root.count;
        clear_side(nd, 1);
    };
    if (n == 0) {
        ph_bvh_node root{};
        clear_side(root, 0);
        clear_side(root, 1);
        nodes_out[0] = root;
        stats.nodes = 1;
        return 1;
    }
    hipStream_t st = build_stream();
    DevBuf<float> d_pos;
    DevBuf<int> d_idx, d_order, d_counter;
    DevBuf<BBox> d_boxes;
    DevBuf<float4> d_centres;
    DevBuf<TNode> d_tree;
    d_pos.upload(positions, size_t(num_vertices) * 3, st);
    d_idx.upload(indices, size_t(n) * 3, st);
    d_order.allocate(size_t(n));
    d_boxes.allocate(size_t(n));
    d_centres.allocate(size_t(n));
    d_tree.allocate(size_t(2) * size_t(n) + 1);
    d_counter.allocate(1);
    hipLaunchKernelGGL(prep_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_pos.ptr, d_idx.ptr, n, d_boxes.ptr, d_centres.ptr, d_order.ptr);
    PH_CHECK(hipGetLastError());

    // Large nodes, level by level. The per-node decisions are the host builder's code on the host.
    struct Big {
        int node, first, count, depth;
    };
    std::vector<Big> level{{0, 0, n, 0}};
    std::vector<Pending> small;
    int next_node = 1;
    if (n <= kSubtree) {
        small.push_back(Pending{0, 0, n, 0});
        level.clear();
    }
    DevBuf<Chunk> d_chunks;
    DevBuf<BBox> d_chunk_boxes;
    DevBuf<AxisParams> d_params;
    DevBuf<int> d_bins, d_counts, d_base, d_lpos, d_rpos;
    DevBuf<Split> d_splits;
    std::vector<std::pair<int, TNode>> big_records;  // (node index, record) of large nodes
    while (!level.empty()) {
        std::vector<Chunk> chunks;
        std::vector<int> chunk_first(level.size() + 1, 0);
        for (size_t i = 0; i < level.size(); i++) {
            chunk_first[i] = int(chunks.size());
            for (int a = level[i].first; a < level[i].first + level[i].count; a += kChunk)
                chunks.push_back(Chunk{int(i), a, std::min(a + kChunk, level[i].first + level[i].count)});
        }
        chunk_first[level.size()] = int(chunks.size());
        const int T = int(chunks.size());
        put(d_chunks, chunks.data(), chunks.size(), st);
        grow(d_chunk_boxes, size_t(2) * size_t(T));
        hipLaunchKernelGGL(chunk_boxes_kernel, dim3(T), dim3(kThreads), 0, st, d_chunks.ptr, d_order.ptr, d_boxes.ptr, d_centres.ptr, d_chunk_boxes.ptr);
        PH_CHECK(hipGetLastError());
        std::vector<BBox> cboxes(size_t(2) * size_t(T));
        PH_CHECK(hipMemcpyAsync(cboxes.data(), d_chunk_boxes.ptr, cboxes.size() * sizeof(BBox), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        std::vector<BBox> box(level.size()), cbox(level.size());
        std::vector<AxisParams> params(level.size());
        std::vector<int> splitting;
        for (size_t i = 0; i < level.size(); i++) {
            BBox bb = empty_box(), cb = empty_box();
            for (int c = chunk_first[i]; c < chunk_first[i + 1]; c++) {
                bb = fold(bb, cboxes[size_t(2 * c)]);
                cb = fold(cb, cboxes[size_t(2 * c + 1)]);
            }
            box[i] = bb;
            cbox[i] = cb;
            AxisParams& ap = params[i];
            ap.nb = bins;
            for (int k = 0; k < 3; k++) {
                ap.lo[k] = cb.lo[k];
                ap.valid[k] = (cb.hi[k] - cb.lo[k] < 1e-12f) ? 0 : 1;
                ap.scale[k] = ap.valid[k] ? float(bins) / (cb.hi[k] - cb.lo[k]) : 0.0f;
            }
        }
        // Large nodes are above max_leaf and depth 60 is far: all split unless no axis has extent.
        put(d_params, params.data(), params.size(), st);
        grow(d_bins, size_t(T) * 3 * kMaxBins * 7);
        hipLaunchKernelGGL(chunk_bins_kernel, dim3(T), dim3(kThreads), 0, st, d_chunks.ptr, d_params.ptr, d_order.ptr, d_boxes.ptr, d_centres.ptr,
                           d_bins.ptr);
        PH_CHECK(hipGetLastError());
        std::vector<int> hbins(size_t(T) * 3 * kMaxBins * 7);
        PH_CHECK(hipMemcpyAsync(hbins.data(), d_bins.ptr, hbins.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        auto value = [](int i) { return __builtin_bit_cast(float, i >= 0 ? i : i ^ 0x7fffffff); };
        std::vector<Split> splits(level.size());
        std::vector<Big> next;
        for (size_t i = 0; i < level.size(); i++) {
            const Big& nd = level[i];
            const int first = nd.first, count = nd.count;
            int bin_n[3][kMaxBins] = {};
            BBox bin_box[3][kMaxBins];
            for (int axis = 0; axis < 3; axis++)
                for (int bb = 0; bb < bins; bb++) bin_box[axis][bb] = empty_box();
            for (int c = chunk_first[i]; c < chunk_first[i + 1]; c++)
                for (int axis = 0; axis < 3; axis++)
                    for (int bb = 0; bb < bins; bb++) {
                        const int* r = &hbins[(size_t(c) * 3 * kMaxBins + size_t(axis) * kMaxBins + size_t(bb)) * 7];
                        bin_n[axis][bb] += r[0];
                        BBox v{{value(r[1]), value(r[2]), value(r[3])}, {value(r[4]), value(r[5]), value(r[6])}};
                        bin_box[axis][bb] = fold(bin_box[axis][bb], v);
                    }
            float best = kBoxInf;
            int best_axis = -1, best_bin = -1;
            const float leaf_cost = float(count);
            const float node_area = surface(box[i]);
            for (int axis = 0; axis < 3; axis++) {
                if (!params[i].valid[axis]) continue;
                float suffix_area[kMaxBins];
                int suffix_n[kMaxBins];
                BBox acc = empty_box();
                int nn = 0;
                for (int bb = bins - 1; bb > 0; bb--) {
                    acc = fold(acc, bin_box[axis][bb]);
                    nn += bin_n[axis][bb];
                    suffix_area[bb] = surface(acc);
                    suffix_n[bb] = nn;
                }
                acc = empty_box();
                nn = 0;
                for (int bb = 0; bb < bins - 1; bb++) {
                    acc = fold(acc, bin_box[axis][bb]);
                    nn += bin_n[axis][bb];
                    const int nr = suffix_n[bb + 1];
                    if (nn == 0 || nr == 0) continue;
                    const float cost = traversal_cost + (surface(acc) * float(nn) + suffix_area[bb + 1] * float(nr)) / node_area;
                    if (cost < best) {
                        best = cost;
                        best_axis = axis;
                        best_bin = bb;
                    }
                }
            }
            const bool forced = count > max_leaf;
            bool leaf = count <= 1 || nd.depth > 60;
            if (!leaf) {
                if (best_axis < 0)
                    leaf = !forced;
                else if (best >= leaf_cost && !forced)
                    leaf = true;
            }
            TNode rec{box[i], first, count, -1, -1};
            Split sp{-1, best_bin, bins, first + count / 2, 0.0f, 0.0f};  // axis -1: no partition
            if (!leaf && best_axis >= 0) {
                sp.axis = best_axis;
                int left_n = 0;
                for (int bb = 0; bb <= best_bin; bb++) left_n += bin_n[best_axis][bb];
                sp.mid = first + left_n;
                sp.lo = params[i].lo[best_axis];
                sp.scale = params[i].scale[best_axis];
                splitting.push_back(int(i));
            }
            splits[i] = sp;
            if (!leaf) {
                // The partition (below) keeps sp.mid unless one side is empty.
                const int m = (best_axis >= 0 && sp.mid != first && sp.mid != first + count) ? sp.mid : first + count / 2;
                rec.left = next_node;
                rec.right = next_node + 1;
                next_node += 2;
                for (int side = 0; side < 2; side++) {
                    const int cf = side == 0 ? first : m, cc = side == 0 ? m - first : first + count - m;
                    const int idx = side == 0 ? rec.left : rec.right;
                    if (cc > kSubtree)
                        next.push_back(Big{idx, cf, cc, nd.depth + 1});
                    else
                        small.push_back(Pending{idx, cf, cc, nd.depth + 1});
                }
            }
            big_records.push_back({nd.node, rec});
        }
        if (!splitting.empty()) {
            put(d_splits, splits.data(), splits.size(), st);
            grow(d_counts, size_t(2) * size_t(T));
            hipLaunchKernelGGL(chunk_misplaced_kernel, dim3(T), dim3(kThreads), 0, st, d_chunks.ptr, d_splits.ptr, d_order.ptr, d_centres.ptr, d_counts.ptr);
            PH_CHECK(hipGetLastError());
            std::vector<int> counts(size_t(2) * size_t(T));
            PH_CHECK(hipMemcpyAsync(counts.data(), d_counts.ptr, counts.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipStreamSynchronize(st));
            // Per chunk: rank of its first left element, rank (from the back) just past its last right
            // element, and the node's offset into the pair arrays.
            std::vector<int> base(size_t(3) * size_t(T), 0);
            int offset = 0;
            std::vector<int> chunk_sel;
            for (int i : splitting) {
                int l = 0, r_total = 0;
                for (int c = chunk_first[size_t(i)]; c < chunk_first[size_t(i) + 1]; c++) r_total += counts[size_t(2 * c + 1)];
                int r_after = r_total;
                for (int c = chunk_first[size_t(i)]; c < chunk_first[size_t(i) + 1]; c++) {
                    base[size_t(3 * c)] = l;
                    base[size_t(3 * c + 1)] = r_after;
                    base[size_t(3 * c + 2)] = offset;
                    l += counts[size_t(2 * c)];
                    r_after -= counts[size_t(2 * c + 1)];
                    chunk_sel.push_back(c);
                }
                if (l != r_total) throw std::runtime_error("device BVH build: unbalanced partition");
                offset += l;
            }
            if (offset > 0) {
                std::vector<Chunk> sel;
                std::vector<int> sel_base;
                for (int c : chunk_sel) {
                    sel.push_back(chunks[size_t(c)]);
                    sel_base.insert(sel_base.end(), {base[size_t(3 * c)], base[size_t(3 * c + 1)], base[size_t(3 * c + 2)]});
                }
                put(d_chunks, sel.data(), sel.size(), st);
                put(d_base, sel_base.data(), sel_base.size(), st);
                grow(d_lpos, size_t(offset));
                grow(d_rpos, size_t(offset));
                hipLaunchKernelGGL(chunk_ranks_kernel, dim3(unsigned(sel.size())), dim3(kThreads), 0, st, d_chunks.ptr, d_splits.ptr, d_order.ptr,
                                   d_centres.ptr, d_base.ptr, d_lpos.ptr, d_rpos.ptr);
                PH_CHECK(hipGetLastError());
                hipLaunchKernelGGL(swap_kernel, dim3((offset + 255) / 256), dim3(256), 0, st, d_order.ptr, d_lpos.ptr, d_rpos.ptr, offset);
                PH_CHECK(hipGetLastError());
            }
        }
        level.swap(next);
    }
    // Large-node records, then every small subtree.
    // Indices [0, next_node) hold the large nodes and the subtree roots; the roots' records are written
    // by the subtree kernel, so one copy of the whole prefix places the large-node records.
    std::vector<TNode> prefix(size_t(next_node), TNode{empty_box(), 0, 0, -1, -1});
    for (auto& r : big_records) prefix[size_t(r.first)] = r.second;
    PH_CHECK(hipMemcpyAsync(d_tree.ptr, prefix.data(), prefix.size() * sizeof(TNode), hipMemcpyHostToDevice, st));
    PH_CHECK(hipStreamSynchronize(st));
    if (!small.empty()) {
        PH_CHECK(hipMemcpyAsync(d_counter.ptr, &next_node, sizeof(int), hipMemcpyHostToDevice, st));
        DevBuf<Pending> d_small, d_lists;
        DevBuf<int> d_tmp;
        d_small.upload(small.data(), small.size(), st);
        d_lists.allocate(small.size() * 2 * size_t(kChunk));
        d_tmp.allocate(size_t(n));
        hipLaunchKernelGGL(subtree_kernel, dim3(unsigned(small.size())), dim3(kSubWaves * 64), 0, st, d_small.ptr, bins, max_leaf, traversal_cost,
                           d_order.ptr, d_boxes.ptr, d_centres.ptr, d_tree.ptr, d_counter.ptr, d_lists.ptr, d_tmp.ptr);
        PH_CHECK(hipGetLastError());
        PH_CHECK(hipMemcpyAsync(&next_node, d_counter.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
    }
    PH_CHECK(hipMemcpyAsync(order_out, d_order.ptr, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, st));
    TNode root;
    PH_CHECK(hipMemcpyAsync(&root, d_tree.ptr, sizeof(TNode), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    for (int k = 0; k < 3; k++) stats.bounds_min[k] = root.box.lo[k], stats.bounds_max[k] = root.box.hi[k];
    if (root.left < 0) {
        write_leaf_root(nodes_out[0], root);
        stats.nodes = 1;
        stats.leaves = 1;
        stats.sah_cost = float(root.count);
    } else {
        // Flattening on the device (kernels above); the host only sums the SAH terms in the walk's order.
        const float root_area = std::max(surface(root.box), 1e-20f);
        const int m = next_node, grid = (m + 255) / 256;
        DevBuf<int> d_parent, d_flag, d_pre, d_sums, d_inner, d_depth;
        DevBuf<ph_bvh_node> d_nodes;
        DevBuf<float2> d_terms;
        const int len = n + 1, tiles = (len + kTile - 1) / kTile;
        d_parent.allocate(size_t(m));
        d_flag.allocate(size_t(len));
        d_pre.allocate(size_t(len));
        d_sums.allocate(size_t(tiles));
        d_inner.allocate(size_t(m));
        d_depth.allocate(1);
        // A full binary tree of m nodes has (m - 1) / 2 internal ones.
        const int internal = (m - 1) / 2;
        d_nodes.allocate(size_t(internal));
        d_terms.allocate(size_t(internal));
        hipLaunchKernelGGL(flat_link_kernel, dim3((std::max(m, len) + 255) / 256), dim3(256), 0, st, d_tree.ptr, m, n, d_parent.ptr, d_flag.ptr,
                           d_depth.ptr);
        PH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(flat_mark_kernel, dim3(grid), dim3(256), 0, st, d_tree.ptr, m, d_flag.ptr);
        PH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(tile_scan_kernel, dim3(tiles), dim3(kThreads), 0, st, d_flag.ptr, len, d_pre.ptr, d_sums.ptr);
        PH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(tile_sums_kernel, dim3(1), dim3(kThreads), 0, st, d_sums.ptr, tiles);
        PH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(flat_inner_kernel, dim3(grid), dim3(256), 0, st, d_tree.ptr, m, d_pre.ptr, d_sums.ptr, d_inner.ptr);
        PH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(flat_place_kernel, dim3(grid), dim3(256), 0, st, d_tree.ptr, m, internal, d_parent.ptr, d_inner.ptr, traversal_cost,
                           root_area, d_nodes.ptr, d_terms.ptr, d_depth.ptr);
        PH_CHECK(hipGetLastError());
        int root_inner = 0;
        std::vector<float2> terms(static_cast<size_t>(internal));
        PH_CHECK(hipMemcpyAsync(&root_inner, d_inner.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipMemcpyAsync(&stats.max_depth, d_depth.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipMemcpyAsync(nodes_out, d_nodes.ptr, size_t(internal) * sizeof(ph_bvh_node), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipMemcpyAsync(terms.data(), d_terms.ptr, terms.size() * sizeof(float2), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        if (root_inner != internal) throw std::runtime_error("device BVH build: the tree is not a full binary tree");
        stats.nodes = internal;
        stats.leaves = internal + 1;
        float sah = traversal_cost;
        for (const float2& t : terms) {
            sah += t.x;
            sah += t.y;
        }
        stats.sah_cost = sah;
    }
    stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return int64_t(stats.nodes);
}

}  // namespace phk
