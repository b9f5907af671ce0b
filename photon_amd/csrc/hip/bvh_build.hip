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
//     host's own code, run on the host for large nodes and by one thread for small ones.
//   * the partition is std::partition's: the k-th element on the wrong side counted from the front
//     swaps with the k-th counted from the back; ranks come from scans, so it runs in parallel.
// Large nodes (more than kChunk triangles) are split level by level by multi-block passes with the
// per-node decisions on the host; every subtree of at most kChunk triangles is then built depth first
// by one block. The node indices of the build tree differ from the host's, but the flattening numbers
// nodes by tree structure only.
#include <algorithm>
#include <chrono>
#include <vector>

#include "internal.h"

// Host and device: the large-node decisions run on the host with the same functions.
#define PH_HD __host__ __device__ inline

namespace phk {

namespace {

constexpr float kBoxInf = 1e30f;  // the host builder's empty box bound
constexpr int kMaxBins = 64;
constexpr int kChunk = 2048;  // triangles per block: large-node passes, and the small-subtree bound
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

// ---- small subtrees: one block each, depth first --------------------------------------------------------

struct Pending {
    int node, first, count, depth;
};

__global__ __launch_bounds__(kThreads) void subtree_kernel(const Pending* roots, int bins, int max_leaf, float traversal_cost, int* order,
                                                            const BBox* boxes, const float4* centres, TNode* tree, int* node_counter) {
    __shared__ BBox scratch[kThreads / 64];
    __shared__ int red[kThreads / 64];
    __shared__ Pending stack[128];
    __shared__ int cnt[3][kMaxBins];
    __shared__ int bx[3][kMaxBins][6];
    __shared__ int lpos[kChunk / 2], rpos[kChunk / 2];
    __shared__ int s_sp, s_split_axis, s_split_bin, s_mid;
    __shared__ float s_lo, s_scale;
    __shared__ BBox s_cbox;
    if (threadIdx.x == 0) {
        stack[0] = roots[blockIdx.x];
        s_sp = 1;
    }
    __syncthreads();
    const int nb = bins;
    while (true) {
        __syncthreads();
        if (s_sp == 0) break;
        const Pending nd = stack[s_sp - 1];
        __syncthreads();
        if (threadIdx.x == 0) s_sp--;
        const int a = nd.first, b = nd.first + nd.count;
        BBox box, cbox;
        range_boxes(order, boxes, centres, a, b, scratch, box, cbox);
        if (threadIdx.x == 0) {
            TNode t;
            t.box = box;
            t.first = nd.first;
            t.count = nd.count;
            t.left = t.right = -1;
            tree[nd.node] = t;
            s_cbox = cbox;
        }
        if (nd.count <= 1 || nd.depth > 60) continue;
        // Bins of the three axes.
        for (int i = threadIdx.x; i < 3 * kMaxBins; i += kThreads) {
            cnt[i / kMaxBins][i % kMaxBins] = 0;
            for (int k = 0; k < 3; k++) {
                bx[i / kMaxBins][i % kMaxBins][k] = ordered(kBoxInf);
                bx[i / kMaxBins][i % kMaxBins][3 + k] = ordered(-kBoxInf);
            }
        }
        __syncthreads();
        for (int p = a + threadIdx.x; p < b; p += kThreads) {
            const int t = order[p];
            const float4 ce = centres[t];
            const float cc[3] = {ce.x, ce.y, ce.z};
            const BBox tb = boxes[t];
            for (int axis = 0; axis < 3; axis++) {
                const float lo = cbox.lo[axis], hi = cbox.hi[axis];
                if (hi - lo < 1e-12f) continue;
                const int bin = bin_of(cc[axis], lo, float(nb) / (hi - lo), nb);
                atomicAdd(&cnt[axis][bin], 1);
                for (int k = 0; k < 3; k++) {
                    atomicMin(&bx[axis][bin][k], ordered(tb.lo[k]));
                    atomicMax(&bx[axis][bin][3 + k], ordered(tb.hi[k]));
                }
            }
        }
        __syncthreads();
        // The split search and leaf test, as the host builder runs them.
        if (threadIdx.x == 0) {
            float best = kBoxInf;
            int best_axis = -1, best_bin = -1;
            const float leaf_cost = float(nd.count);
            const float node_area = surface(box);
            for (int axis = 0; axis < 3; axis++) {
                const float lo = s_cbox.lo[axis], hi = s_cbox.hi[axis];
                if (hi - lo < 1e-12f) continue;
                float suffix_area[kMaxBins];
                int suffix_n[kMaxBins];
                BBox acc = empty_box();
                int n = 0;
                for (int bb = nb - 1; bb > 0; bb--) {
                    BBox bin_box;
                    for (int k = 0; k < 3; k++) {
                        bin_box.lo[k] = from_ordered(bx[axis][bb][k]);
                        bin_box.hi[k] = from_ordered(bx[axis][bb][3 + k]);
                    }
                    acc = fold(acc, bin_box);
                    n += cnt[axis][bb];
                    suffix_area[bb] = surface(acc);
                    suffix_n[bb] = n;
                }
                acc = empty_box();
                n = 0;
                for (int bb = 0; bb < nb - 1; bb++) {
                    BBox bin_box;
                    for (int k = 0; k < 3; k++) {
                        bin_box.lo[k] = from_ordered(bx[axis][bb][k]);
                        bin_box.hi[k] = from_ordered(bx[axis][bb][3 + k]);
                    }
                    acc = fold(acc, bin_box);
                    n += cnt[axis][bb];
                    const int nr = suffix_n[bb + 1];
                    if (n == 0 || nr == 0) continue;
                    const float cost = traversal_cost + (surface(acc) * float(n) + suffix_area[bb + 1] * float(nr)) / node_area;
                    if (cost < best) {
                        best = cost;
                        best_axis = axis;
                        best_bin = bb;
                    }
                }
            }
            const bool forced = nd.count > max_leaf;
            bool leaf = false;
            if (best_axis < 0) {
                leaf = !forced;
            } else if (best >= leaf_cost && !forced) {
                leaf = true;
            }
            s_split_axis = leaf ? -2 : best_axis;
            s_split_bin = best_bin;
            int left_n = nd.count / 2;
            if (!leaf && best_axis >= 0) {
                left_n = 0;
                for (int bb = 0; bb <= best_bin; bb++) left_n += cnt[best_axis][bb];
                s_lo = s_cbox.lo[best_axis];
                s_scale = float(nb) / (s_cbox.hi[best_axis] - s_cbox.lo[best_axis]);
            }
            s_mid = a + left_n;
        }
        __syncthreads();
        const int axis = s_split_axis;
        if (axis == -2) continue;
        int mid = s_mid;
        if (axis >= 0) {
            // std::partition: the k-th wrong-side element from the front swaps with the k-th from the back.
            int lbase = 0, rbase = 0;
            for (int p0 = a; p0 < b; p0 += kThreads) {
                const int p = p0 + int(threadIdx.x);
                bool lm = false, rm = false;
                if (p < b) {
                    const float4 ce = centres[order[p]];
                    const float cc[3] = {ce.x, ce.y, ce.z};
                    const bool pred = bin_of(cc[axis], s_lo, s_scale, nb) <= s_split_bin;
                    lm = p < mid && !pred;
                    rm = p >= mid && pred;
                }
                int le, re;
                const int lt = block_scan(lm ? 1 : 0, red, le);
                const int rt = block_scan(rm ? 1 : 0, red, re);
                if (lm) lpos[lbase + le] = p;
                if (rm) rpos[rbase + re] = p;
                lbase += lt;
                rbase += rt;
            }
            __syncthreads();
            // rpos holds the right ones front to back: its k-th from the back is rpos[rbase - 1 - k].
            for (int k = threadIdx.x; k < lbase; k += kThreads) {
                const int pa = lpos[k], pb = rpos[rbase - 1 - k];
                const int t = order[pa];
                order[pa] = order[pb];
                order[pb] = t;
            }
            __syncthreads();
            if (mid == a || mid == b) mid = a + nd.count / 2;
        }
        if (threadIdx.x == 0) {
            const int kids = atomicAdd(node_counter, 2);
            tree[nd.node].left = kids;
            tree[nd.node].right = kids + 1;
            stack[s_sp] = Pending{kids + 1, mid, b - mid, nd.depth + 1};
            stack[s_sp + 1] = Pending{kids, a, mid - a, nd.depth + 1};
            s_sp += 2;
        }
    }
}

}  // namespace

int64_t bvh_build_device(const float* positions, int num_vertices, const int* indices, int n, int bins, int max_leaf, float traversal_cost,
                         std::vector<ph_bvh_node>& nodes_out, std::vector<int>& order_out, ph_bvh_build_stats& stats) {
    const auto t0 = std::chrono::steady_clock::now();
    stats = ph_bvh_build_stats{};
    order_out.assign(size_t(n), 0);
    nodes_out.clear();
    auto clear_side = [](ph_bvh_node& nd, int side) {
        float* lo = side == 0 ? nd.lmin : nd.rmin;
        float* hi = side == 0 ? nd.lmax : nd.rmax;
        for (int k = 0; k < 3; k++) lo[k] = kBoxInf, hi[k] = -kBoxInf;
        (side == 0 ? nd.left : nd.right) = -1;
        (side == 0 ? nd.lcount : nd.rcount) = 0;
    };
    if (n == 0) {
        ph_bvh_node root{};
        clear_side(root, 0);
        clear_side(root, 1);
        nodes_out.push_back(root);
        stats.nodes = 1;
        return 1;
    }
    hipStream_t st;
    PH_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
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
    if (n <= kChunk) {
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
        d_chunks.upload(chunks.data(), chunks.size(), st);
        d_chunk_boxes.allocate(size_t(2) * size_t(T));
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
        d_params.upload(params.data(), params.size(), st);
        d_bins.allocate(size_t(T) * 3 * kMaxBins * 7);
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
                    if (cc > kChunk)
                        next.push_back(Big{idx, cf, cc, nd.depth + 1});
                    else
                        small.push_back(Pending{idx, cf, cc, nd.depth + 1});
                }
            }
            big_records.push_back({nd.node, rec});
        }
        if (!splitting.empty()) {
            d_splits.upload(splits.data(), splits.size(), st);
            d_counts.allocate(size_t(2) * size_t(T));
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
                d_chunks.upload(sel.data(), sel.size(), st);
                d_base.upload(sel_base.data(), sel_base.size(), st);
                d_lpos.allocate(size_t(offset));
                d_rpos.allocate(size_t(offset));
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
    for (auto& r : big_records) PH_CHECK(hipMemcpyAsync(d_tree.ptr + r.first, &r.second, sizeof(TNode), hipMemcpyHostToDevice, st));
    PH_CHECK(hipStreamSynchronize(st));
    if (!small.empty()) {
        PH_CHECK(hipMemcpyAsync(d_counter.ptr, &next_node, sizeof(int), hipMemcpyHostToDevice, st));
        DevBuf<Pending> d_small;
        d_small.upload(small.data(), small.size(), st);
        hipLaunchKernelGGL(subtree_kernel, dim3(unsigned(small.size())), dim3(kThreads), 0, st, d_small.ptr, bins, max_leaf, traversal_cost, d_order.ptr,
                           d_boxes.ptr, d_centres.ptr, d_tree.ptr, d_counter.ptr);
        PH_CHECK(hipGetLastError());
        PH_CHECK(hipMemcpyAsync(&next_node, d_counter.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
    }
    std::vector<TNode> tree(static_cast<size_t>(next_node));
    PH_CHECK(hipMemcpyAsync(tree.data(), d_tree.ptr, tree.size() * sizeof(TNode), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(order_out.data(), d_order.ptr, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));

    // Flattening, as the host builder does it: a node's two children numbered consecutively, the right
    // child's subtree emitted before the left child's.
    auto write_side = [&](ph_bvh_node& nd, int side, const TNode& c, int child) {
        float* lo = side == 0 ? nd.lmin : nd.rmin;
        float* hi = side == 0 ? nd.lmax : nd.rmax;
        for (int k = 0; k < 3; k++) lo[k] = c.box.lo[k], hi[k] = c.box.hi[k];
        const bool leaf = c.left < 0;
        (side == 0 ? nd.left : nd.right) = leaf ? c.first : child;
        (side == 0 ? nd.lcount : nd.rcount) = leaf ? c.count : 0;
    };
    const TNode& root = tree[0];
    for (int k = 0; k < 3; k++) stats.bounds_min[k] = root.box.lo[k], stats.bounds_max[k] = root.box.hi[k];
    const float root_area = std::max(surface(root.box), 1e-20f);
    nodes_out.push_back(ph_bvh_node{});
    if (root.left < 0) {
        write_side(nodes_out[0], 0, root, 0);
        clear_side(nodes_out[0], 1);
        stats.nodes = 1;
        stats.leaves = 1;
        stats.sah_cost = float(root.count);
    } else {
        struct Item {
            int tree, slot, depth;
        };
        std::vector<Item> stack{{0, 0, 0}};
        stats.sah_cost = traversal_cost;
        while (!stack.empty()) {
            const Item it = stack.back();
            stack.pop_back();
            stats.nodes++;
            stats.max_depth = std::max(stats.max_depth, it.depth);
            const TNode& tn = tree[size_t(it.tree)];
            for (int side = 0; side < 2; side++) {
                const int ci = side == 0 ? tn.left : tn.right;
                const TNode& c = tree[size_t(ci)];
                int slot = -1;
                if (c.left >= 0) {
                    slot = int(nodes_out.size());
                    nodes_out.push_back(ph_bvh_node{});
                    stack.push_back({ci, slot, it.depth + 1});
                    stats.sah_cost += traversal_cost * surface(c.box) / root_area;
                } else {
                    stats.leaves++;
                    stats.sah_cost += float(c.count) * surface(c.box) / root_area;
                }
                write_side(nodes_out[size_t(it.slot)], side, c, slot);
            }
        }
    }
    stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return int64_t(nodes_out.size());
}

}  // namespace phk
