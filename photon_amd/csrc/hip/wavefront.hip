// Wavefront path tracing engine.
//
// Paths in flight live in structure-of-arrays state in HBM, one "slot" per path. A slot owns one pixel
// of one job at a time and runs that pixel's samples one after another, so each pixel's accumulation
// is the same sequence of float additions as in the one-thread-per-pixel kernel (bit-identical images,
// whatever else shares the wavefront); when a pixel's samples are done the slot takes the next work
// item (regeneration). A work item is one pixel of a pass: a range of sample indices of one job. A
// job's next pass is queued only once its previous pass has finished, so the samples of a pixel still
// run in order. Between runs the host lays out the passes' unstarted items in priority order (chunks of
// consecutive item indices after those already handed out), so large jobs yield to new ones.
//
// Every slot carries a one-byte stage (waiting to be traced, missed, hit a material class, finished,
// idle) plus a pending-shadow-ray bit. Queues are not appended to: after each stage an ordered
// compaction (a count pass and a scatter pass over the stage bytes) lists the slots of each next stage
// in slot order. Slots are handed their pixels in 8x8 tiles, so the 64 lanes of a wave work on one
// tile: camera rays stay coherent, and the per-slot state is read and written in long runs.
//
// One bounce of every path in flight is one iteration of these stages:
//   bookkeep  fold ray counts (one wave)
//   extend    closest-hit traversal of the extension queue; the stage records miss or material class
//   compact   -> miss queue and one queue per material class
//   miss      environment contribution of escaped paths
//   shade x4  surface interaction, NEE light sample, BSDF sample, Russian roulette; one launch per
//             material class (metal / opaque / glass / mixed) so a wave runs one lobe structure
//   compact   -> shadow queue and finished-path queue (serving: also idle slots while work waits)
//   shadow    any-hit traversal of the NEE queue; unoccluded contributions are added to L
//   regen     finished paths: add L to the pixel, start the next sample or the next work item
//   compact   -> extension queue of the next iteration
//   finish    (a single render, once every item is handed out and few paths are left) each remaining
//             path and the rest of its work item in one thread, as the per-pixel kernel runs them
// Traversal, shading and regeneration kernels are persistent: a fixed grid whose waves take the
// queue's 64-entry batches in turn (no atomics: atomics on one address serialize). Traversal lanes are
// refilled one ray at a time from the wave's batches, so a wave is not held by its longest ray.
// The bounce loop can be captured once into a HIP graph of G iterations and replayed; kernels read
// their sizes from device memory, so the graph stays valid while jobs come and go.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>

#include "device_scene.h"
#include "wavefront.h"

using namespace phd;

namespace phk {

namespace {

// Queue lengths (written by the compactions; each set is contiguous) and work counters.
enum : int {
    C_EXT = 0,     // extension queue
    C_SHADOW = 1,  // shadow queue
    C_TERM = 2,    // finished paths (serving: and idle slots while work waits)
    C_MISS = 3,
    C_CLASS = 4,    // 4 queues: metal, opaque, glass, mixed
    C_CHUNKS = 10,  // entries of the chunk table
    C_PARTS = 11,   // 1, or 8 for per-XCD parts of the queues (wave_share)
    C_NEED = 12,    // the last C_NEED entries of the finished-path queue take a new work item
    C_FINISH = 13,  // finish_kernel: extension-queue entries handed out so far
    C_COUNT = 16
};

// Slot stages: the stage code (low 3 bits), the extension ray's bin (bits 3-6: 0 camera rays, 1-8
// continuing rays by direction octant) or, for a finished path, whether its work item is finished too
// (S_NEED: the slot takes a new item), and the pending-shadow-ray bit.
enum : uint8_t { S_IDLE = 0, S_EXT = 1, S_MISS = 2, S_CLASS = 3, S_TERM = 7, S_CODE = 7, S_NEED = 0x08, S_SHADOW = 0x80 };
constexpr int kBins = 9;

PH_DEV uint8_t ext_stage(int bin) { return uint8_t(S_EXT | (bin << 3)); }

// Bin of a continuing ray: its direction octant, or one shared bin when binning is off.
PH_DEV int ray_bin(f3 d, int bins) { return bins ? 1 + (d.x < 0 ? 1 : 0) + (d.y < 0 ? 2 : 0) + (d.z < 0 ? 4 : 0) : 1; }

constexpr int kClasses = 4;

// Path state is read and written once per stage: streaming (non-temporal) accesses, so that it does not
// push the BVH and triangle lines out of the caches.
using v4f = float __attribute__((ext_vector_type(4)));
PH_DEV float4 ld_stream(const float4* p) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
PH_DEV void st_stream(float4* p, float4 a) { __builtin_nontemporal_store(v4f{a.x, a.y, a.z, a.w}, reinterpret_cast<v4f*>(p)); }
constexpr int kMaxChunks = 1 << 16;
// Ordered compaction: a block of 256 threads covers 4096 slots, 16 consecutive slots per thread.
constexpr int kCompactThreads = 256, kCompactPerThread = 16, kCompactChunk = kCompactThreads * kCompactPerThread;
constexpr int kMaxCategories = 9;

struct JobParams {
    DScene scene;
    ph_camera_view camera;
    int width, height, spp, pixels;
    int split, span;  // sub-streams of samples per pixel (serving), samples per sub-stream
    uint32_t frame0, seed;
    PathOpts opt;
    float4* accum;
    float4* albedo;
    float4* nd;
    const int* mat_class;
};

struct State {
    int* job;
    int* pixel;
    int* sample;
    int* sample_end;  // end of the slot's pass (serving)
    int* part;        // the slot's sub-stream (serving: accumulates into partial image `part`)
    uint8_t* stage;
    uint64_t* rng;
    float4* ro;  // origin, prev_pdf
    float4* rd;  // direction, flags (depth | prev_delta << 16)
    float4* beta;
    float4* L;
    float4* hit;
    float4* so;  // shadow origin, tmax
    float4* sd;  // shadow direction
    float4* sc;  // shadow contribution
    int* q_ext;
    int* q_shadow;
    int* q_term;
    int* q_miss;
    int* q_class[kClasses];
};

// Work items in queue order: chunk c holds items [chunk_first[c], ...), pixels of one pass, described
// by chunk_pass[c] = (table entry, first sample, end sample, local index of the chunk's first item).
// Item indices are 64-bit: a server hands out items for as long as it runs. items[0] is the next item
// to hand out, items[1] the end of the items laid out so far.
struct Work {
    const JobParams* jobs;
    unsigned long long* items;
    const long long* chunk_first;
    const int4* chunk_pass;
    int* job_done;  // pixel passes finished, per table entry
};

struct TraceTotals {
    unsigned long long nodes, tris, alpha;
};

PH_DEV int lane_id() { return __lane_id(); }

PH_DEV unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Adds this wave's traversal counters to the global totals (one atomic per counter per wave).
PH_DEV void flush_counts(const TraceCount& c, TraceTotals* totals) {
    const unsigned long long n = wave_sum(c.nodes), t = wave_sum(c.tris), a = wave_sum(c.alpha);
    if (lane_id() == 0) {
        atomicAdd(&totals->nodes, n);
        atomicAdd(&totals->tris, t);
        atomicAdd(&totals->alpha, a);
    }
}

// Adds to counters[key] the number of lanes with `pred` and that key: one atomic per distinct key in
// the wave instead of one per lane (lanes finishing work for one job would otherwise serialize on its
// counter). All lanes must call.
PH_DEV void wave_count(int* counters, int key, bool pred) {
    uint64_t mask = __ballot(pred);
    while (mask) {
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        const int k = __shfl(key, leader);
        const uint64_t same = __ballot(pred && key == k) & mask;
        if (lane_id() == leader) atomicAdd(counters + k, __popcll(same));
        mask &= ~same;
    }
}

// Work item -> pixel of a w x h image: 8-row bands, 8-column tiles inside a band (the last tile and band
// may be narrower), row-major inside a tile, so a wave's 64 lanes start on one 8x8 pixel block.
PH_DEV int work_pixel(int w, int width, int height) {
    const int band = w / (8 * width);
    const int bh = min(8, height - band * 8);
    const int in_band = w - band * 8 * width;
    const int tc = in_band / (8 * bh);
    const int tw = min(8, width - tc * 8);
    const int off = in_band - tc * 8 * bh;
    const int x = tc * 8 + off % tw, y = band * 8 + off / tw;
    return y * width + x;
}

// Table entry owning work item `w` (the last chunk that starts at or before it).
template <bool Multi>
PH_DEV int chunk_of(const Work& work, int chunks, long long w) {
    if (!Multi) return 0;
    int lo = 0, hi = chunks;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) / 2;
        if (work.chunk_first[mid] <= w)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// A kernel's copy of the single job (all lanes share it; unused when serving). The job table is written
// only by the host between launches, so it is read through the constant address space: scalar loads,
// once per kernel.
template <bool Multi>
PH_DEV JobParams single_job(const Work& work) {
    JobParams j{};
    if (!Multi) {
        const uint32_t __attribute__((address_space(4)))* src = (const uint32_t __attribute__((address_space(4)))*)work.jobs;
        uint32_t* dst = reinterpret_cast<uint32_t*>(&j);
#pragma unroll
        for (int i = 0; i < int(sizeof(JobParams) / 4); i++) dst[i] = src[i];
    }
    return j;
}

template <bool Multi>
PH_DEV const JobParams& job_of(const Work& work, const State& st, int slot, const JobParams& single) {
    return Multi ? work.jobs[st.job[slot]] : single;
}

// Stage of a path that ends with its current sample: S_TERM, plus S_NEED when that sample is the last of
// the slot's work item (the pixel's samples, or the pass or sub-stream the slot runs).
template <bool Multi>
PH_DEV uint8_t term_stage(const JobParams& jp, const State& st, int slot) {
    const int end = (Multi || jp.split > 1) ? st.sample_end[slot] : jp.spp;
    return uint8_t(S_TERM | (st.sample[slot] + 1 >= end ? S_NEED : 0));
}

PH_DEV void start_sample(const JobParams& jp, const State& st, int slot, int pixel, int k) {
    const int x = pixel % jp.width, y = pixel / jp.width;
    Pcg rng = pixel_rng(uint32_t(pixel), jp.frame0 + uint32_t(k), jp.seed);
    f2 jitter = pcg_uniform2(rng);
    f2 lens = pcg_uniform2(rng);
    f3 o, d;
    eye_ray(jp.camera, float(x) + jitter.x, float(y) + jitter.y, jp.width, jp.height, lens, o, d);
    st.pixel[slot] = pixel;
    st.sample[slot] = k;
    st.rng[slot] = rng.state;
    st_stream(st.ro + slot, make_float4(o.x, o.y, o.z, 0.0f));
    st_stream(st.rd + slot, make_float4(d.x, d.y, d.z, __int_as_float(1 << 16)));
    st_stream(st.beta + slot, make_float4(1.0f, 1.0f, 1.0f, 0.0f));
    st_stream(st.L + slot, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
}

// Starts work item `w` on `slot`: finds its job and pixel, records them, seeds the first sample.
template <bool Multi>
PH_DEV void start_work(const Work& work, const State& st, int chunks, int slot, long long w, const JobParams& single) {
    const int c = chunk_of<Multi>(work, chunks, w);
    const int4 pass = Multi ? work.chunk_pass[c] : make_int4(0, 0, 0, 0);
    const JobParams& jp = Multi ? work.jobs[pass.x] : single;
    int local = int(w - (Multi ? work.chunk_first[c] : 0)) + pass.w;
    int k0 = pass.y;
    if (Multi) {
        // Items of a pass: sub-stream `part` of every pixel, for parts 0 .. split-1 in turn; the pass's
        // sample range is relative to the sub-stream's first sample.
        const int part = local / jp.pixels;
        local -= part * jp.pixels;
        k0 += part * jp.span;
        st.job[slot] = pass.x;
        st.part[slot] = part;
        st.sample_end[slot] = part * jp.span + pass.z;
    } else if (jp.split > 1) {
        // A single render in sub-streams: item = part * pixels + pixel item; the part's samples in order.
        const int part = local / jp.pixels;
        local -= part * jp.pixels;
        k0 = part * jp.span;
        st.part[slot] = part;
        st.sample_end[slot] = k0 + jp.span;
    }
    start_sample(jp, st, slot, work_pixel(local, jp.width, jp.height), k0);
}

PH_DEV Pcg slot_rng(const State& st, int slot, int pixel) { return Pcg{st.rng[slot], (uint64_t(uint32_t(pixel)) << 1u) | 1u}; }

// Single job: slot s starts work item s.
__global__ void init_kernel(Work work, State st, int slots) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= slots) return;
    const bool has = slot < (long long)work.items[1];
    if (has) start_work<false>(work, st, 1, slot, slot, work.jobs[0]);
    st.stage[slot] = has ? ext_stage(0) : S_IDLE;
}

// Serving: every slot starts idle.
__global__ void init_idle_kernel(State st, int slots) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot < slots) st.stage[slot] = S_IDLE;
}

// Starts an iteration: folds the ray counts (this iteration's extension queue, the previous iteration's
// shadow queue) into the 64-bit totals.
__global__ void bookkeep_kernel(const int* counters, unsigned long long* rays) {
    if (threadIdx.x == 0) {
        rays[0] += (unsigned long long)counters[C_EXT];
        rays[1] += (unsigned long long)counters[C_SHADOW];
    }
}

// ---- ordered compaction -----------------------------------------------------------------------------
enum { K_EXT = 0, K_HIT = 1, K_DONE = 2 };

template <int Kind>
struct Compact {
    static constexpr int cats = Kind == K_EXT ? kBins : Kind == K_HIT ? 5 : 3;
    // Output list of a category; the categories of one list follow each other in category order. The
    // extension queue is one list of bins; the hit queues are one list each; the finished-path queue
    // holds the paths that continue their work item, then those that take a new one (and idle slots).
    static PH_DEV constexpr int list(int k) { return Kind == K_EXT ? 0 : Kind == K_HIT ? k : (k == 0 ? 0 : 1); }
    static constexpr int lists = Kind == K_EXT ? 1 : Kind == K_HIT ? 5 : 2;
    static PH_DEV constexpr int counter(int l) { return Kind == K_EXT ? C_EXT : Kind == K_HIT ? C_MISS + l : (l == 0 ? C_SHADOW : C_TERM); }
    // Category bits of a stage byte: extension bins; miss + 4 classes; shadow, finished path continuing
    // its item, finished path taking a new item (or idle slot while work waits).
    static PH_DEV unsigned bits(uint8_t s, bool idle_takes_work) {
        const int code = s & S_CODE;
        if (Kind == K_EXT) return code == S_EXT ? 1u << ((s >> 3) & 15) : 0u;
        if (Kind == K_HIT) return (code >= S_MISS && code < S_CLASS + kClasses) ? 1u << (code - S_MISS) : 0u;
        const unsigned term = code == S_TERM ? ((s & S_NEED) ? 4u : 2u) : (idle_takes_work && code == S_IDLE) ? 4u : 0u;
        return ((s & S_SHADOW) ? 1u : 0u) | term;
    }
};

struct QueueOut {
    int* q[kMaxCategories];
};

// The 16 stage bytes of one thread (zero past the end: no category, since idle slots only count when
// the caller asks for them and the padding is never idle).
PH_DEV void load_stages(const uint8_t* stage, int first, int n, uint8_t* b, bool& valid_all) {
    valid_all = first + kCompactPerThread <= n;
    if (valid_all) {
        const uint4 v = *reinterpret_cast<const uint4*>(stage + first);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        for (int j = 0; j < 16; j++) b[j] = uint8_t(w[j >> 2] >> (8 * (j & 3)));
    } else {
        for (int j = 0; j < 16; j++) b[j] = first + j < n ? stage[first + j] : uint8_t(0);
    }
}

template <int Kind, bool Multi>
PH_DEV bool idle_takes_work(const unsigned long long* items) {
    return Multi && Kind == K_DONE && items[0] < items[1];
}

// Per-chunk counts of every category: partial[k * nchunks + chunk].
template <int Kind, bool Multi>
__global__ __launch_bounds__(kCompactThreads) void compact_count_kernel(const uint8_t* stage, int n, const unsigned long long* items, int* partial,
                                                                       int nchunks) {
    constexpr int K = Compact<Kind>::cats;
    constexpr int W = kCompactThreads / 64;
    __shared__ int red[K][W];
    const bool idle = idle_takes_work<Kind, Multi>(items);
    const int first = blockIdx.x * kCompactChunk + threadIdx.x * kCompactPerThread;
    uint8_t b[16];
    bool full;
    load_stages(stage, first, n, b, full);
    int cnt[K] = {};
    for (int j = 0; j < 16; j++) {
        const unsigned m = (first + j < n) ? Compact<Kind>::bits(b[j], idle) : 0u;
        for (int k = 0; k < K; k++) cnt[k] += (m >> k) & 1u;
    }
    const int wave = threadIdx.x >> 6;
    for (int k = 0; k < K; k++) {
        int v = cnt[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id() == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (int(threadIdx.x) < K) {
        int t = 0;
        for (int w = 0; w < W; w++) t += red[threadIdx.x][w];
        partial[threadIdx.x * nchunks + blockIdx.x] = t;
    }
}

// Each chunk's first output position in each category: the totals of the earlier categories of the same
// list, plus the exclusive prefix of the chunk's count within its category. One block per category
// (each also sums the earlier categories of its list); the positions go to `first_pos`, laid out like
// `partial`. The block of a list's last category writes the list's length, and for the finished paths
// the block of category 2 writes how many take new work.
constexpr int kScanThreads = 1024;

template <int Kind>
__global__ __launch_bounds__(kScanThreads) void compact_scan_kernel(const int* partial, int* first_pos, int nchunks, int* counters) {
    constexpr int K = Compact<Kind>::cats;
    constexpr int W = kScanThreads / 64;
    __shared__ int wsum[W];
    const int k = blockIdx.x;
    const int l = Compact<Kind>::list(k);
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    // Sum over the block (all threads get it).
    auto block_total = [&](int v) {
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) wsum[wave] = v;
        __syncthreads();
        int t = 0;
        for (int w = 0; w < W; w++) t += wsum[w];
        __syncthreads();
        return t;
    };
    int earlier = 0;
    for (int k2 = 0; k2 < k; k2++)
        if (Compact<Kind>::list(k2) == l)
            for (int i = threadIdx.x; i < nchunks; i += kScanThreads) earlier += partial[k2 * nchunks + i];
    const int base = block_total(earlier);
    // Exclusive scan of this category's counts: consecutive runs per thread, then over the block.
    const int per = (nchunks + kScanThreads - 1) / kScanThreads;
    const int lo = min(nchunks, int(threadIdx.x) * per), hi = min(nchunks, lo + per);
    const int* p = partial + k * nchunks;
    int s = 0;
    for (int i = lo; i < hi; i++) s += p[i];
    int v = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < W; w++) {
        const int t = wsum[w];
        if (w < wave) before += t;
        total += t;
    }
    int run = base + before + v - s;
    for (int i = lo; i < hi; i++) {
        first_pos[k * nchunks + i] = run;
        run += p[i];
    }
    if (threadIdx.x == 0) {
        if (k == K - 1 || Compact<Kind>::list(k + 1) != l) counters[Compact<Kind>::counter(l)] = base + total;
        if (Kind == K_DONE && k == 2) counters[C_NEED] = total;  // finished paths taking new work
    }
}

// Writes the slots of each category in slot order. The block first places its entries in LDS, category
// after category (positions from a scan of the per-thread counts), then copies each category's run to
// its place in the output list (the chunk's first position from compact_scan_kernel) with coalesced
// stores.
template <int Kind, bool Multi>
__global__ __launch_bounds__(kCompactThreads) void compact_scatter_kernel(const uint8_t* stage, int n, const unsigned long long* items,
                                                                         const int* first_pos, int nchunks, QueueOut out) {
    constexpr int K = Compact<Kind>::cats;
    constexpr int W = kCompactThreads / 64;
    // A slot is in one category, except that a finished path can also have a shadow ray pending.
    constexpr int kMaxEntries = (Kind == K_DONE ? 2 : 1) * kCompactChunk;
    __shared__ int wave_tot[K][W];
    __shared__ int seg[K + 1];  // the block's entries of category k are buf[seg[k] .. seg[k + 1])
    __shared__ int dst[K];      // where they go in their list
    __shared__ int buf[kMaxEntries];
    const bool idle = idle_takes_work<Kind, Multi>(items);
    const int c = blockIdx.x;
    const int first = c * kCompactChunk + threadIdx.x * kCompactPerThread;
    uint8_t b[16];
    bool full;
    load_stages(stage, first, n, b, full);
    uint16_t m[16];
    int cnt[K] = {};
    for (int j = 0; j < 16; j++) {
        m[j] = uint16_t((first + j < n) ? Compact<Kind>::bits(b[j], idle) : 0u);
        for (int k = 0; k < K; k++) cnt[k] += (m[j] >> k) & 1;
    }
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    // Inclusive scans over the wave, then over the waves of the block.
    int incl[K];
    for (int k = 0; k < K; k++) {
        int v = cnt[k];
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(v, off);
            if (lane >= off) v += u;
        }
        incl[k] = v;
        if (lane == 63) wave_tot[k][wave] = v;
    }
    if (int(threadIdx.x) < K) dst[threadIdx.x] = first_pos[threadIdx.x * nchunks + c];
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < K; k++) {
            seg[k] = acc;
            for (int w = 0; w < W; w++) acc += wave_tot[k][w];
        }
        seg[K] = acc;
    }
    __syncthreads();
    int pos[K];
    for (int k = 0; k < K; k++) {
        int prev = 0;
        for (int w = 0; w < wave; w++) prev += wave_tot[k][w];
        pos[k] = seg[k] + prev + incl[k] - cnt[k];
    }
    for (int j = 0; j < 16; j++)
        for (int k = 0; k < K; k++)
            if ((m[j] >> k) & 1) buf[pos[k]++] = first + j;
    __syncthreads();
    const int total = seg[K];
    for (int e = threadIdx.x; e < total; e += kCompactThreads) {
        int k = 0;
        while (k + 1 < K && e >= seg[k + 1]) k++;
        out.q[Compact<Kind>::list(k)][dst[k] + (e - seg[k])] = buf[e];
    }
}

// Traversal of a queue of slots (trace_stream). One job: every lane shares the scene, whose traversal
// arrays sit in scalar registers and whose top `lds_nodes` nodes the block stages in LDS. Serving: each
// lane reads its slot's scene.
template <bool AnyHit, bool Count, bool Multi, class Stack, class Load, class Finish>
PH_DEV void trace_queue(const Work& work, const State& st, const int* queue, int n, int parts, int refill, int lds_nodes, float4* top, Stack& stack,
                        TraceCount& cnt, Load load, Finish finish) {
    TraceScene ts0{};
    if (!Multi) {
        const JobParams j0 = single_job<false>(work);
        ts0 = TraceScene{uniform_ptr(j0.scene.nodes), uniform_ptr(j0.scene.tris), &work.jobs[0].scene, top, min(lds_nodes, j0.scene.top_nodes)};
        stage_top_nodes(top, ts0.nodes, ts0.top_n);
    }
    int slot = 0, entry = 0;
    trace_stream<AnyHit, Count>(
        n, parts, refill, ts0, stack, cnt,
        [&](int i, TraceScene& ts, Walker& w, HitRec& hit) {
            slot = queue[i];
            if (Multi) {
                entry = st.job[slot];
                ts = trace_scene(work.jobs[entry].scene);
            }
            load(slot, w, hit);
        },
        [&](const TraceScene& ts, const Walker& w, const HitRec& hit) { finish(slot, entry, ts, w, hit); });
}

template <int Block, int Depth, bool Count, bool Multi>
__global__ __launch_bounds__(Block) void extend_kernel(Work work, State st, int* counters, int sort, int refill, int lds_nodes, TraceTotals* totals) {
    __shared__ int lds[Depth * Block];
    extern __shared__ float4 top_nodes[];
    SpillArea<Depth> spill;
    LaneStack<Depth> stack{lds + threadIdx.x, Block, spill.v};
    TraceCount cnt{0, 0, 0};
    trace_queue<false, Count, Multi>(
        work, st, st.q_ext, counters[C_EXT], counters[C_PARTS], refill, lds_nodes, top_nodes, stack, cnt,
        [&](int slot, Walker& w, HitRec& hit) {
            const float4 o4 = ld_stream(st.ro + slot), d4 = ld_stream(st.rd + slot);
            w = walker(xyz(o4), xyz(d4), kInf);
            hit = HitRec{kInf, -1, 0.0f, 0.0f};
        },
        [&](int slot, int entry, const TraceScene& ts, const Walker&, const HitRec& hit) {
            st_stream(st.hit + slot, make_float4(hit.t, __int_as_float(hit.tri), hit.u, hit.v));
            uint8_t next = S_MISS;
            if (hit.tri >= 0) {
                const int* cls = work.jobs[entry].mat_class;
                next = uint8_t(S_CLASS + (sort ? ldg(cls + __float_as_int(ldg(ts.tris + 3 * hit.tri).w)) : kClasses - 1));
            }
            st.stage[slot] = next;
        });
    if (Count) flush_counts(cnt, totals);
}

template <int Block, int Depth, bool Count, bool Multi>
__global__ __launch_bounds__(Block) void shadow_kernel(Work work, State st, int* counters, int refill, int lds_nodes, TraceTotals* totals) {
    __shared__ int lds[Depth * Block];
    extern __shared__ float4 top_nodes[];
    SpillArea<Depth> spill;
    LaneStack<Depth> stack{lds + threadIdx.x, Block, spill.v};
    TraceCount cnt{0, 0, 0};
    trace_queue<true, Count, Multi>(
        work, st, st.q_shadow, counters[C_SHADOW], counters[C_PARTS], refill, lds_nodes, top_nodes, stack, cnt,
        [&](int slot, Walker& w, HitRec& hit) {
            const float4 o4 = ld_stream(st.so + slot), d4 = ld_stream(st.sd + slot);
            w = walker(xyz(o4), xyz(d4), o4.w);
            hit = HitRec{kInf, -1, 0.0f, 0.0f};
        },
        [&](int slot, int, const TraceScene&, const Walker& w, const HitRec&) {
            if (w.found) return;
            const float4 c = ld_stream(st.sc + slot);
            float4 L = ld_stream(st.L + slot);
            L.x += c.x;
            L.y += c.y;
            L.z += c.z;
            st_stream(st.L + slot, L);
        });
    if (Count) flush_counts(cnt, totals);
}

template <bool Multi>
__global__ __launch_bounds__(256) void miss_kernel(Work work, State st, int* counters) {
    const int n = counters[C_MISS];
    const JobParams single = single_job<Multi>(work);
    const WaveShare ws = wave_share(n, counters[C_PARTS]);
    for (int b = ws.wave; b < ws.batches; b += ws.waves) {
        const int i = (ws.first + b) * 64 + lane_id();
        if (i < n) {
            const int slot = st.q_miss[i];
            const JobParams& jp = job_of<Multi>(work, st, slot, single);
            const DScene& s = jp.scene;
            const float4 o4 = ld_stream(st.ro + slot), d4 = ld_stream(st.rd + slot);
            const f3 rd = xyz(d4);
            const int flags = __float_as_int(d4.w);
            const int depth = flags & 0xFFFF;
            const bool prev_delta = (flags >> 16) & 1;
            const float4 b4 = ld_stream(st.beta + slot);
            const float4 L4 = ld_stream(st.L + slot);
            const f3 le = sky_radiance(s, rd);
            const float w = (prev_delta || s.p_env <= 0) ? 1.0f : mis_power(o4.w, sky_pdf(s, rd));
            const f3 L = xyz(L4) + clamp_contribution(xyz(b4) * le * w, depth, jp.opt.clamp);
            st_stream(st.L + slot, make_float4(L.x, L.y, L.z, 0.0f));
            if (depth == 0 && st.sample[slot] == 0 && jp.albedo) {
                const int p = st.pixel[slot];
                jp.albedo[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                jp.nd[p] = make_float4(-rd.x, -rd.y, -rd.z, 0.0f);
            }
            st.stage[slot] = term_stage<Multi>(jp, st, slot);
        }
    }
}

// One path vertex (integrator.h:178-260) for the slots of one material class.
template <int Lobes, bool Multi>
__global__ __launch_bounds__(256) void shade_kernel(Work work, State st, int* counters, int cls, int bins) {
    const int n = counters[C_CLASS + cls];
    const int* queue = st.q_class[cls];
    const JobParams single = single_job<Multi>(work);
    const WaveShare ws = wave_share(n, counters[C_PARTS]);
    for (int b = ws.wave; b < ws.batches; b += ws.waves) {
        const int i = (ws.first + b) * 64 + lane_id();
        if (i < n) {
            const int slot = queue[i];
            bool cont = false, shadow = false;
            int bin = 1;
            const JobParams& jp = job_of<Multi>(work, st, slot, single);
            const DScene& s = jp.scene;
            const PathOpts opt = jp.opt;
            const bool has_lights = s.p_env + s.p_emissive + s.p_punctual > 0;
            const int pixel = st.pixel[slot];
            const float4 o4 = ld_stream(st.ro + slot), d4 = ld_stream(st.rd + slot), h4 = ld_stream(st.hit + slot);
            const f3 rd = xyz(d4);
            const int flags = __float_as_int(d4.w);
            const int depth = flags & 0xFFFF;
            const bool prev_delta = (flags >> 16) & 1;
            const float prev_pdf = o4.w;
            f3 beta = xyz(ld_stream(st.beta + slot));
            f3 L = xyz(ld_stream(st.L + slot));
            const int tri = __float_as_int(h4.y);
            const Surface it = shade_point(s, xyz(o4), rd, h4.x, tri, h4.z, h4.w);
            const f3 wo_world = -rd;
            if (!black(it.emission)) {
                float w = 1.0f;
                if (!prev_delta) w = mis_power(prev_pdf, emitter_pdf(s, tri, rd, h4.x));
                L = L + clamp_contribution(beta * it.emission * w, depth, opt.clamp);
            }
            const Mix bsdf = make_mix(it.params);
            if (depth == 0 && st.sample[slot] == 0 && jp.albedo) {
                const f3 a = mix_albedo(bsdf);
                jp.albedo[pixel] = make_float4(a.x, a.y, a.z, 1.0f);
                jp.nd[pixel] = make_float4(it.ns.x, it.ns.y, it.ns.z, h4.x);
            }
            if (depth + 1 < opt.max_depth) {
                Pcg rng = slot_rng(st, slot, pixel);
                const f3 wo = to_local(it.frame, wo_world);
                const bool nee = opt.nee && has_lights && light_samplable(bsdf);
                if (nee) {
                    float ug = pcg_uniform(rng);
                    f2 ul = pcg_uniform2(rng);
                    float uk = pcg_uniform(rng);
                    LightDraw ls = draw_light(s, it.p, ug, ul, uk);
                    if (ls.pdf > 0 && !black(ls.radiance)) {
                        const f3 wi = to_local(it.frame, ls.wi);
                        float pdf_b;
                        const f3 f = mix_eval<Lobes>(bsdf, wo, wi, &pdf_b);
                        if (!black(f)) {
                            const f3 so = nudge(it.p, dot(ls.wi, it.ng) > 0 ? it.ng : -it.ng);
                            const float stmax = ls.dist == kInf ? kInf : ls.dist * (1.0f - 1e-4f);
                            const float w = ls.delta ? 1.0f : mis_power(ls.pdf, pdf_b);
                            const f3 c = clamp_contribution(beta * f * fabsf(wi.z) * ls.radiance * (w / ls.pdf), depth, opt.clamp);
                            st_stream(st.so + slot, make_float4(so.x, so.y, so.z, stmax));
                            st_stream(st.sd + slot, make_float4(ls.wi.x, ls.wi.y, ls.wi.z, 0.0f));
                            st_stream(st.sc + slot, make_float4(c.x, c.y, c.z, 0.0f));
                            shadow = true;
                        }
                    }
                }
                float u_lobe = pcg_uniform(rng);
                f2 u = pcg_uniform2(rng);
                float u_sub = pcg_uniform(rng);
                const Scatter bs = mix_sample<Lobes>(bsdf, wo, u_lobe, u, u_sub);
                if (bs.valid) {
                    beta = beta * (bs.f * (fabsf(bs.wi.z) / bs.pdf));
                    if (finite3(beta) && !black(beta)) {
                        const f3 wi_world = to_world(it.frame, bs.wi);
                        const float side = dot(wi_world, it.ng) * dot(wo_world, it.ng);
                        if (!(bs.transmitted ? side >= 0 : side <= 0)) {
                            const f3 no = nudge(it.p, dot(wi_world, it.ng) > 0 ? it.ng : -it.ng);
                            const f3 nd = unit(wi_world);
                            cont = true;
                            if (depth + 1 >= opt.rr_depth) {
                                float q = fmaxf(0.05f, 1.0f - maxc(beta));
                                if (pcg_uniform(rng) < q)
                                    cont = false;
                                else
                                    beta = beta * (1.0f / (1.0f - q));
                            }
                            if (cont) {
                                bin = ray_bin(nd, bins);
                                st_stream(st.ro + slot, make_float4(no.x, no.y, no.z, bs.pdf));
                                st_stream(st.rd + slot, make_float4(nd.x, nd.y, nd.z, __int_as_float((depth + 1) | ((nee ? 0 : 1) << 16))));
                                st_stream(st.beta + slot, make_float4(beta.x, beta.y, beta.z, 0.0f));
                            }
                        }
                    }
                }
                st.rng[slot] = rng.state;
            }
            st_stream(st.L + slot, make_float4(L.x, L.y, L.z, 0.0f));
            st.stage[slot] = uint8_t((cont ? ext_stage(bin) : term_stage<Multi>(jp, st, slot)) | (shadow ? S_SHADOW : 0));
        }
    }
}

// Finished paths: add the sample to its pixel (render_kernel's accumulation order), then start the
// pixel's next sample, or take the next work item, or go idle. Serving also lists idle slots here while
// work waits; they only take work.
template <bool Multi>
__global__ __launch_bounds__(256) void regen_kernel(Work work, State st, int* counters) {
    const int n = counters[C_TERM];
    // The queue's last C_NEED entries take new work items, in queue order: entry i gets item
    // base + (i - first_need). No atomics: advance_items_kernel moves the counter on afterwards.
    const int first_need = n - counters[C_NEED];
    const long long base = (long long)work.items[0];
    const long long total = (long long)work.items[1];
    const int chunks = counters[C_CHUNKS];
    const JobParams single = single_job<Multi>(work);
    const WaveShare ws = wave_share(n, counters[C_PARTS]);
    for (int b = ws.wave; b < ws.batches; b += ws.waves) {
        const int i = (ws.first + b) * 64 + lane_id();
        const bool active = i < n;
        const int slot = active ? st.q_term[i] : 0;
        bool need_work = false, ext = false, item_done = false;
        int job = 0;
        if (active && (!Multi || (st.stage[slot] & S_CODE) == S_TERM)) {
            const JobParams& jp = job_of<Multi>(work, st, slot, single);
            const int pixel = st.pixel[slot];
            int k = st.sample[slot];
            f3 L = xyz(ld_stream(st.L + slot));
            if (!finite3(L) || L.x < 0 || L.y < 0 || L.z < 0) L = mk3(0);
            const bool parts = Multi || jp.split > 1;
            float4* acc = jp.accum + (parts ? st.part[slot] * jp.pixels : 0) + pixel;
            float4 a = *acc;
            a.x += L.x;
            a.y += L.y;
            a.z += L.z;
            const float l = lum(L);
            a.w += l * l;
            *acc = a;
            k++;
            need_work = k >= (parts ? st.sample_end[slot] : jp.spp);
            if (need_work) {
                item_done = true;
                if (Multi) job = st.job[slot];
            } else {
                start_sample(jp, st, slot, pixel, k);
                ext = true;
            }
        } else if (active) {
            need_work = true;  // an idle slot
        }
        // Finished items per job (serving progress); a single render needs no count.
        if (Multi) wave_count(work.job_done, job, item_done);
        const long long w = base + (i - first_need);
        const bool got = need_work && i >= first_need && w < total;
        if (got) start_work<Multi>(work, st, chunks, slot, w, single);
        if (active) st.stage[slot] = (ext || got) ? ext_stage(0) : S_IDLE;
    }
}

// After regeneration: the items handed out by rank move the hand-out counter on.
__global__ void advance_items_kernel(unsigned long long* items, const int* counters) {
    if (threadIdx.x == 0) items[0] += (unsigned long long)counters[C_NEED];
}

// ---- the end of a single render ------------------------------------------------------------------------
// Once every work item is handed out, the paths still in flight drain over many bounces, each with few
// paths and the fixed cost of every launch. When the extension queue holds at most `limit` paths and no
// item is left, finish_kernel instead runs each of them to the end of its work item in one thread, as
// the per-pixel kernel runs a pixel: the path from its pending extension ray (trace_path_from, same
// vertices, same draws), then the item's remaining samples, accumulated as regen_kernel accumulates
// them. The queue is then cleared (finish_clear_kernel) and the rest of the iteration finds no work.
constexpr int kFinishBlock = 128, kFinishDepth = 32;

PH_DEV bool finish_now(const int* counters, const unsigned long long* items, int limit) {
    const int n = counters[C_EXT];
    return n > 0 && n <= limit && items[0] >= items[1];
}

// tally: [0] extension rays, [1] shadow rays, [2] paths finished. Lanes take queue entries as they become
// free (one atomic per wave and refill, on C_FINISH), and run one sample per loop turn, so a lane whose
// item ends takes the next entry instead of idling while its wave finishes longer paths.
__global__ __launch_bounds__(kFinishBlock) __attribute__((amdgpu_waves_per_eu(4))) void finish_kernel(Work work, State st, int* counters, int limit,
                                                                                                        unsigned long long* rays) {
    if (!finish_now(counters, work.items, limit)) return;
    __shared__ int lds[kFinishDepth * kFinishBlock];
    SpillArea<kFinishDepth> spill;
    LaneStack<kFinishDepth> stack{lds + threadIdx.x, kFinishBlock, spill.v};
    const int n = counters[C_EXT];
    const JobParams jp = single_job<false>(work);
    const bool parts = jp.split > 1;
    RayTally tally{0, 0};
    bool live = false, resumed = false;
    int slot = 0, pixel = 0, k = 0, end = 0;
    float4* acc = nullptr;
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (;;) {
        const uint64_t idle = __ballot(!live);
        if (idle) {
            const int leader = __ffsll(static_cast<unsigned long long>(idle)) - 1;
            int base = 0;
            if (lane_id() == leader) base = atomicAdd(counters + C_FINISH, __popcll(idle));
            base = __shfl(base, leader);
            const int i = base + __popcll(idle & ((1ull << lane_id()) - 1ull));
            if (!live && i < n) {
                slot = st.q_ext[i];
                pixel = st.pixel[slot];
                k = st.sample[slot];
                end = parts ? st.sample_end[slot] : jp.spp;
                acc = jp.accum + (parts ? st.part[slot] * jp.pixels : 0) + pixel;
                a = *acc;
                live = true;
                resumed = false;
            }
        }
        if (__ballot(live) == 0) break;
        if (live) {
            f3 L;
            if (!resumed) {
                // The path in flight, from its pending extension ray and its generator's current state.
                const float4 o4 = ld_stream(st.ro + slot), d4 = ld_stream(st.rd + slot);
                const int flags = __float_as_int(d4.w);
                const PathState ps{xyz(ld_stream(st.L + slot)), xyz(ld_stream(st.beta + slot)), o4.w, ((flags >> 16) & 1) != 0, flags & 0xFFFF};
                const bool first = ps.depth == 0 && k == 0 && jp.albedo;
                PrimaryOut prim{mk3(0), 0.0f, mk3(0), -1, mk3(1), mk3(0)};
                Pcg rng = slot_rng(st, slot, pixel);
                L = trace_path_from<LaneStack<kFinishDepth>, false>(jp.scene, xyz(o4), xyz(d4), rng, jp.opt, ps, first ? &prim : nullptr, stack,
                                                                    nullptr, &tally);
                if (first) {
                    jp.albedo[pixel] = make_float4(prim.albedo.x, prim.albedo.y, prim.albedo.z, 1.0f);
                    jp.nd[pixel] = make_float4(prim.normal.x, prim.normal.y, prim.normal.z, prim.depth);
                }
                resumed = true;
            } else {
                L = pixel_sample<LaneStack<kFinishDepth>, false>(jp.scene, jp.camera, pixel % jp.width, pixel / jp.width, jp.width, jp.height,
                                                                 jp.frame0 + uint32_t(k), jp.seed, jp.opt, nullptr, stack, nullptr, &tally);
            }
            if (!finite3(L) || L.x < 0 || L.y < 0 || L.z < 0) L = mk3(0);
            a.x += L.x;
            a.y += L.y;
            a.z += L.z;
            const float l = lum(L);
            a.w += l * l;
            if (++k >= end) {
                *acc = a;
                st.stage[slot] = S_IDLE;
                live = false;
            }
        }
    }
    const unsigned long long e = wave_sum(tally.ext), sh = wave_sum(tally.shadow);
    if (lane_id() == 0) {
        atomicAdd(rays, e);
        atomicAdd(rays + 1, sh);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(rays + 2, (unsigned long long)n);
}

__global__ void finish_clear_kernel(int* counters, const unsigned long long* items, int limit) {
    if (threadIdx.x == 0 && finish_now(counters, items, limit)) counters[C_EXT] = 0;
}

// The end of a single render in sub-streams: adds the partial sums to the accumulation, sub-stream 0
// first.
__global__ void add_parts_kernel(const float4* parts, int pixels, int count, float4* accum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    float4 a = accum[i];
    for (int p = 0; p < count; p++) {
        const float4 s = parts[size_t(p) * size_t(pixels) + size_t(i)];
        a.x += s.x;
        a.y += s.y;
        a.z += s.z;
        a.w += s.w;
    }
    accum[i] = a;
}

int device_cus() {
    int dev = 0, cus = 0;
    PH_CHECK(hipGetDevice(&dev));
    PH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return cus;
}

using TraceLaunch = void (*)(hipStream_t, int, Work, State, int*, int, int, int, TraceTotals*, bool);

template <int Block, int Depth, bool Multi>
void launch_extend(hipStream_t st, int grid, Work w, State s, int* counters, int sort, int refill, int lds_nodes, TraceTotals* tot, bool count) {
    const size_t lds = Multi ? 0 : size_t(lds_nodes) * 64;
    if (count)
        hipLaunchKernelGGL((extend_kernel<Block, Depth, true, Multi>), dim3(grid), dim3(Block), lds, st, w, s, counters, sort, refill, Multi ? 0 : lds_nodes,
                           tot);
    else
        hipLaunchKernelGGL((extend_kernel<Block, Depth, false, Multi>), dim3(grid), dim3(Block), lds, st, w, s, counters, sort, refill, Multi ? 0 : lds_nodes,
                           tot);
}

template <int Block, int Depth, bool Multi>
void launch_shadow(hipStream_t st, int grid, Work w, State s, int* counters, int, int refill, int lds_nodes, TraceTotals* tot, bool count) {
    const size_t lds = Multi ? 0 : size_t(lds_nodes) * 64;
    if (count)
        hipLaunchKernelGGL((shadow_kernel<Block, Depth, true, Multi>), dim3(grid), dim3(Block), lds, st, w, s, counters, refill, Multi ? 0 : lds_nodes, tot);
    else
        hipLaunchKernelGGL((shadow_kernel<Block, Depth, false, Multi>), dim3(grid), dim3(Block), lds, st, w, s, counters, refill, Multi ? 0 : lds_nodes, tot);
}

struct ShapeFns {
    TraceLaunch extend, shadow;
};

template <bool Multi>
bool shape_fns(int block, int depth, ShapeFns& f) {
#define PH_WF_SHAPE(B, D)                                                       \
    if (block == B && depth == D) {                                             \
        f = ShapeFns{&launch_extend<B, D, Multi>, &launch_shadow<B, D, Multi>}; \
        return true;                                                            \
    }
    PH_WF_SHAPE(64, 8)
    PH_WF_SHAPE(64, 16)
    PH_WF_SHAPE(64, 24)
    PH_WF_SHAPE(128, 8)
    PH_WF_SHAPE(128, 16)
    PH_WF_SHAPE(128, 24)
    PH_WF_SHAPE(128, 32)
    PH_WF_SHAPE(256, 8)
    PH_WF_SHAPE(256, 16)
    PH_WF_SHAPE(256, 24)
#undef PH_WF_SHAPE
    return false;
}

int material_class(const float4* m5) {
    const float metallic = m5[1].w, transmission = m5[2].z;
    const int mr_tex = __builtin_bit_cast(int, m5[4].x);
    // A textured metallic factor can take any value below the constant: treat it as mixed.
    if (mr_tex >= 0) return 3;
    const float m = fminf(fmaxf(metallic, 0.0f), 1.0f), t = fminf(fmaxf(transmission, 0.0f), 1.0f);
    const bool metal = m > 0, glass = (1 - m) * t > 0, opaque = (1 - m) * (1 - t) > 0;
    if (metal && !glass && !opaque) return 0;
    if (opaque && !metal && !glass) return 1;
    if (glass && !metal && !opaque) return 2;
    return 3;
}

}  // namespace

std::vector<int> material_classes(const DScene& scene, hipStream_t st) {
    const int nm = scene.num_materials;
    std::vector<float4> mats(size_t(nm) * 5);
    PH_CHECK(hipMemcpyAsync(mats.data(), scene.materials, mats.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    std::vector<int> cls(static_cast<size_t>(nm));
    for (int i = 0; i < nm; i++) cls[size_t(i)] = material_class(&mats[size_t(i) * 5]);
    return cls;
}

struct WavefrontEngine::Impl {
    ph_wavefront_config cfg{128, 24, 0, 16, 1, 1, 16, 16, 0, 1, 1, 1, 10};
    int slots = 0;
    int cus = 0;
    DevBuf<int> job, pixel, sample, sample_end, part, counters, q_ext, q_shadow, q_term, q_miss, q_cls[kClasses], mat_class, partial, first_pos;
    DevBuf<uint8_t> stage;
    DevBuf<int> job_done;
    DevBuf<long long> chunk_first;
    DevBuf<unsigned long long> items;  // next work item, items queued
    DevBuf<int4> chunk_pass;
    DevBuf<uint64_t> rng;
    DevBuf<float4> ro, rd, beta, L, hit, so, sd, sc;
    DevBuf<float4> part_sums;  // partial sums of a single render in sub-streams
    DevBuf<JobParams> jobs;
    DevBuf<TraceTotals> totals;             // [0] extension rays, [1] shadow rays
    DevBuf<unsigned long long> ray_totals;  // rays traced: [0] extension, [1] shadow; finish_kernel: [2] extension, [3] shadow, [4] paths
    int* host_counters = nullptr;
    int* host_done = nullptr;
    int* host_sizes = nullptr;
    unsigned long long* host_items = nullptr;  // after a batch: next item, items queued
    unsigned long long* host_stage = nullptr;  // staging of item counts for upload
    std::vector<hipEvent_t> events;
    hipGraphExec_t graph_exec = nullptr;
    std::tuple<int, int, int, int, int, bool> graph_key{};
    std::map<std::tuple<const void*, int, int, int, int>, ph_wavefront_config> tuned;
    // Serving bookkeeping (host side).
    bool serving = false;
    std::vector<bool> entry_used;
    std::vector<int> entry_pixels;
    // A queued pass: its unstarted items are [next, pixels) of the job's pixel order. While laid out,
    // its items occupy the segments (first, length) of the item index space, in pixel order: one
    // segment for a contiguous layout, one per round for a round-robin one.
    struct Pass {
        int entry, k0, k1, pixels, next;
        double priority;
        uint64_t order;
        std::vector<std::pair<long long, long long>> segments;
    };
    int share_quantum = 0;  // > 0: round-robin layout, this many items of each pass per round
    std::vector<Pass> passes;  // laid out in this order at the last run
    uint64_t pass_counter = 0;
    bool chunks_dirty = false;
    int64_t total_work = 0;  // end of the item layout

    Impl() {
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_counters), C_COUNT * sizeof(int), hipHostMallocDefault));
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_done), kMaxJobs * sizeof(int), hipHostMallocDefault));
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_sizes), 2 * sizeof(int), hipHostMallocDefault));
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_items), 2 * sizeof(unsigned long long), hipHostMallocDefault));
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_stage), 2 * sizeof(unsigned long long), hipHostMallocDefault));
        std::memset(host_items, 0, 2 * sizeof(unsigned long long));
        counters.allocate(C_COUNT);
        jobs.allocate(kMaxJobs);
        chunk_first.allocate(kMaxChunks);
        items.allocate(2);
        chunk_pass.allocate(kMaxChunks);
        job_done.allocate(kMaxJobs);
        totals.allocate(2);
        ray_totals.allocate(5);
        cus = device_cus();
        entry_used.assign(kMaxJobs, false);
        entry_pixels.assign(kMaxJobs, 0);
    }
    ~Impl() {
        drop_graph();
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (host_counters) (void)hipHostFree(host_counters);
        if (host_done) (void)hipHostFree(host_done);
        if (host_sizes) (void)hipHostFree(host_sizes);
        if (host_items) (void)hipHostFree(host_items);
        if (host_stage) (void)hipHostFree(host_stage);
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }

    void ensure_slots(int n) {
        if (n == slots) return;
        drop_graph();
        for (DevBuf<int>* b : {&job, &pixel, &sample, &sample_end, &part, &q_ext, &q_shadow, &q_term, &q_miss}) b->allocate(size_t(n));
        for (auto& q : q_cls) q.allocate(size_t(n));
        stage.allocate(size_t(n + 15) / 16 * 16);
        partial.allocate(size_t(kMaxCategories) * size_t(chunks_of(n)));
        first_pos.allocate(size_t(kMaxCategories) * size_t(chunks_of(n)));
        rng.allocate(size_t(n));
        for (DevBuf<float4>* b : {&ro, &rd, &beta, &L, &hit, &so, &sd, &sc}) b->allocate(size_t(n));
        slots = n;
    }

    State state() {
        State s;
        s.job = job.ptr;
        s.pixel = pixel.ptr;
        s.sample = sample.ptr;
        s.sample_end = sample_end.ptr;
        s.part = part.ptr;
        s.rng = rng.ptr;
        s.ro = ro.ptr;
        s.rd = rd.ptr;
        s.beta = beta.ptr;
        s.L = L.ptr;
        s.hit = hit.ptr;
        s.so = so.ptr;
        s.sd = sd.ptr;
        s.sc = sc.ptr;
        s.stage = stage.ptr;
        s.q_ext = q_ext.ptr;
        s.q_shadow = q_shadow.ptr;
        s.q_term = q_term.ptr;
        s.q_miss = q_miss.ptr;
        for (int c = 0; c < kClasses; c++) s.q_class[c] = q_cls[c].ptr;
        return s;
    }

    Work work() { return Work{jobs.ptr, items.ptr, chunk_first.ptr, chunk_pass.ptr, job_done.ptr}; }

    ShapeFns shapes(bool multi) const {
        ShapeFns f{};
        const bool ok = multi ? shape_fns<true>(cfg.block, cfg.stack_depth, f) : shape_fns<false>(cfg.block, cfg.stack_depth, f);
        if (!ok) throw std::invalid_argument("unsupported wavefront launch shape");
        return f;
    }

    static int chunks_of(int n) { return (n + kCompactChunk - 1) / kCompactChunk; }

    template <int Kind, bool Multi>
    void compact(hipStream_t st, const State& s) {
        const int nch = chunks_of(slots);
        QueueOut out{};
        if (Kind == K_EXT) out.q[0] = s.q_ext;
        if (Kind == K_HIT) {
            out.q[0] = s.q_miss;
            for (int c = 0; c < kClasses; c++) out.q[1 + c] = s.q_class[c];
        }
        if (Kind == K_DONE) {
            out.q[0] = s.q_shadow;
            out.q[1] = s.q_term;
        }
        hipLaunchKernelGGL((compact_count_kernel<Kind, Multi>), dim3(nch), dim3(kCompactThreads), 0, st, s.stage, slots, items.ptr, partial.ptr, nch);
        hipLaunchKernelGGL((compact_scan_kernel<Kind>), dim3(Compact<Kind>::cats), dim3(kScanThreads), 0, st, partial.ptr, first_pos.ptr, nch,
                           counters.ptr);
        hipLaunchKernelGGL((compact_scatter_kernel<Kind, Multi>), dim3(nch), dim3(kCompactThreads), 0, st, s.stage, slots, items.ptr, first_pos.ptr,
                           nch, out);
    }

    // One bounce. Event pairs bracket the traversal launches when timing.
    template <bool Multi>
    void iteration(hipStream_t st, const ShapeFns& fns, bool count, hipEvent_t* ev) {
        const int tgrid = cus * std::max(1, cfg.blocks_per_cu);
        // Shading: 4 blocks of 256 threads per CU. The shading kernels use 232 VGPRs (2 waves per SIMD), so
        // half of the blocks start as the others finish their share. Compiling them for 6 or 8 waves (with a
        // grid to match) spills and was slower: 940 and 897 against 995 Msamples/s.
        const int sgrid = cus * 4;
        const State s = state();
        const Work w = work();
        int* cnt = counters.ptr;
        hipLaunchKernelGGL(bookkeep_kernel, dim3(1), dim3(64), 0, st, cnt, ray_totals.ptr);
        if (ev) PH_CHECK(hipEventRecord(ev[0], st));
        const int refill = std::clamp(cfg.refill_lanes, 1, 64);
        fns.extend(st, tgrid, w, s, cnt, cfg.sort_materials ? 1 : 0, refill, cfg.lds_nodes, totals.ptr, count);
        if (ev) PH_CHECK(hipEventRecord(ev[1], st));
        compact<K_HIT, Multi>(st, s);
        hipLaunchKernelGGL(miss_kernel<Multi>, dim3(sgrid), dim3(256), 0, st, w, s, cnt);
        const int bins = cfg.ray_bins;
        hipLaunchKernelGGL((shade_kernel<1, Multi>), dim3(sgrid), dim3(256), 0, st, w, s, cnt, 0, bins);
        hipLaunchKernelGGL((shade_kernel<2, Multi>), dim3(sgrid), dim3(256), 0, st, w, s, cnt, 1, bins);
        hipLaunchKernelGGL((shade_kernel<4, Multi>), dim3(sgrid), dim3(256), 0, st, w, s, cnt, 2, bins);
        hipLaunchKernelGGL((shade_kernel<7, Multi>), dim3(sgrid), dim3(256), 0, st, w, s, cnt, 3, bins);
        compact<K_DONE, Multi>(st, s);
        if (ev) PH_CHECK(hipEventRecord(ev[2], st));
        fns.shadow(st, tgrid, w, s, cnt, 0, refill, cfg.lds_nodes, totals.ptr + 1, count);
        if (ev) PH_CHECK(hipEventRecord(ev[3], st));
        hipLaunchKernelGGL(regen_kernel<Multi>, dim3(sgrid), dim3(256), 0, st, w, s, cnt);
        hipLaunchKernelGGL(advance_items_kernel, dim3(1), dim3(64), 0, st, items.ptr, cnt);
        compact<K_EXT, Multi>(st, s);
        const int finish_limit = int(int64_t(slots) * std::clamp(cfg.finish_percent, 0, 100) / 100);
        if (!Multi && finish_limit > 0) {
            hipLaunchKernelGGL(finish_kernel, dim3(cus * 8), dim3(kFinishBlock), 0, st, w, s, cnt, finish_limit, ray_totals.ptr + 2);
            hipLaunchKernelGGL(finish_clear_kernel, dim3(1), dim3(64), 0, st, cnt, items.ptr, finish_limit);
        }
        PH_CHECK(hipGetLastError());
    }

    void ensure_events(int n) {
        for (size_t i = events.size(); i < size_t(n); i++) {
            hipEvent_t e;
            PH_CHECK(hipEventCreate(&e));
            events.push_back(e);
        }
    }

    // Runs G bounces (G even) from a graph or directly; returns (ms_extend, ms_shadow) when timed.
    std::pair<double, double> run_batch(hipStream_t st, int G, bool multi, bool counting) {
        const ShapeFns fns = shapes(multi);
        const bool graph = cfg.use_graph != 0 && !counting;
        if (graph) {
            auto key = std::make_tuple(cfg.block, cfg.stack_depth, cfg.blocks_per_cu, G, slots, multi);
            if (!graph_exec || key != graph_key) {
                drop_graph();
                hipGraph_t g = nullptr;
                PH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                for (int it = 0; it < G; it++) {
                    if (multi)
                        iteration<true>(st, fns, false, nullptr);
                    else
                        iteration<false>(st, fns, false, nullptr);
                }
                PH_CHECK(hipStreamEndCapture(st, &g));
                PH_CHECK(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                graph_key = key;
            }
            PH_CHECK(hipGraphLaunch(graph_exec, st));
            PH_CHECK(hipMemcpyAsync(host_counters, counters.ptr, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipMemcpyAsync(host_items, items.ptr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            PH_CHECK(hipStreamSynchronize(st));
            return {0.0, 0.0};
        }
        // Launched directly: HIP events bracket every traversal launch (graph-recorded events cannot
        // be timed, so a graph replay reports only its total).
        ensure_events(4 * G);
        for (int it = 0; it < G; it++) {
            if (multi)
                iteration<true>(st, fns, counting, &events[size_t(4 * it)]);
            else
                iteration<false>(st, fns, counting, &events[size_t(4 * it)]);
        }
        PH_CHECK(hipMemcpyAsync(host_counters, counters.ptr, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipMemcpyAsync(host_items, items.ptr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        double me = 0, ms = 0;
        for (int it = 0; it < G; it++) {
            float a = 0, b = 0;
            PH_CHECK(hipEventElapsedTime(&a, events[size_t(4 * it)], events[size_t(4 * it + 1)]));
            PH_CHECK(hipEventElapsedTime(&b, events[size_t(4 * it + 2)], events[size_t(4 * it + 3)]));
            me += a;
            ms += b;
        }
        return {me, ms};
    }

    void reset_counters(hipStream_t st) {
        PH_CHECK(hipMemsetAsync(counters.ptr, 0, C_COUNT * sizeof(int), st));
        PH_CHECK(hipStreamSynchronize(st));  // the pinned staging value may feed an earlier copy
        host_sizes[1] = cfg.xcd_parts == 8 ? 8 : 1;
        PH_CHECK(hipMemcpyAsync(counters.ptr + C_PARTS, host_sizes + 1, sizeof(int), hipMemcpyHostToDevice, st));
        PH_CHECK(hipMemsetAsync(items.ptr, 0, 2 * sizeof(unsigned long long), st));
        PH_CHECK(hipMemsetAsync(totals.ptr, 0, 2 * sizeof(TraceTotals), st));
        PH_CHECK(hipMemsetAsync(ray_totals.ptr, 0, 5 * sizeof(unsigned long long), st));
    }

    JobParams params_of(const WavefrontJob& job, const int* cls) const {
        JobParams jp{};
        jp.scene = job.scene;
        jp.camera = job.camera;
        jp.width = job.width;
        jp.height = job.height;
        jp.spp = job.spp;
        jp.pixels = job.width * job.height;
        jp.split = std::max(1, job.split);
        jp.span = job.spp / jp.split;
        jp.frame0 = job.frame0;
        jp.seed = job.seed;
        jp.opt = job.opt;
        jp.accum = job.accum;
        jp.albedo = job.albedo;
        jp.nd = job.nd;
        jp.mat_class = cls;
        return jp;
    }
};

WavefrontEngine::WavefrontEngine() : impl_(std::make_unique<Impl>()) {}
WavefrontEngine::~WavefrontEngine() = default;

void WavefrontEngine::set_config(const ph_wavefront_config& c) {
    ShapeFns f;
    if (!shape_fns<false>(c.block, c.stack_depth, f)) throw std::invalid_argument("unsupported block/stack depth for the wavefront engine");
    if (c.path_slots < 0 || c.blocks_per_cu <= 0 || c.blocks_per_cu > 64) throw std::invalid_argument("invalid wavefront configuration");
    if (c.graph_iterations < 1 || c.graph_iterations > 256) throw std::invalid_argument("graph_iterations must be in [1, 256]");
    if (c.refill_lanes < 1 || c.refill_lanes > 64) throw std::invalid_argument("refill_lanes must be in [1, 64]");
    if (c.ray_bins != 0 && c.ray_bins != 1) throw std::invalid_argument("ray_bins must be 0 or 1");
    if (c.sample_split < 1 || c.sample_split > 256) throw std::invalid_argument("sample_split must be in [1, 256]");
    if (c.xcd_parts != 1 && c.xcd_parts != 8) throw std::invalid_argument("xcd_parts must be 1 or 8");
    if (c.finish_percent < 0 || c.finish_percent > 100) throw std::invalid_argument("finish_percent must be in [0, 100]");
    if (c.lds_nodes < 0 || c.lds_nodes > DeviceScene::kTopNodes || size_t(c.block) * size_t(c.stack_depth) * 4 + size_t(c.lds_nodes) * 64 > 65536)
        throw std::invalid_argument("lds_nodes does not fit the block's LDS with its stacks");
    impl_->cfg = c;
    impl_->drop_graph();
}

ph_wavefront_config WavefrontEngine::config() const { return impl_->cfg; }

void WavefrontEngine::render(hipStream_t st, const WavefrontJob& job, bool counting, ph_render_stats& stats) {
    Impl& m = *impl_;
    stats = ph_render_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    const int pixels = job.width * job.height;
    if (pixels <= 0 || job.spp <= 0) return;
    if (m.serving) throw std::runtime_error("the engine is serving; a progressive render needs its own engine");
    // Sub-streams per pixel: the largest divisor of spp not above the configured split.
    int split = 1;
    for (int k = 1; k <= std::min(m.cfg.sample_split, job.spp); k++)
        if (job.spp % k == 0) split = k;
    const int64_t items = int64_t(pixels) * split;
    if (items > int64_t(INT32_MAX)) throw std::invalid_argument("pixels x sub-streams must stay below 2^31");
    // Paths in flight: one per pixel up to a cap of 4M slots (~900 MB of state), never more than items.
    const int slots = int(std::min<int64_t>(m.cfg.path_slots > 0 ? m.cfg.path_slots : std::min(pixels, 1 << 22), items));
    m.ensure_slots(slots);
    if (!job.mat_class) {
        std::vector<int> cls = material_classes(job.scene, st);
        m.mat_class.upload(cls.data(), cls.size(), st);
    }
    WavefrontJob sub = job;
    sub.split = split;
    if (split > 1) {
        m.part_sums.allocate(size_t(items));
        m.part_sums.zero(st);
        sub.accum = m.part_sums.ptr;
    }
    const JobParams jp = m.params_of(sub, job.mat_class ? job.mat_class : m.mat_class.ptr);
    PH_CHECK(hipMemcpyAsync(m.jobs.ptr, &jp, sizeof(jp), hipMemcpyHostToDevice, st));
    m.reset_counters(st);
    // Work items 0..slots-1 go to the slots directly; regeneration hands out the rest.
    PH_CHECK(hipStreamSynchronize(st));  // the staging buffer may still feed an earlier copy
    m.host_stage[0] = (unsigned long long)slots;
    m.host_stage[1] = (unsigned long long)items;
    PH_CHECK(hipMemcpyAsync(m.items.ptr, m.host_stage, 2 * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    PH_CHECK(hipMemsetAsync(m.job_done.ptr, 0, sizeof(int), st));
    hipLaunchKernelGGL(init_kernel, dim3((slots + 255) / 256), dim3(256), 0, st, m.work(), m.state(), slots);
    m.compact<K_EXT, false>(st, m.state());
    PH_CHECK(hipGetLastError());

    const int G = m.cfg.graph_iterations;
    // A sample takes at most max_depth bounces and a slot runs ceil(items / slots) items of spp/split samples.
    const int64_t bound = int64_t(job.spp / split) * int64_t(std::max(1, job.opt.max_depth)) * ((items + slots - 1) / slots) + 2 * G;
    double ms_extend = 0, ms_shadow = 0;
    int64_t iterations = 0;
    for (;;) {
        auto t = m.run_batch(st, G, false, counting);
        ms_extend += t.first;
        ms_shadow += t.second;
        iterations += G;
        if (m.host_counters[C_EXT] == 0) break;
        if (iterations > bound) throw std::runtime_error("wavefront loop did not drain");
    }
    if (split > 1) {
        hipLaunchKernelGGL(add_parts_kernel, dim3((pixels + 255) / 256), dim3(256), 0, st, m.part_sums.ptr, pixels, split, job.accum);
        PH_CHECK(hipGetLastError());
    }
    TraceTotals tot[2] = {};
    if (counting) PH_CHECK(hipMemcpy(tot, m.totals.ptr, sizeof(tot), hipMemcpyDeviceToHost));
    unsigned long long rays[5] = {0, 0, 0, 0, 0};
    PH_CHECK(hipMemcpy(rays, m.ray_totals.ptr, sizeof(rays), hipMemcpyDeviceToHost));
    stats.rays_extend = int64_t(rays[0]);
    stats.rays_shadow = int64_t(rays[1]) + m.host_counters[C_SHADOW];  // the last iteration's queue
    stats.rays_finish = int64_t(rays[2] + rays[3]);
    stats.finish_paths = int64_t(rays[4]);
    stats.ms_extend = ms_extend;
    stats.ms_shadow = ms_shadow;
    stats.iterations = iterations;
    stats.extend_launches = iterations;
    stats.shadow_launches = iterations;
    stats.shade_launches = iterations * kClasses;
    stats.samples = int64_t(pixels) * job.spp;
    stats.node_visits = int64_t(tot[0].nodes);
    stats.tri_tests = int64_t(tot[0].tris);
    stats.alpha_tests = int64_t(tot[0].alpha);
    stats.shadow_node_visits = int64_t(tot[1].nodes);
    stats.shadow_tri_tests = int64_t(tot[1].tris);
    stats.shadow_alpha_tests = int64_t(tot[1].alpha);
    stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

ph_wavefront_config WavefrontEngine::autotune(hipStream_t st, const WavefrontJob& job) {
    Impl& m = *impl_;
    auto key = std::make_tuple(static_cast<const void*>(job.scene.nodes), job.width, job.height, job.opt.max_depth, job.spp);
    auto found = m.tuned.find(key);
    if (found != m.tuned.end()) {
        set_config(found->second);
        return found->second;
    }
    const size_t accum_bytes = size_t(job.width) * size_t(job.height) * sizeof(float4);
    auto time_of = [&](const ph_wavefront_config& c) {
        set_config(c);
        ph_render_stats s{};
        render(st, job, false, s);  // warm: allocation, graph capture
        PH_CHECK(hipMemsetAsync(job.accum, 0, accum_bytes, st));
        const auto t0 = std::chrono::steady_clock::now();
        render(st, job, false, s);
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    ph_wavefront_config best = m.cfg;
    double best_ms = time_of(best);
    // A candidate replaces the current best only when it is faster by more than the run-to-run spread of
    // one render (about 1%), so that near-ties keep the earlier choice and repeated tunings agree.
    auto consider = [&](const ph_wavefront_config& c) {
        const double t = time_of(c);
        if (t < best_ms * 0.98) best_ms = t, best = c;
    };
    // Coordinate descent, largest effects first. Sub-streams per pixel (the caller's spp decides which
    // divide it) and paths in flight: with sub-streams there are more work items than pixels, more slots
    // than pixels pay, and every traversal launch gets more rays to hide latency behind.
    const int pixels = job.width * job.height;
    for (int split : {2, 4, 8, 16, 32, 64}) {
        if (split > job.spp || job.spp % split) continue;
        ph_wavefront_config c = best;
        c.sample_split = split;
        consider(c);
    }
    for (int64_t slots : {int64_t(pixels) / 4, int64_t(pixels) / 2, int64_t(0), int64_t(pixels) * 2, int64_t(pixels) * 4, int64_t(pixels) * 8,
                           int64_t(pixels) * 12, int64_t(pixels) * 15}) {
        if (slots > 0 && slots < 65536) continue;
        if (slots > pixels && (best.sample_split < 2 || slots > int64_t(pixels) * best.sample_split || slots > (int64_t(1) << 25))) continue;
        ph_wavefront_config c = best;
        c.path_slots = int(slots);
        consider(c);
    }
    // Traversal shape, persistent grid size, lane refill.
    for (int block : {64, 128, 256})
        for (int depth : {8, 16, 24}) {
            ph_wavefront_config c = best;
            c.block = block;
            c.stack_depth = depth;
            consider(c);
        }
    for (int bpc : {4, 8, 16, 24, 32}) {
        ph_wavefront_config c = best;
        c.blocks_per_cu = bpc;
        consider(c);
    }
    for (int refill : {4, 16, 32, 64}) {
        ph_wavefront_config c = best;
        c.refill_lanes = refill;
        consider(c);
    }
    for (int bins : {0, 1}) {
        ph_wavefront_config c = best;
        c.ray_bins = bins;
        consider(c);
    }
    for (int parts : {1, 8}) {
        ph_wavefront_config c = best;
        c.xcd_parts = parts;
        consider(c);
    }
    for (int top : {0, 128, 256, 512}) {
        ph_wavefront_config c = best;
        c.lds_nodes = top;
        if (size_t(c.block) * size_t(c.stack_depth) * 4 + size_t(top) * 64 > 65536) continue;
        consider(c);
    }
    // Graph replay and its length (the loop checks for an empty queue once per replay).
    for (int g : {0, 1}) {
        ph_wavefront_config c = best;
        c.use_graph = g;
        consider(c);
    }
    if (best.use_graph)
        for (int gi : {1, 2, 4, 5, 8, 10, 16}) {
            ph_wavefront_config c = best;
            c.graph_iterations = gi;
            consider(c);
        }
    // Path completion threshold: short renders (few bounces per path, few samples) gain from finishing more of
    // their paths one per thread.
    for (int fp : {0, 30, 60}) {
        ph_wavefront_config c = best;
        c.finish_percent = fp;
        consider(c);
    }
    set_config(best);
    m.tuned[key] = best;
    PH_CHECK(hipMemsetAsync(job.accum, 0, accum_bytes, st));
    PH_CHECK(hipStreamSynchronize(st));
    return best;
}

void WavefrontEngine::serve_begin(hipStream_t st, int slots) {
    Impl& m = *impl_;
    if (slots <= 0) throw std::invalid_argument("the serving engine needs at least one path slot");
    m.ensure_slots(slots);
    m.serving = true;
    m.entry_used.assign(kMaxJobs, false);
    m.entry_pixels.assign(kMaxJobs, 0);
    m.passes.clear();
    m.chunks_dirty = false;
    m.total_work = 0;
    m.reset_counters(st);
    PH_CHECK(hipMemsetAsync(m.job_done.ptr, 0, kMaxJobs * sizeof(int), st));
    hipLaunchKernelGGL(init_idle_kernel, dim3((slots + 255) / 256), dim3(256), 0, st, m.state(), slots);
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipMemcpyAsync(m.host_counters, m.counters.ptr, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    std::memset(m.host_done, 0, kMaxJobs * sizeof(int));
    m.host_items[0] = m.host_items[1] = 0;
}

int WavefrontEngine::serve_admit(hipStream_t st, const WavefrontJob& job) {
    Impl& m = *impl_;
    if (!m.serving) throw std::runtime_error("serve_begin was not called");
    if (job.width <= 0 || job.height <= 0 || job.spp <= 0) throw std::invalid_argument("empty job");
    if (!job.mat_class) throw std::invalid_argument("serving jobs need a per-scene material class table");
    if (job.split < 1 || job.spp % job.split) throw std::invalid_argument("the sub-stream count must divide the sample count");
    if (int64_t(job.width) * job.height * job.split >= int64_t(INT32_MAX)) throw std::invalid_argument("job too large");
    int entry = -1;
    for (int e = 0; e < kMaxJobs; e++)
        if (!m.entry_used[size_t(e)]) {
            entry = e;
            break;
        }
    if (entry < 0) return -1;
    const JobParams jp = m.params_of(job, job.mat_class);
    PH_CHECK(hipMemcpyAsync(m.jobs.ptr + entry, &jp, sizeof(jp), hipMemcpyHostToDevice, st));
    PH_CHECK(hipMemsetAsync(m.job_done.ptr + entry, 0, sizeof(int), st));
    m.host_done[entry] = 0;
    m.entry_used[size_t(entry)] = true;
    m.entry_pixels[size_t(entry)] = job.width * job.height * job.split;
    return entry;
}

bool WavefrontEngine::serve_enqueue(int entry, int k0, int k1, double priority) {
    Impl& m = *impl_;
    if (entry < 0 || entry >= kMaxJobs || !m.entry_used[size_t(entry)]) throw std::invalid_argument("not an admitted job");
    if (k0 < 0 || k1 <= k0) throw std::invalid_argument("empty pass");
    if (int(m.passes.size()) >= kMaxChunks) return false;
    m.passes.push_back(Impl::Pass{entry, k0, k1, m.entry_pixels[size_t(entry)], 0, priority, m.pass_counter++, {}});
    m.chunks_dirty = true;
    return true;
}

void WavefrontEngine::serve_reprioritize(int entry, double priority) {
    Impl& m = *impl_;
    for (Impl::Pass& p : m.passes)
        if (p.entry == entry && p.priority != priority) {
            p.priority = priority;
            m.chunks_dirty = true;
        }
}

void WavefrontEngine::serve_share(int quantum) {
    impl_->share_quantum = std::max(0, quantum);
    impl_->chunks_dirty = true;
}

void WavefrontEngine::serve_run(hipStream_t st, int iterations) {
    Impl& m = *impl_;
    if (!m.serving) throw std::runtime_error("serve_begin was not called");
    // Items below the counter have been handed to slots (and items past the layout's end were asked
    // for when there was none). Account for what each laid-out pass handed out, drop finished passes,
    // then lay the rest out again from the counter on, by priority.
    const long long handed = (long long)m.host_items[0];
    bool changed = m.chunks_dirty;
    std::vector<Impl::Pass> keep;
    for (Impl::Pass p : m.passes) {
        long long got = 0;
        for (const auto& [first, len] : p.segments) got += std::clamp<long long>(handed - first, 0, len);
        if (got > 0) changed = true;
        p.next += int(got);
        if (p.next < p.pixels) keep.push_back(p);
    }
    if (changed) {
        std::stable_sort(keep.begin(), keep.end(),
                         [](const Impl::Pass& a, const Impl::Pass& b) { return a.priority != b.priority ? a.priority < b.priority : a.order < b.order; });
        std::vector<long long> starts;
        std::vector<int4> descs;
        long long at = handed;
        auto place = [&](Impl::Pass& p, int from, long long count) {
            p.segments.push_back({at, count});
            starts.push_back(at);
            descs.push_back(make_int4(p.entry, p.k0, p.k1, from));
            at += count;
        };
        std::vector<int> pos(keep.size());
        for (size_t i = 0; i < keep.size(); i++) {
            pos[i] = keep[i].next;
            keep[i].segments.clear();
        }
        if (m.share_quantum > 0) {
            // Processor sharing: rounds of `share_quantum` items from every pass with work left, so slots
            // freed during a run go to the running jobs in equal shares. When the chunk table would
            // overflow, the rest of each pass follows in one piece, in priority order.
            const size_t budget = size_t(kMaxChunks) - keep.size();
            bool left = true;
            while (left && starts.size() + keep.size() <= budget) {
                left = false;
                for (size_t i = 0; i < keep.size(); i++) {
                    const long long rest = keep[i].pixels - pos[i];
                    if (rest <= 0) continue;
                    const long long q = std::min<long long>(rest, m.share_quantum);
                    place(keep[i], pos[i], q);
                    pos[i] += int(q);
                    left = left || q < rest;
                }
            }
        }
        for (size_t i = 0; i < keep.size(); i++)
            if (keep[i].pixels > pos[i]) place(keep[i], pos[i], keep[i].pixels - pos[i]);
        if (starts.size() > size_t(kMaxChunks)) throw std::runtime_error("serving: more queued passes than the chunk table holds");
        m.total_work = at;
        if (!starts.empty()) {
            PH_CHECK(hipMemcpyAsync(m.chunk_first.ptr, starts.data(), starts.size() * sizeof(long long), hipMemcpyHostToDevice, st));
            PH_CHECK(hipMemcpyAsync(m.chunk_pass.ptr, descs.data(), descs.size() * sizeof(int4), hipMemcpyHostToDevice, st));
        }
        // Staged in pinned memory (the stream was synchronized at the end of the previous run).
        m.host_stage[1] = (unsigned long long)m.total_work;
        m.host_sizes[0] = int(starts.size());
        PH_CHECK(hipMemcpyAsync(m.items.ptr + 1, m.host_stage + 1, sizeof(unsigned long long), hipMemcpyHostToDevice, st));
        PH_CHECK(hipMemcpyAsync(m.counters.ptr + C_CHUNKS, m.host_sizes, sizeof(int), hipMemcpyHostToDevice, st));
        // The tables are read by the run below; the host vectors must outlive the copies.
        PH_CHECK(hipStreamSynchronize(st));
    }
    m.passes.swap(keep);
    m.chunks_dirty = false;
    m.run_batch(st, std::max(1, iterations), true, false);
    PH_CHECK(hipMemcpyAsync(m.host_done, m.job_done.ptr, kMaxJobs * sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
}

int WavefrontEngine::serve_pixels_done(int entry) const { return impl_->host_done[entry]; }

void WavefrontEngine::serve_release(int entry) {
    if (entry >= 0 && entry < kMaxJobs) impl_->entry_used[size_t(entry)] = false;
}

int WavefrontEngine::serve_in_flight() const { return impl_->host_counters[C_EXT]; }

int64_t WavefrontEngine::serve_work_remaining() const { return impl_->total_work - (long long)impl_->host_items[0]; }

}  // namespace phk
