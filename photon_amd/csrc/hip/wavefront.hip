// Wavefront path tracing engine.
//
// Paths in flight live in structure-of-arrays state in HBM, one "slot" per path. A slot owns one pixel
// at a time and runs that pixel's samples one after another, so each pixel's accumulation is the same
// sequence of float additions as in the one-thread-per-pixel kernel (bit-identical images); when a
// pixel's samples are done the slot takes the next pixel from a global work counter (regeneration).
//
// One bounce of every path in flight is one iteration of six stages:
//   bookkeep  reset the per-iteration queue counters and work cursors (one wave)
//   extend    closest-hit traversal of the extension queue; hits are binned by material class
//   miss      environment contribution of escaped paths
//   shade x4  surface interaction, NEE light sample, BSDF sample, Russian roulette; one launch per
//             material class (metal / opaque / glass / mixed) so a wave runs one lobe structure
//   shadow    any-hit traversal of the NEE queue; unoccluded contributions are added to L
//   regen     terminated paths: add L to the pixel, start the next sample or the next pixel
// The traversal, shading and regeneration kernels are persistent: a fixed grid whose waves pull
// batches of 64 queue entries from a cursor until the queue is drained, so uneven rays never leave a
// wave idle. Queue appends are wave-aggregated (one atomic per wave). The bounce loop can be captured
// once into a HIP graph of G iterations and replayed until the queues are empty.
#include <chrono>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "wavefront.h"

using namespace phd;

namespace phk {

namespace {

enum : int {
    C_EXT0 = 0,
    C_EXT1 = 1,
    C_SHADOW = 2,
    C_TERM = 3,
    C_MISS = 4,
    C_CLASS = 5,  // 4 counters: metal, opaque, glass, mixed
    C_WORK = 9,   // next work item (pixel) to hand out
    C_CUR = 10,   // cursors: 10 extend, 11 miss, 12..15 shade, 16 shadow, 17 regen
    C_RETIRED = 18,
    C_COUNT = 32
};

constexpr int kClasses = 4;

struct JobParams {
    DScene scene;
    ph_camera_view camera;
    int width, height, spp, total_work;
    uint32_t frame0, seed;
    PathOpts opt;
    float4* accum;
    float4* albedo;
    float4* nd;
    const int* mat_class;  // per material: index into the class queues
};

struct State {
    int* pixel;
    int* sample;
    uint64_t* rng;
    float4* ro;  // origin, prev_pdf
    float4* rd;  // direction, flags (depth | prev_delta << 16)
    float4* beta;
    float4* L;
    float4* hit;
    float4* so;  // shadow origin, tmax
    float4* sd;  // shadow direction
    float4* sc;  // shadow contribution
    int* q[2];
    int* q_shadow;
    int* q_term;
    int* q_miss;
    int* q_class[kClasses];
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

// Appends `value` for every lane with `pred` using one atomic per wave. All lanes of the wave must call.
PH_DEV void wave_append(int* queue, int* count, bool pred, int value) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = lane_id();
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, leader);
    if (pred) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Persistent work distribution: a wave takes 64 consecutive queue entries per round.
PH_DEV int wave_fetch(int* cursor) {
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(cursor, 64);
    return __shfl(base, 0);
}

// Work item -> pixel: 8-row bands, 8-column tiles inside a band (the last tile and band may be
// narrower), row-major inside a tile, so the 64 lanes of a wave start on an 8x8 pixel block.
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
    st.ro[slot] = make_float4(o.x, o.y, o.z, 0.0f);
    st.rd[slot] = make_float4(d.x, d.y, d.z, __int_as_float(1 << 16));
    st.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    st.L[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

PH_DEV Pcg slot_rng(const State& st, int slot, int pixel) { return Pcg{st.rng[slot], (uint64_t(uint32_t(pixel)) << 1u) | 1u}; }

__global__ void init_kernel(const JobParams* jpp, State st, int* counters, int slots) {
    const JobParams& jp = *jpp;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const bool has = slot < slots && slot < jp.total_work;
    if (has) start_sample(jp, st, slot, work_pixel(slot, jp.width, jp.height), 0);
    wave_append(st.q[0], counters + C_EXT0, has, slot);
}

// Starts iteration `par`: folds the ray counts (this iteration's extension queue, the previous
// iteration's shadow queue) into the 64-bit totals, then clears the per-iteration counters.
__global__ void bookkeep_kernel(int* counters, int par, unsigned long long* rays) {
    const int t = threadIdx.x;
    if (t == 0) {
        rays[0] += (unsigned long long)counters[C_EXT0 + par];
        rays[1] += (unsigned long long)counters[C_SHADOW];
    }
    __syncthreads();
    if (t == 0) counters[C_EXT0 + (1 - par)] = 0;
    if (t >= C_SHADOW && t < C_WORK) counters[t] = 0;
    if (t >= C_CUR && t < C_CUR + 8) counters[t] = 0;
}

template <int Block, int Depth, bool Count>
__global__ __launch_bounds__(Block) void extend_kernel(const JobParams* jpp, State st, int* counters, int par, TraceTotals* totals) {
    __shared__ int lds[Depth * Block];
    LaneStack<Depth> stack{lds + threadIdx.x, Block};
    const DScene s = jpp->scene;
    const int* mat_class = jpp->mat_class;
    const int n = counters[C_EXT0 + par];
    const int* queue = st.q[par];
    TraceCount cnt{0, 0, 0};
    for (;;) {
        const int base = wave_fetch(counters + C_CUR + 0);
        if (base >= n) break;
        const int i = base + lane_id();
        const bool active = i < n;
        int slot = 0, cls = -1;
        if (active) {
            slot = queue[i];
            const float4 o4 = st.ro[slot], d4 = st.rd[slot];
            HitRec hit{kInf, -1, 0.0f, 0.0f};
            traverse<false, Count>(s, xyz(o4), xyz(d4), kInf, hit, stack, &cnt);
            st.hit[slot] = make_float4(hit.t, __int_as_float(hit.tri), hit.u, hit.v);
            if (hit.tri >= 0) cls = mat_class[__float_as_int(s.tris[3 * hit.tri].w)];
        }
        wave_append(st.q_miss, counters + C_MISS, active && cls < 0, slot);
        for (int c = 0; c < kClasses; c++) wave_append(st.q_class[c], counters + C_CLASS + c, active && cls == c, slot);
    }
    if (Count) flush_counts(cnt, totals);
}

template <int Block, int Depth, bool Count>
__global__ __launch_bounds__(Block) void shadow_kernel(const JobParams* jpp, State st, int* counters, TraceTotals* totals) {
    __shared__ int lds[Depth * Block];
    LaneStack<Depth> stack{lds + threadIdx.x, Block};
    const DScene s = jpp->scene;
    const int n = counters[C_SHADOW];
    TraceCount cnt{0, 0, 0};
    for (;;) {
        const int base = wave_fetch(counters + C_CUR + 6);
        if (base >= n) break;
        const int i = base + lane_id();
        if (i < n) {
            const int slot = st.q_shadow[i];
            const float4 o4 = st.so[slot], d4 = st.sd[slot];
            HitRec hit{kInf, -1, 0.0f, 0.0f};
            if (!traverse<true, Count>(s, xyz(o4), xyz(d4), o4.w, hit, stack, &cnt)) {
                const float4 c = st.sc[slot];
                float4 L = st.L[slot];
                L.x += c.x;
                L.y += c.y;
                L.z += c.z;
                st.L[slot] = L;
            }
        }
    }
    if (Count) flush_counts(cnt, totals);
}

PH_DEV f3 add_clamped(f3 L, f3 c, int depth, float clamp) { return L + clamp_contribution(c, depth, clamp); }

__global__ __launch_bounds__(256) void miss_kernel(const JobParams* jpp, State st, int* counters) {
    const JobParams& jp = *jpp;
    const DScene& s = jp.scene;
    const int n = counters[C_MISS];
    for (;;) {
        const int base = wave_fetch(counters + C_CUR + 1);
        if (base >= n) break;
        const int i = base + lane_id();
        const bool active = i < n;
        int slot = 0;
        if (active) {
            slot = st.q_miss[i];
            const float4 o4 = st.ro[slot], d4 = st.rd[slot];
            const f3 rd = xyz(d4);
            const int flags = __float_as_int(d4.w);
            const int depth = flags & 0xFFFF;
            const bool prev_delta = (flags >> 16) & 1;
            const float4 b4 = st.beta[slot];
            float4 L4 = st.L[slot];
            const f3 le = sky_radiance(s, rd);
            const float w = (prev_delta || s.p_env <= 0) ? 1.0f : mis_power(o4.w, sky_pdf(s, rd));
            const f3 L = add_clamped(xyz(L4), xyz(b4) * le * w, depth, jp.opt.clamp);
            st.L[slot] = make_float4(L.x, L.y, L.z, 0.0f);
            if (depth == 0 && st.sample[slot] == 0) {
                const int p = st.pixel[slot];
                jp.albedo[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                jp.nd[p] = make_float4(-rd.x, -rd.y, -rd.z, 0.0f);
            }
        }
        wave_append(st.q_term, counters + C_TERM, active, slot);
    }
}

// One path vertex (integrator.h:178-260) for the slots of one material class.
template <int Lobes>
__global__ __launch_bounds__(256) void shade_kernel(const JobParams* jpp, State st, int* counters, int par, int cls) {
    const JobParams& jp = *jpp;
    const DScene& s = jp.scene;
    const PathOpts opt = jp.opt;
    const int n = counters[C_CLASS + cls];
    const int* queue = st.q_class[cls];
    const bool has_lights = s.p_env + s.p_emissive + s.p_punctual > 0;
    for (;;) {
        const int base = wave_fetch(counters + C_CUR + 2 + cls);
        if (base >= n) break;
        const int i = base + lane_id();
        const bool active = i < n;
        int slot = 0;
        bool cont = false, shadow = false;
        if (active) {
            slot = queue[i];
            const int pixel = st.pixel[slot];
            const float4 o4 = st.ro[slot], d4 = st.rd[slot], h4 = st.hit[slot];
            const f3 rd = xyz(d4);
            const int flags = __float_as_int(d4.w);
            const int depth = flags & 0xFFFF;
            const bool prev_delta = (flags >> 16) & 1;
            const float prev_pdf = o4.w;
            f3 beta = xyz(st.beta[slot]);
            f3 L = xyz(st.L[slot]);
            const int tri = __float_as_int(h4.y);
            const Surface it = shade_point(s, xyz(o4), rd, h4.x, tri, h4.z, h4.w);
            const f3 wo_world = -rd;
            if (!black(it.emission)) {
                float w = 1.0f;
                if (!prev_delta) w = mis_power(prev_pdf, emitter_pdf(s, tri, rd, h4.x));
                L = add_clamped(L, beta * it.emission * w, depth, opt.clamp);
            }
            const Mix bsdf = make_mix(it.params);
            if (depth == 0 && st.sample[slot] == 0) {
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
                            st.so[slot] = make_float4(so.x, so.y, so.z, stmax);
                            st.sd[slot] = make_float4(ls.wi.x, ls.wi.y, ls.wi.z, 0.0f);
                            st.sc[slot] = make_float4(c.x, c.y, c.z, 0.0f);
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
                                st.ro[slot] = make_float4(no.x, no.y, no.z, bs.pdf);
                                st.rd[slot] = make_float4(nd.x, nd.y, nd.z, __int_as_float((depth + 1) | ((nee ? 0 : 1) << 16)));
                                st.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.0f);
                            }
                        }
                    }
                }
                st.rng[slot] = rng.state;
            }
            st.L[slot] = make_float4(L.x, L.y, L.z, 0.0f);
        }
        wave_append(st.q_shadow, counters + C_SHADOW, active && shadow, slot);
        wave_append(st.q[1 - par], counters + C_EXT0 + (1 - par), active && cont, slot);
        wave_append(st.q_term, counters + C_TERM, active && !cont, slot);
    }
}

// Finished paths: add the sample to its pixel (render_kernel's accumulation order), then start the
// pixel's next sample or take the next pixel.
__global__ __launch_bounds__(256) void regen_kernel(const JobParams* jpp, State st, int* counters, int par) {
    const JobParams& jp = *jpp;
    const int n = counters[C_TERM];
    for (;;) {
        const int base = wave_fetch(counters + C_CUR + 7);
        if (base >= n) break;
        const int i = base + lane_id();
        const bool active = i < n;
        int slot = 0, pixel = -1, k = 0;
        bool need_pixel = false;
        if (active) {
            slot = st.q_term[i];
            pixel = st.pixel[slot];
            k = st.sample[slot];
            f3 L = xyz(st.L[slot]);
            if (!finite3(L) || L.x < 0 || L.y < 0 || L.z < 0) L = mk3(0);
            float4 a = jp.accum[pixel];
            a.x += L.x;
            a.y += L.y;
            a.z += L.z;
            const float l = lum(L);
            a.w += l * l;
            jp.accum[pixel] = a;
            k++;
            need_pixel = k >= jp.spp;
        }
        // Lanes whose pixel is finished take new work items with one atomic per wave.
        const uint64_t mask = __ballot(need_pixel);
        int work = 0;
        if (mask) {
            const int lane = lane_id();
            const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
            int first = 0;
            if (lane == leader) first = atomicAdd(counters + C_WORK, __popcll(mask));
            first = __shfl(first, leader);
            work = first + __popcll(mask & ((1ull << lane) - 1ull));
        }
        bool again = false;
        if (active) {
            if (need_pixel) {
                if (work < jp.total_work) {
                    start_sample(jp, st, slot, work_pixel(work, jp.width, jp.height), 0);
                    again = true;
                }
            } else {
                start_sample(jp, st, slot, pixel, k);
                again = true;
            }
        }
        wave_append(st.q[1 - par], counters + C_EXT0 + (1 - par), again, slot);
    }
}

int device_cus() {
    int dev = 0, cus = 0;
    PH_CHECK(hipGetDevice(&dev));
    PH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return cus;
}

using TraceLaunch = void (*)(hipStream_t, int, const JobParams*, State, int*, int, TraceTotals*, bool);

template <int Block, int Depth>
void launch_extend(hipStream_t st, int grid, const JobParams* jp, State s, int* counters, int par, TraceTotals* tot, bool count) {
    if (count)
        hipLaunchKernelGGL((extend_kernel<Block, Depth, true>), dim3(grid), dim3(Block), 0, st, jp, s, counters, par, tot);
    else
        hipLaunchKernelGGL((extend_kernel<Block, Depth, false>), dim3(grid), dim3(Block), 0, st, jp, s, counters, par, tot);
}

template <int Block, int Depth>
void launch_shadow(hipStream_t st, int grid, const JobParams* jp, State s, int* counters, int, TraceTotals* tot, bool count) {
    if (count)
        hipLaunchKernelGGL((shadow_kernel<Block, Depth, true>), dim3(grid), dim3(Block), 0, st, jp, s, counters, tot);
    else
        hipLaunchKernelGGL((shadow_kernel<Block, Depth, false>), dim3(grid), dim3(Block), 0, st, jp, s, counters, tot);
}

struct ShapeFns {
    TraceLaunch extend, shadow;
};

bool shape_fns(int block, int depth, ShapeFns& f) {
#define PH_WF_SHAPE(B, D)                                       \
    if (block == B && depth == D) {                             \
        f = ShapeFns{&launch_extend<B, D>, &launch_shadow<B, D>}; \
        return true;                                            \
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

// Material -> class queue index: 0 metal only, 1 opaque only, 2 glass only, 3 anything else.
// A textured metallic factor can take any value below the constant, so it counts as mixed.
int material_class(const float4* m5) {
    const float metallic = m5[1].w, transmission = m5[2].z;
    const int mr_tex = __builtin_bit_cast(int, m5[4].x);
    if (mr_tex >= 0) return 3;
    const float m = fminf(fmaxf(metallic, 0.0f), 1.0f), t = fminf(fmaxf(transmission, 0.0f), 1.0f);
    const bool metal = m > 0, glass = (1 - m) * t > 0, opaque = (1 - m) * (1 - t) > 0;
    if (metal && !glass && !opaque) return 0;
    if (opaque && !metal && !glass) return 1;
    if (glass && !metal && !opaque) return 2;
    return 3;
}

}  // namespace

struct WavefrontEngine::Impl {
    ph_wavefront_config cfg{128, 16, 0, 8, 1, 1, 16};
    int slots = 0;
    DevBuf<int> pixel, sample, counters, q0, q1, q_shadow, q_term, q_miss, q_cls[kClasses], mat_class;
    DevBuf<uint64_t> rng;
    DevBuf<float4> ro, rd, beta, L, hit, so, sd, sc;
    DevBuf<JobParams> params;
    DevBuf<TraceTotals> totals;
    DevBuf<unsigned long long> ray_totals;  // [0] extension rays traced, [1] shadow rays traced
    int* host_counters = nullptr;
    int cus = 0;
    std::vector<hipEvent_t> events;
    // Captured bounce loop and the key it was captured for.
    hipGraphExec_t graph_exec = nullptr;
    std::tuple<int, int, int, int, int, int, bool, bool, int> graph_key{};
    std::map<std::tuple<const void*, int, int, int>, ph_wavefront_config> tuned;

    Impl() {
        PH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_counters), C_COUNT * sizeof(int), hipHostMallocDefault));
        counters.allocate(C_COUNT);
        params.allocate(1);
        totals.allocate(2);  // [0] extension rays, [1] shadow rays
        ray_totals.allocate(2);
        cus = device_cus();
    }
    ~Impl() {
        drop_graph();
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (host_counters) (void)hipHostFree(host_counters);
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }

    void ensure_slots(int n) {
        if (n == slots) return;
        drop_graph();
        for (DevBuf<int>* b : {&pixel, &sample, &q0, &q1, &q_shadow, &q_term, &q_miss}) b->allocate(size_t(n));
        for (auto& q : q_cls) q.allocate(size_t(n));
        rng.allocate(size_t(n));
        for (DevBuf<float4>* b : {&ro, &rd, &beta, &L, &hit, &so, &sd, &sc}) b->allocate(size_t(n));
        slots = n;
    }

    State state() {
        State s;
        s.pixel = pixel.ptr;
        s.sample = sample.ptr;
        s.rng = rng.ptr;
        s.ro = ro.ptr;
        s.rd = rd.ptr;
        s.beta = beta.ptr;
        s.L = L.ptr;
        s.hit = hit.ptr;
        s.so = so.ptr;
        s.sd = sd.ptr;
        s.sc = sc.ptr;
        s.q[0] = q0.ptr;
        s.q[1] = q1.ptr;
        s.q_shadow = q_shadow.ptr;
        s.q_term = q_term.ptr;
        s.q_miss = q_miss.ptr;
        for (int c = 0; c < kClasses; c++) s.q_class[c] = q_cls[c].ptr;
        return s;
    }

    // One bounce (iteration parity `par`). Event pairs bracket the traversal launches when timing.
    void iteration(hipStream_t st, const ShapeFns& fns, int par, bool count, hipEvent_t* ev) {
        const int tgrid = cus * std::max(1, cfg.blocks_per_cu);
        const int sgrid = cus * 4;
        State s = state();
        const JobParams* jp = params.ptr;
        int* cnt = counters.ptr;
        hipLaunchKernelGGL(bookkeep_kernel, dim3(1), dim3(64), 0, st, cnt, par, ray_totals.ptr);
        if (ev) PH_CHECK(hipEventRecord(ev[0], st));
        fns.extend(st, tgrid, jp, s, cnt, par, totals.ptr, count);
        if (ev) PH_CHECK(hipEventRecord(ev[1], st));
        hipLaunchKernelGGL(miss_kernel, dim3(sgrid), dim3(256), 0, st, jp, s, cnt);
        hipLaunchKernelGGL(shade_kernel<1>, dim3(sgrid), dim3(256), 0, st, jp, s, cnt, par, 0);
        hipLaunchKernelGGL(shade_kernel<2>, dim3(sgrid), dim3(256), 0, st, jp, s, cnt, par, 1);
        hipLaunchKernelGGL(shade_kernel<4>, dim3(sgrid), dim3(256), 0, st, jp, s, cnt, par, 2);
        hipLaunchKernelGGL(shade_kernel<7>, dim3(sgrid), dim3(256), 0, st, jp, s, cnt, par, 3);
        if (ev) PH_CHECK(hipEventRecord(ev[2], st));
        fns.shadow(st, tgrid, jp, s, cnt, par, totals.ptr + 1, count);
        if (ev) PH_CHECK(hipEventRecord(ev[3], st));
        hipLaunchKernelGGL(regen_kernel, dim3(sgrid), dim3(256), 0, st, jp, s, cnt, par);
        PH_CHECK(hipGetLastError());
    }
};

WavefrontEngine::WavefrontEngine() : impl_(std::make_unique<Impl>()) {}
WavefrontEngine::~WavefrontEngine() = default;

void WavefrontEngine::set_config(const ph_wavefront_config& c) {
    ShapeFns f;
    if (!shape_fns(c.block, c.stack_depth, f)) throw std::invalid_argument("unsupported block/stack depth for the wavefront engine");
    if (c.path_slots < 0 || c.blocks_per_cu <= 0 || c.blocks_per_cu > 64) throw std::invalid_argument("invalid wavefront configuration");
    if (c.use_graph && (c.graph_iterations < 2 || c.graph_iterations > 256 || c.graph_iterations % 2))
        throw std::invalid_argument("graph_iterations must be even and in [2, 256]");
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
    ShapeFns fns;
    if (!shape_fns(m.cfg.block, m.cfg.stack_depth, fns)) throw std::invalid_argument("unsupported wavefront launch shape");
    // Paths in flight: one per pixel up to a cap of 4M slots (~730 MB of state).
    int slots = m.cfg.path_slots > 0 ? std::min(m.cfg.path_slots, pixels) : std::min(pixels, 1 << 22);
    m.ensure_slots(slots);

    // Per-material class table and job parameters (device copies, read by every kernel).
    const int nm = job.scene.num_materials;
    std::vector<float4> mats(size_t(nm) * 5);
    PH_CHECK(hipMemcpyAsync(mats.data(), job.scene.materials, mats.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    std::vector<int> cls(static_cast<size_t>(nm));
    for (int i = 0; i < nm; i++) cls[size_t(i)] = material_class(&mats[size_t(i) * 5]);
    m.mat_class.upload(cls.data(), cls.size(), st);
    JobParams jp{};
    jp.scene = job.scene;
    jp.camera = job.camera;
    jp.width = job.width;
    jp.height = job.height;
    jp.spp = job.spp;
    jp.total_work = pixels;
    jp.frame0 = job.frame0;
    jp.seed = job.seed;
    jp.opt = job.opt;
    jp.accum = job.accum;
    jp.albedo = job.albedo;
    jp.nd = job.nd;
    jp.mat_class = m.mat_class.ptr;
    PH_CHECK(hipMemcpyAsync(m.params.ptr, &jp, sizeof(jp), hipMemcpyHostToDevice, st));
    PH_CHECK(hipMemsetAsync(m.counters.ptr, 0, C_COUNT * sizeof(int), st));
    PH_CHECK(hipMemsetAsync(m.totals.ptr, 0, 2 * sizeof(TraceTotals), st));
    PH_CHECK(hipMemsetAsync(m.ray_totals.ptr, 0, 2 * sizeof(unsigned long long), st));
    int* cnt = m.counters.ptr;
    {
        // Work items 0..slots-1 go to the slots directly; the rest are handed out by regeneration.
        int first = std::min(slots, pixels);
        PH_CHECK(hipMemcpyAsync(cnt + C_WORK, &first, sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(init_kernel, dim3((slots + 255) / 256), dim3(256), 0, st, m.params.ptr, m.state(), cnt, slots);
        PH_CHECK(hipGetLastError());
    }

    const int G = std::max(2, m.cfg.graph_iterations);
    const bool timing = true;
    if (m.events.size() < size_t(4 * G)) {
        for (size_t i = m.events.size(); i < size_t(4 * G); i++) {
            hipEvent_t e;
            PH_CHECK(hipEventCreate(&e));
            m.events.push_back(e);
        }
    }
    const bool graph = m.cfg.use_graph != 0 && !counting;
    if (graph) {
        auto key = std::make_tuple(m.cfg.block, m.cfg.stack_depth, m.cfg.blocks_per_cu, G, slots, m.cfg.sort_materials, counting, timing,
                                   int(reinterpret_cast<uintptr_t>(st) & 0x7fffffff));
        if (!m.graph_exec || key != m.graph_key) {
            m.drop_graph();
            hipGraph_t g = nullptr;
            PH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < G; it++) m.iteration(st, fns, it & 1, false, timing ? &m.events[size_t(4 * it)] : nullptr);
            PH_CHECK(hipStreamEndCapture(st, &g));
            PH_CHECK(hipGraphInstantiate(&m.graph_exec, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            m.graph_key = key;
        }
    }
    double ms_extend = 0, ms_shadow = 0;
    int64_t iterations = 0;
    for (;;) {
        if (graph) {
            PH_CHECK(hipGraphLaunch(m.graph_exec, st));
        } else {
            for (int it = 0; it < G; it++) m.iteration(st, fns, it & 1, counting, timing ? &m.events[size_t(4 * it)] : nullptr);
        }
        PH_CHECK(hipMemcpyAsync(m.host_counters, cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        for (int it = 0; it < G; it++) {
            float a = 0, b = 0;
            PH_CHECK(hipEventElapsedTime(&a, m.events[size_t(4 * it)], m.events[size_t(4 * it + 1)]));
            PH_CHECK(hipEventElapsedTime(&b, m.events[size_t(4 * it + 2)], m.events[size_t(4 * it + 3)]));
            ms_extend += a;
            ms_shadow += b;
        }
        iterations += G;
        stats.extend_launches += G;
        stats.shadow_launches += G;
        // G is even, so the next iteration reads queue 0.
        if (m.host_counters[C_EXT0] == 0) break;
        // A sample takes at most max_depth iterations and a slot runs ceil(pixels / slots) pixels.
        const int64_t bound = int64_t(job.spp) * int64_t(std::max(1, job.opt.max_depth)) * int64_t((pixels + slots - 1) / slots) + 2 * G;
        if (iterations > bound) throw std::runtime_error("wavefront loop did not drain");
    }
    TraceTotals tot[2] = {};
    if (counting) PH_CHECK(hipMemcpy(tot, m.totals.ptr, sizeof(tot), hipMemcpyDeviceToHost));
    unsigned long long rays[2] = {0, 0};
    PH_CHECK(hipMemcpy(rays, m.ray_totals.ptr, sizeof(rays), hipMemcpyDeviceToHost));
    stats.rays_extend = int64_t(rays[0]);
    stats.rays_shadow = int64_t(rays[1]) + m.host_counters[C_SHADOW];  // the last iteration's queue
    stats.ms_extend = ms_extend;
    stats.ms_shadow = ms_shadow;
    stats.iterations = iterations;
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
    auto key = std::make_tuple(static_cast<const void*>(job.scene.nodes), job.width, job.height, job.opt.max_depth);
    auto found = m.tuned.find(key);
    if (found != m.tuned.end()) {
        set_config(found->second);
        return found->second;
    }
    const ph_wavefront_config start = m.cfg;
    auto time_of = [&](const ph_wavefront_config& c) {
        set_config(c);
        ph_render_stats s{};
        render(st, job, false, s);  // warm (graph capture, allocation)
        PH_CHECK(hipMemsetAsync(job.accum, 0, size_t(job.width) * size_t(job.height) * sizeof(float4), st));
        render(st, job, false, s);
        return s.ms_total;
    };
    ph_wavefront_config best = start;
    double best_ms = time_of(best);
    // Coordinate descent: traversal shape, then persistent grid size, then paths in flight.
    for (int block : {64, 128, 256})
        for (int depth : {8, 16, 24}) {
            ph_wavefront_config c = best;
            c.block = block;
            c.stack_depth = depth;
            double t = time_of(c);
            if (t < best_ms) best_ms = t, best = c;
        }
    for (int bpc : {2, 4, 8, 16}) {
        ph_wavefront_config c = best;
        c.blocks_per_cu = bpc;
        double t = time_of(c);
        if (t < best_ms) best_ms = t, best = c;
    }
    const int pixels = job.width * job.height;
    for (int slots : {pixels / 4, pixels / 2, 0}) {
        if (slots < 0 || (slots > 0 && slots < 65536)) continue;
        ph_wavefront_config c = best;
        c.path_slots = slots;
        double t = time_of(c);
        if (t < best_ms) best_ms = t, best = c;
    }
    set_config(best);
    m.tuned[key] = best;
    PH_CHECK(hipMemsetAsync(job.accum, 0, size_t(job.width) * size_t(job.height) * sizeof(float4), st));
    PH_CHECK(hipStreamSynchronize(st));
    return best;
}

}  // namespace phk
