// Per-pixel kernels: the plain one-thread-per-pixel path tracer (render_kernel's counterpart,
// src/render/gpu_renderer.cu:60-80), the real-time G-buffer sample (sample_kernel, :83-95), the BVH
// heat map (:98-106), the display pass (:129-175), and the ray-buffer tracer of the traversal workload.
#include "dgbuf.h"
#include "internal.h"

using namespace phd;

namespace phk {

namespace {

constexpr int kBX = 16, kBY = 8, kPixelBlock = kBX * kBY;
constexpr int kMegaDepth = 32;  // LDS stack entries per lane in the per-pixel kernels
// Waves per SIMD the per-pixel kernels are compiled for. The whole integrator is live in one kernel,
// and left alone the compiler gives it registers for 2 waves per SIMD (228-246 VGPRs), too few to hide
// the traversal's memory latency. At 4 (128 VGPRs, the rest spilled), with bit-identical output:
//   * a 1600x900 one-sample G-buffer frame of the 1M-triangle field takes 3.96 ms instead of 4.96 ms
//     (3 waves: 4.08 ms, 6 waves: 4.88 ms);
//   * the 1080p render of that field runs at 479 Msamples/s instead of 329 (3 waves: 416).
constexpr int kSampleWaves = 4;

__global__ __launch_bounds__(kPixelBlock) __attribute__((amdgpu_waves_per_eu(kSampleWaves))) void render_kernel(DScene s, ph_camera_view cam, int w, int h, uint32_t frame0, int spp, uint32_t seed,
                                                             PathOpts opt, float4* accum, float4* albedo, float4* nd) {
    __shared__ int lds[kMegaDepth * kPixelBlock];
    const int tid = threadIdx.y * kBX + threadIdx.x;
    SpillArea<kMegaDepth> spill;
    LaneStack<kMegaDepth> stack{lds + tid, kPixelBlock, spill.v};
    const int x = blockIdx.x * kBX + threadIdx.x, y = blockIdx.y * kBY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    float4 a = accum[i];
    for (int k = 0; k < spp; k++) {
        PrimaryOut prim{mk3(0), 0.0f, mk3(0), -1, mk3(1), mk3(0)};
        f3 L = pixel_sample<LaneStack<kMegaDepth>, false>(s, cam, x, y, w, h, frame0 + uint32_t(k), seed, opt, k == 0 ? &prim : nullptr, stack,
                                                          nullptr);
        a.x += L.x;
        a.y += L.y;
        a.z += L.z;
        float l = lum(L);
        a.w += l * l;
        if (k == 0) {
            albedo[i] = make_float4(prim.albedo.x, prim.albedo.y, prim.albedo.z, 1.0f);
            nd[i] = make_float4(prim.normal.x, prim.normal.y, prim.normal.z, prim.depth);
        }
    }
    accum[i] = a;
}

__global__ __launch_bounds__(kPixelBlock) __attribute__((amdgpu_waves_per_eu(kSampleWaves))) void sample_kernel(DScene s, ph_camera_view cam, int w, int h, uint32_t frame, uint32_t seed, PathOpts opt,
                                                             float4* color, float4* albedo, float4* nd, float4* position, float4* emission) {
    __shared__ int lds[kMegaDepth * kPixelBlock];
    const int tid = threadIdx.y * kBX + threadIdx.x;
    SpillArea<kMegaDepth> spill;
    LaneStack<kMegaDepth> stack{lds + tid, kPixelBlock, spill.v};
    const int x = blockIdx.x * kBX + threadIdx.x, y = blockIdx.y * kBY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    PrimaryOut p{mk3(0), 0.0f, mk3(0), -1, mk3(1), mk3(0)};
    f3 L = pixel_sample<LaneStack<kMegaDepth>, false>(s, cam, x, y, w, h, frame, seed, opt, &p, stack, nullptr);
    color[i] = make_float4(L.x, L.y, L.z, 1.0f);
    albedo[i] = make_float4(p.albedo.x, p.albedo.y, p.albedo.z, 1.0f);
    nd[i] = make_float4(p.normal.x, p.normal.y, p.normal.z, p.depth);
    position[i] = make_float4(p.position.x, p.position.y, p.position.z, float(p.material));
    emission[i] = make_float4(p.emission.x, p.emission.y, p.emission.z, 1.0f);
}

// sample_kernel's sample, written as the compact G-buffer.
__global__ __launch_bounds__(kPixelBlock) __attribute__((amdgpu_waves_per_eu(kSampleWaves))) void sample_compact_kernel(DScene s, ph_camera_view cam, int w, int h, uint32_t frame, uint32_t seed,
                                                                     PathOpts opt, float* color, float* albedo, float* emission, float* depth,
                                                                     uint32_t* normal, int* material) {
    __shared__ int lds[kMegaDepth * kPixelBlock];
    const int tid = threadIdx.y * kBX + threadIdx.x;
    SpillArea<kMegaDepth> spill;
    LaneStack<kMegaDepth> stack{lds + tid, kPixelBlock, spill.v};
    const int x = blockIdx.x * kBX + threadIdx.x, y = blockIdx.y * kBY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    PrimaryOut p{mk3(0), 0.0f, mk3(0), -1, mk3(1), mk3(0)};
    f3 L = pixel_sample<LaneStack<kMegaDepth>, false>(s, cam, x, y, w, h, frame, seed, opt, &p, stack, nullptr);
    color[3 * i] = L.x;
    color[3 * i + 1] = L.y;
    color[3 * i + 2] = L.z;
    albedo[3 * i] = p.albedo.x;
    albedo[3 * i + 1] = p.albedo.y;
    albedo[3 * i + 2] = p.albedo.z;
    emission[3 * i] = p.emission.x;
    emission[3 * i + 1] = p.emission.y;
    emission[3 * i + 2] = p.emission.z;
    depth[i] = p.depth;
    normal[i] = oct_encode(p.normal);
    material[i] = p.material;
}

__global__ void cdf_guide_kernel(const float* cdf, int rows, int n, size_t stride, int m, int* guide) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * (m + 1)) return;
    const int r = int(t / (m + 1)), k = int(t - (long long)r * (m + 1));
    guide[t] = cdf_search(cdf + size_t(r) * stride, n, float(k) / float(m));
}

__global__ __launch_bounds__(kPixelBlock) void heatmap_kernel(DScene s, ph_camera_view cam, int w, int h, float* out) {
    __shared__ int lds[kMegaDepth * kPixelBlock];
    const int tid = threadIdx.y * kBX + threadIdx.x;
    SpillArea<kMegaDepth> spill;
    LaneStack<kMegaDepth> stack{lds + tid, kPixelBlock, spill.v};
    const int x = blockIdx.x * kBX + threadIdx.x, y = blockIdx.y * kBY + threadIdx.y;
    if (x >= w || y >= h) return;
    f3 o, d;
    eye_ray(cam, float(x) + 0.5f, float(y) + 0.5f, w, h, mk2(0.5f, 0.5f), o, d);
    HitRec hit{kInf, -1, 0.0f, 0.0f};
    TraceCount c{0, 0, 0};
    traverse<false, true>(s, o, d, kInf, hit, stack, &c);
    out[y * w + x] = float(c.nodes);
}

PH_DEV f3 heat_color(float t) {
    t = sat(t);
    const f3 c0 = mk3(0.05f, 0.05f, 0.35f), c1 = mk3(0.0f, 0.6f, 0.4f), c2 = mk3(0.95f, 0.85f, 0.1f), c3 = mk3(0.85f, 0.1f, 0.05f);
    if (t < 1.0f / 3) return lerp3(c0, c1, t * 3);
    if (t < 2.0f / 3) return lerp3(c1, c2, t * 3 - 1);
    return lerp3(c2, c3, t * 3 - 2);
}

__global__ void display_kernel(DisplayParams p, const float4* accum, const float4* denoised, const float4* albedo, const float4* nd,
                               const float4* variance, const float* heat, uint32_t* out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.w || y >= p.h) return;
    const int i = y * p.w + x;
    f3 c;
    switch (p.mode) {
        case PH_DISPLAY_ALBEDO: c = xyz(albedo[i]); break;
        case PH_DISPLAY_NORMAL: {
            float4 v = nd[i];
            c = v.w > 0 ? xyz(v) * 0.5f + mk3(0.5f) : mk3(0);
            break;
        }
        case PH_DISPLAY_DEPTH: {
            float d = nd[i].w;
            c = d > 0 ? mk3(1.0f - sat(d / p.max_depth)) : mk3(0);
            break;
        }
        case PH_DISPLAY_VARIANCE: {
            float v;
            if (p.denoised) {
                v = variance[i].w;
            } else {
                float4 a = accum[i];
                float mean = lum(xyz(a)) * p.inv_samples;
                v = p.samples > 1 ? fmaxf(0.0f, a.w * p.inv_samples - mean * mean) / (p.samples - 1) : 0.0f;
            }
            float t = __builtin_sqrtf(v);
            c = mk3(t / (1.0f + t));
            break;
        }
        case PH_DISPLAY_HEATMAP: c = heat_color(heat[i] * p.heat_scale); break;
        default:
            c = p.denoised ? xyz(denoised[i]) : xyz(accum[i]) * p.inv_samples;
            c = tonemap(c, p.exposure, p.tonemap);
            break;
    }
    // Debug views are display values already; packing re-encodes them, so decode first.
    if (p.mode != PH_DISPLAY_COLOR && p.mode != PH_DISPLAY_ALBEDO) c = mk3(srgb_decode(c.x), srgb_decode(c.y), srgb_decode(c.z));
    out[i] = pack8(c);
}

// ---- ray-buffer traversal --------------------------------------------------------------------------
PH_DEV void trace_one(const DScene& s, const ph_ray* rays, int i, bool anyhit_flag, ph_hit* hits, uint32_t* counts, auto& stack, auto count_tag,
                      auto any_tag) {
    constexpr bool Count = decltype(count_tag)::value;
    constexpr bool AnyHit = decltype(any_tag)::value;
    const float* r = reinterpret_cast<const float*>(rays + i);
    f3 o = mk3(r[0], r[1], r[2]), d = mk3(r[3], r[4], r[5]);
    float tmax = r[6];
    HitRec hit{kInf, -1, 0.0f, 0.0f};
    TraceCount c{0, 0, 0};
    traverse<AnyHit, Count>(s, o, d, tmax, hit, stack, &c);
    reinterpret_cast<float4*>(hits)[i] = make_float4(hit.t, __int_as_float(hit.tri), hit.u, hit.v);
    if (Count) {
        counts[3 * i] = c.nodes;
        counts[3 * i + 1] = c.tris;
        counts[3 * i + 2] = c.alpha;
    }
    (void)anyhit_flag;
}

template <bool V>
struct Tag {
    static constexpr bool value = V;
};

template <int Block, int Depth, bool AnyHit, bool Count, bool Persistent>
__global__ __launch_bounds__(Block) void trace_kernel(DScene s, const ph_ray* rays, int n, ph_hit* hits, uint32_t* counts, int lds_nodes) {
    __shared__ int lds[Depth * Block];
    extern __shared__ float4 top_nodes[];
    SpillArea<Depth> spill;
    LaneStack<Depth> stack{lds + threadIdx.x, Block, spill.v};
    if (Persistent) {
        // A fixed grid; each wave refills its lanes from its share of the rays (trace_stream), with the
        // top of the tree staged in LDS.
        TraceScene ts = trace_scene(s);
        ts.top = top_nodes;
        ts.top_n = min(lds_nodes, s.top_nodes);
        stage_top_nodes(top_nodes, s.nodes, ts.top_n);
        TraceCount c{0, 0, 0};
        int idx = 0;  // this lane's ray
        trace_stream<AnyHit, false>(
            n, 1, 16, ts, stack, c,
            [&](int i, TraceScene&, Walker& w, HitRec& hit) {
                idx = i;
                const float* r = reinterpret_cast<const float*>(rays + i);
                w = walker(mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6]);
                hit = HitRec{kInf, -1, 0.0f, 0.0f};
            },
            [&](const TraceScene&, const Walker&, const HitRec& hit) {
                reinterpret_cast<float4*>(hits)[idx] = make_float4(hit.t, __int_as_float(hit.tri), hit.u, hit.v);
            });
    } else {
        const int i = blockIdx.x * Block + threadIdx.x;
        if (i < n) trace_one(s, rays, i, AnyHit, hits, counts, stack, Tag<Count>{}, Tag<AnyHit>{});
    }
}

template <int Block, int Depth>
void trace_shape(hipStream_t st, const DScene& s, const ph_ray* rays, int n, bool anyhit, ph_hit* hits, uint32_t* counts, bool persistent,
                 int lds_nodes, int grid_blocks) {
    const int blocks = persistent ? grid_blocks : (n + Block - 1) / Block;
    if (blocks <= 0) return;
    const int top = persistent && !counts ? lds_nodes : 0;
#define PH_TRACE_LAUNCH(A, C, P) \
    hipLaunchKernelGGL((trace_kernel<Block, Depth, A, C, P>), dim3(blocks), dim3(Block), size_t(top) * 64, st, s, rays, n, hits, counts, top)
    if (counts) {
        if (anyhit)
            PH_TRACE_LAUNCH(true, true, false);
        else
            PH_TRACE_LAUNCH(false, true, false);
    } else if (persistent) {
        if (anyhit)
            PH_TRACE_LAUNCH(true, false, true);
        else
            PH_TRACE_LAUNCH(false, false, true);
    } else {
        if (anyhit)
            PH_TRACE_LAUNCH(true, false, false);
        else
            PH_TRACE_LAUNCH(false, false, false);
    }
#undef PH_TRACE_LAUNCH
}

}  // namespace

void launch_render(hipStream_t st, const DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame0, int spp, uint32_t seed, const PathOpts& opt,
                   float4* accum, float4* albedo, float4* nd) {
    dim3 block(kBX, kBY), grid((w + kBX - 1) / kBX, (h + kBY - 1) / kBY);
    hipLaunchKernelGGL(render_kernel, grid, block, 0, st, s, cam, w, h, frame0, spp, seed, opt, accum, albedo, nd);
    PH_CHECK(hipGetLastError());
}

void launch_sample(hipStream_t st, const DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame, uint32_t seed, const PathOpts& opt,
                   float4* color, float4* albedo, float4* nd, float4* position, float4* emission) {
    dim3 block(kBX, kBY), grid((w + kBX - 1) / kBX, (h + kBY - 1) / kBY);
    hipLaunchKernelGGL(sample_kernel, grid, block, 0, st, s, cam, w, h, frame, seed, opt, color, albedo, nd, position, emission);
    PH_CHECK(hipGetLastError());
}

void build_cdf_guides(hipStream_t st, const float* cdf, int rows, int n, size_t stride, int m, int* guide) {
    const long long total = (long long)rows * (m + 1);
    hipLaunchKernelGGL(cdf_guide_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, cdf, rows, n, stride, m, guide);
    PH_CHECK(hipGetLastError());
}

void launch_sample_compact(hipStream_t st, const DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame, uint32_t seed, const PathOpts& opt,
                           const ph_compact_gbuffer& out) {
    dim3 block(kBX, kBY), grid((w + kBX - 1) / kBX, (h + kBY - 1) / kBY);
    hipLaunchKernelGGL(sample_compact_kernel, grid, block, 0, st, s, cam, w, h, frame, seed, opt, static_cast<float*>(out.color_dev),
                       static_cast<float*>(out.albedo_dev), static_cast<float*>(out.emission_dev), static_cast<float*>(out.depth_dev),
                       static_cast<uint32_t*>(out.normal_dev), static_cast<int*>(out.material_dev));
    PH_CHECK(hipGetLastError());
}

void launch_heatmap(hipStream_t st, const DScene& s, const ph_camera_view& cam, int w, int h, float* out) {
    dim3 block(kBX, kBY), grid((w + kBX - 1) / kBX, (h + kBY - 1) / kBY);
    hipLaunchKernelGGL(heatmap_kernel, grid, block, 0, st, s, cam, w, h, out);
    PH_CHECK(hipGetLastError());
}

void launch_display(hipStream_t st, const DisplayParams& p, const float4* accum, const float4* denoised, const float4* albedo, const float4* nd,
                    const float4* variance, const float* heat, uint32_t* out) {
    dim3 block(kBX, kBY), grid((p.w + kBX - 1) / kBX, (p.h + kBY - 1) / kBY);
    hipLaunchKernelGGL(display_kernel, grid, block, 0, st, p, accum, denoised, albedo, nd, variance, heat, out);
    PH_CHECK(hipGetLastError());
}

bool launch_trace(hipStream_t st, const DScene& s, const ph_ray* rays, int n, bool anyhit, ph_hit* hits, uint32_t* counts, int block, int depth,
                  bool persistent, int lds_nodes, int grid_blocks) {
#define PH_SHAPE(B, D)                                                                           \
    if (block == B && depth == D) {                                                              \
        trace_shape<B, D>(st, s, rays, n, anyhit, hits, counts, persistent, lds_nodes, grid_blocks); \
        PH_CHECK(hipGetLastError());                                                        \
        return true;                                                                        \
    }
    PH_SHAPE(64, 8)
    PH_SHAPE(64, 16)
    PH_SHAPE(64, 32)
    PH_SHAPE(128, 8)
    PH_SHAPE(128, 16)
    PH_SHAPE(128, 24)
    PH_SHAPE(128, 32)
    PH_SHAPE(256, 8)
    PH_SHAPE(256, 16)
    PH_SHAPE(256, 32)
    PH_SHAPE(128, 64)
#undef PH_SHAPE
    return false;
}

}  // namespace phk
