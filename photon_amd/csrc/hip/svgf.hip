// SVGF denoiser on the device (Schied et al. 2017), following the reference's arithmetic and pass order
// (src/denoise/svgf.cu:19-255): temporal accumulation -> variance -> a-trous pass 0 -> history ->
// a-trous passes 1..4 -> remodulation.
//
// MI355X layout: the reference's planes are float4 per pixel (16-byte loads). The variance estimate and
// the a-trous passes stage their footprint through LDS: a 16x16 output tile reads a (16 + 2r)^2 halo of
// the G-buffer planes it needs once per block instead of once per tap. The larger a-trous steps (4, 8,
// 16) gather directly through the caches: a wave's loads for one tap cover 4 rows of 16 neighbouring
// pixels, so they stay coalesced, and staging by residue class mod the step (so that a block's taps
// share one lattice halo) measured slower, its staging loads being strided.
//
// Two input formats share the kernels through an accessor policy: the reference's five float4 planes
// (RefIn), and the real-time workload's compact G-buffer (CompactIn: 12-byte colour, albedo and emission,
// fp32 depth, fp16 octahedral normal, int32 material; positions rebuilt from depth along each pixel's
// centre ray). A compact frame is denoised exactly as its decoded float4 planes would be; only the
// loads, and the format of the stored previous-frame geometry, differ.
#include <mutex>
#include <string>

#include "internal.h"
#include "dgbuf.h"
#include "svgf_state.h"

using namespace phd;

namespace phk {

namespace {

constexpr int TX = 16, TY = 16;

// Edge-stopping weights. The denoiser's tolerance is 1e-3 relative, so they use the hardware's
// single-precision exp2/log2 (a few ulp) instead of correctly rounded exp/pow; a power-of-two normal
// exponent (the default 128) is evaluated by exact repeated squaring.
PH_DEV float weight_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

PH_DEV float weight_pow(float x, float e) {
    if (x <= 0.0f) return e > 0.0f ? 0.0f : 1.0f;
    const int ie = int(e);
    if (float(ie) == e && ie > 0 && ie <= 1024 && (ie & (ie - 1)) == 0) {
        float r = x;
        for (int k = ie; k > 1; k >>= 1) r *= r;
        return r;
    }
    return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}

PH_DEV f3 demod(f3 color, f3 emission, f3 albedo) {
    f3 illum = vmax3(color - emission, mk3(0));
    return mk3(illum.x / fmaxf(albedo.x, 1e-3f), illum.y / fmaxf(albedo.y, 1e-3f), illum.z / fmaxf(albedo.z, 1e-3f));
}

PH_DEV bool reproject(const ph_camera_view& c, f3 p, int w, int h, float& px, float& py, float& dist) {
    f3 d = p - ld3(c.position);
    float z = dot(d, ld3(c.forward));
    if (z <= 1e-4f) return false;
    float sx = dot(d, ld3(c.right)) / (z * c.tan_half_fov * c.aspect);
    float sy = dot(d, ld3(c.up)) / (z * c.tan_half_fov);
    px = (sx + 1.0f) * 0.5f * float(w) - 0.5f;
    py = (1.0f - sy) * 0.5f * float(h) - 0.5f;
    dist = __builtin_sqrtf(dot(d, d));
    return true;
}

// The reference's five float4 planes (SvgfInput, src/denoise/svgf.h:24-30) and its previous-frame
// normal+depth and position+material planes.
struct RefIn {
    const float4 *color, *albedo, *nd, *pos, *emission;
    PH_DEV f3 color3(int i) const { return xyz(color[i]); }
    PH_DEV f3 albedo3(int i) const { return xyz(albedo[i]); }
    PH_DEV f3 emission3(int i) const { return xyz(emission[i]); }
    PH_DEV float4 ndz(int i) const { return nd[i]; }
    PH_DEV float mat(int i) const { return pos[i].w; }
    PH_DEV f3 position(int i, int, int) const { return xyz(pos[i]); }
    PH_DEV float4 raw_color(int i) const { return color[i]; }
};

struct RefHist {
    float4 *nd, *pos;
    PH_DEV float4 ndz(int j) const { return nd[j]; }
    PH_DEV float mat(int j) const { return pos[j].w; }
    PH_DEV void store(const RefIn& in, int i) const {
        nd[i] = in.nd[i];
        pos[i] = in.pos[i];
    }
};

// The compact G-buffer (ph_compact_gbuffer) and its previous-frame depth, normal code and material.
struct CompactIn {
    const float *color, *albedo, *emission, *depth;
    const uint32_t* normal;
    const int* material;
    ph_camera_view cam;  // the view the G-buffer was traced from (positions are rebuilt from depth)
    int w, h;
    PH_DEV f3 color3(int i) const { return ld3(color + 3 * i); }
    PH_DEV f3 albedo3(int i) const { return ld3(albedo + 3 * i); }
    PH_DEV f3 emission3(int i) const { return ld3(emission + 3 * i); }
    PH_DEV float4 ndz(int i) const {
        const f3 n = oct_decode(normal[i]);
        return make_float4(n.x, n.y, n.z, depth[i]);
    }
    PH_DEV float mat(int i) const { return float(material[i]); }
    PH_DEV f3 position(int i, int x, int y) const { return centre_position(cam, x, y, w, h, depth[i]); }
    PH_DEV float4 raw_color(int i) const {
        const f3 c = color3(i);
        return make_float4(c.x, c.y, c.z, 1.0f);
    }
};

struct CompactHist {
    float* depth;
    uint32_t* normal;
    int* material;
    PH_DEV float4 ndz(int j) const {
        const f3 n = oct_decode(normal[j]);
        return make_float4(n.x, n.y, n.z, depth[j]);
    }
    PH_DEV float mat(int j) const { return float(material[j]); }
    PH_DEV void store(const CompactIn& in, int i) const {
        depth[i] = in.depth[i];
        normal[i] = in.normal[i];
        material[i] = in.material[i];
    }
};

template <class In, class Hist>
__global__ __launch_bounds__(TX* TY) void temporal_kernel(In in, ph_camera_view prev_cam, int w, int h, int has_history, float alpha_c,
                                                           float alpha_m, const float4* prev_illum, const float4* prev_moments, Hist prev,
                                                           float4* out_illum, float4* out_moments) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    const float mat = in.mat(i);
    const f3 cur = demod(in.color3(i), in.emission3(i), in.albedo3(i));
    const float l = lum(cur);
    f3 hist_c = mk3(0);
    float hm0 = 0, hm1 = 0, hist_len = 0, wsum = 0;
    if (has_history && mat >= 0) {
        float px, py, dist;
        if (reproject(prev_cam, in.position(i, x, y), w, h, px, py, dist)) {
            const f3 n = xyz(in.ndz(i));
            const int x0 = int(floorf(px)), y0 = int(floorf(py));
            const float fx = px - float(x0), fy = py - float(y0);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int xx = x0 + (k & 1), yy = y0 + (k >> 1);
                if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                const int j = yy * w + xx;
                if (prev.mat(j) != mat) continue;
                const float4 pn = prev.ndz(j);
                if (fabsf(pn.w - dist) > 0.05f * dist + 0.01f) continue;
                if (dot(xyz(pn), n) < 0.9f) continue;
                const float bw = ((k & 1) ? fx : 1 - fx) * ((k >> 1) ? fy : 1 - fy);
                const float4 pi = prev_illum[j];
                const float4 pm = prev_moments[j];
                hist_c = hist_c + xyz(pi) * bw;
                hm0 = hm0 + pm.x * bw;
                hm1 = hm1 + pm.y * bw;
                hist_len += pi.w * bw;
                wsum += bw;
            }
        }
    }
    f3 c;
    float m0, m1, len;
    if (wsum > 1e-3f) {
        hist_c = hist_c / wsum;
        const float inv = 1.0f / wsum;
        hm0 = hm0 * inv;
        hm1 = hm1 * inv;
        len = fminf(hist_len / wsum + 1.0f, 64.0f);
        const float ac = fmaxf(alpha_c, 1.0f / len), am = fmaxf(alpha_m, 1.0f / len);
        c = lerp3(hist_c, cur, ac);
        m0 = hm0 + (l - hm0) * am;
        m1 = hm1 + (l * l - hm1) * am;
    } else {
        len = 1.0f;
        c = cur;
        m0 = l;
        m1 = l * l;
    }
    const float variance = fmaxf(0.0f, m1 - m0 * m0);
    out_illum[i] = make_float4(c.x, c.y, c.z, len);
    out_moments[i] = make_float4(m0, m1, variance, 0.0f);
}

// 7x7 bilateral variance estimate while the history is short (svgf.cu:96-132). The block stages the
// 22x22 halo of normal+depth and position.w (material) in LDS; moments and illumination are read only
// for accepted taps of short-history pixels.
template <class In>
__global__ __launch_bounds__(TX* TY) void variance_kernel(In in, int w, int h, const float4* illum, const float4* moments, float phi_normal,
                                                           float4* out) {
    constexpr int R = 3, SX = TX + 2 * R, SY = TY + 2 * R;
    __shared__ float4 s_nd[SY][SX];
    __shared__ float s_mat[SY][SX];
    const int bx = blockIdx.x * TX - R, by = blockIdx.y * TY - R;
    for (int t = threadIdx.y * TX + threadIdx.x; t < SX * SY; t += TX * TY) {
        const int sy = t / SX, sx = t % SX;
        const int gx = bx + sx, gy = by + sy;
        if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
            const int j = gy * w + gx;
            s_nd[sy][sx] = in.ndz(j);
            s_mat[sy][sx] = in.mat(j);
        }
    }
    __syncthreads();
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    float4 c = illum[i];
    const float len = c.w;
    float var = moments[i].z;
    const float mat = s_mat[threadIdx.y + R][threadIdx.x + R];
    if (len < 4.0f && mat >= 0) {
        const float4 ndc = s_nd[threadIdx.y + R][threadIdx.x + R];
        const f3 n = xyz(ndc);
        const float z = ndc.w;
        float wsum = 0, m1 = 0, m2 = 0;
        f3 acc = mk3(0);
        for (int dy = -R; dy <= R; dy++)
            for (int dx = -R; dx <= R; dx++) {
                const int xx = x + dx, yy = y + dy;
                if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                const int ly = threadIdx.y + R + dy, lx = threadIdx.x + R + dx;
                if (s_mat[ly][lx] != mat) continue;
                const float4 nd = s_nd[ly][lx];
                const float wn = weight_pow(fmaxf(0.0f, dot(n, xyz(nd))), phi_normal);
                const float wz = weight_exp(-fabsf(nd.w - z) / (0.05f * z + 1e-3f));
                const float wt = wn * wz;
                const int j = yy * w + xx;
                const float4 m = moments[j];
                m1 += m.x * wt;
                m2 += m.y * wt;
                acc = acc + xyz(illum[j]) * wt;
                wsum += wt;
            }
        if (wsum > 0) {
            m1 /= wsum;
            m2 /= wsum;
            var = fmaxf(0.0f, m2 - m1 * m1) * (4.0f / len);
            f3 a = acc / wsum;
            c = make_float4(a.x, a.y, a.z, len);
        }
    }
    out[i] = make_float4(c.x, c.y, c.z, var);
}

// One a-trous pass (svgf.cu:134-186). For steps 1 and 2 the block stages src, normal+depth and the
// material id of its (16 + 4*step + 2)^2 footprint in LDS; larger steps read through the caches.
template <int Step, class In>
__global__ __launch_bounds__(TX* TY) void atrous_kernel(In in, int w, int h, int step_rt, float phi_color, float phi_normal, float phi_depth,
                                                         const float4* src, float4* dst) {
    constexpr bool Tiled = Step > 0;
    constexpr int R = Tiled ? 2 * Step + 1 : 1;  // taps reach 2*step; the 3x3 prefilter and gradient need 1
    constexpr int SX = Tiled ? TX + 2 * R : 1, SY = Tiled ? TY + 2 * R : 1;
    __shared__ float4 s_src[SY][SX];
    __shared__ float4 s_nd[SY][SX];
    __shared__ float s_mat[SY][SX];
    const int step = Tiled ? Step : step_rt;
    if (Tiled) {
        const int bx = blockIdx.x * TX - R, by = blockIdx.y * TY - R;
        for (int t = threadIdx.y * TX + threadIdx.x; t < SX * SY; t += TX * TY) {
            const int sy = t / SX, sx = t % SX;
            // Clamp to the image: the prefilter and the gradient clamp their coordinates, and the
            // a-trous taps skip out-of-image positions before reading.
            const int gx = min(max(blockIdx.x * TX - R + sx, 0), w - 1), gy = min(max(blockIdx.y * TY - R + sy, 0), h - 1);
            const int j = gy * w + gx;
            s_src[sy][sx] = src[j];
            s_nd[sy][sx] = in.ndz(j);
            s_mat[sy][sx] = in.mat(j);
            (void)bx;
            (void)by;
        }
        __syncthreads();
    }
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    const int lx = threadIdx.x + R, ly = threadIdx.y + R;
    // Loads through LDS (tiled) or global memory; (gx, gy) are image coordinates already in range.
    auto SRC = [&](int gx, int gy) { return Tiled ? s_src[ly + gy - y][lx + gx - x] : src[gy * w + gx]; };
    auto ND = [&](int gx, int gy) { return Tiled ? s_nd[ly + gy - y][lx + gx - x] : in.ndz(gy * w + gx); };
    auto MAT = [&](int gx, int gy) { return Tiled ? s_mat[ly + gy - y][lx + gx - x] : in.mat(gy * w + gx); };
    const float4 center = SRC(x, y);
    const float mat = MAT(x, y);
    if (mat < 0) {
        dst[i] = center;
        return;
    }
    float var = 0, vw = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const int xx = min(max(x + dx, 0), w - 1), yy = min(max(y + dy, 0), h - 1);
            const float k = (dx == 0 && dy == 0) ? 0.25f : ((dx == 0 || dy == 0) ? 0.125f : 0.0625f);
            var += SRC(xx, yy).w * k;
            vw += k;
        }
    var /= vw;
    const float4 ndc = ND(x, y);
    const f3 n = xyz(ndc);
    const float z = ndc.w;
    const float zx = ND(min(x + 1, w - 1), y).w - ND(max(x - 1, 0), y).w;
    const float zy = ND(x, min(y + 1, h - 1)).w - ND(x, max(y - 1, 0)).w;
    const float grad = 0.5f * fmaxf(fabsf(zx), fabsf(zy));
    const float lum_c = lum(xyz(center));
    const float sigma_l = phi_color * __builtin_sqrtf(fmaxf(var, 0.0f)) + 1e-4f;
    const float kernel[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    f3 sum = mk3(0);
    float wsum = 0, vsum = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++)
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            const int xx = x + dx * step, yy = y + dy * step;
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            if (MAT(xx, yy) != mat) continue;
            const float4 sv = SRC(xx, yy);
            const float4 nd = ND(xx, yy);
            const float k = kernel[dx < 0 ? -dx : dx] * kernel[dy < 0 ? -dy : dy];
            const float wl = weight_exp(-fabsf(lum(xyz(sv)) - lum_c) / sigma_l);
            const float wn = weight_pow(fmaxf(0.0f, dot(n, xyz(nd))), phi_normal);
            const float dist = __builtin_sqrtf(float(dx * dx + dy * dy)) * float(step);
            const float wz = weight_exp(-fabsf(nd.w - z) / (phi_depth * grad * dist + 1e-3f * z + 1e-4f));
            const float wt = k * wl * wn * wz;
            sum = sum + xyz(sv) * wt;
            vsum += wt * wt * sv.w;
            wsum += wt;
        }
    if (wsum > 0) {
        const f3 a = sum / wsum;
        dst[i] = make_float4(a.x, a.y, a.z, vsum / (wsum * wsum));
    } else {
        dst[i] = center;
    }
}

template <class In, class Hist>
__global__ void history_kernel(In in, int n, const float4* filtered, const float4* illum, const float4* moments, float4* prev_illum,
                               float4* prev_moments, Hist prev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 f = filtered[i];
    prev_illum[i] = make_float4(f.x, f.y, f.z, illum[i].w);
    prev_moments[i] = moments[i];
    prev.store(in, i);
}

template <class In>
__global__ void remodulate_kernel(In in, int n, const float4* filtered, float4* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (in.mat(i) < 0) {
        out[i] = in.raw_color(i);
        return;
    }
    const f3 a = in.albedo3(i);
    const f3 c = xyz(filtered[i]) * vmax3(a, mk3(1e-3f)) + in.emission3(i);
    out[i] = make_float4(c.x, c.y, c.z, 1.0f);
}

// The planes one frame works on: the previous frame's illumination and moments (history, read when
// `has_history`, rewritten after a-trous pass 0) and four frame-local planes.
struct PassPlanes {
    int w, h;
    bool has_history;
    float4 *prev_illum, *prev_moments;
    float4 *illum, *moments, *filter0, *filter1;
};

// The nine launches of one frame (svgf.cu:238-255) for either input format.
template <class In, class Hist>
void run_passes(const PassPlanes& pl, hipStream_t st, const In& in, const Hist& prev, const ph_camera_view& prev_cam, const ph_svgf_settings& s,
                float4* output) {
    const int w = pl.w, h = pl.h, n = w * h;
    float4 *prev_illum = pl.prev_illum, *prev_moments = pl.prev_moments;
    float4 *illum = pl.illum, *moments = pl.moments, *filter[2] = {pl.filter0, pl.filter1};
    const dim3 tile(TX, TY), grid((w + TX - 1) / TX, (h + TY - 1) / TY);
    const int lin = 256, lgrid = (n + lin - 1) / lin;
    hipLaunchKernelGGL((temporal_kernel<In, Hist>), grid, tile, 0, st, in, prev_cam, w, h, int(pl.has_history), s.alpha_color, s.alpha_moments,
                       prev_illum, prev_moments, prev, illum, moments);
    hipLaunchKernelGGL((variance_kernel<In>), grid, tile, 0, st, in, w, h, illum, moments, s.phi_normal, filter[0]);
    int src = 0;
    for (int it = 0; it < s.atrous_iterations; it++) {
        const int step = 1 << it;
        if (step == 1)
            hipLaunchKernelGGL((atrous_kernel<1, In>), grid, tile, 0, st, in, w, h, step, s.phi_color, s.phi_normal, s.phi_depth, filter[src],
                               filter[1 - src]);
        else if (step == 2)
            hipLaunchKernelGGL((atrous_kernel<2, In>), grid, tile, 0, st, in, w, h, step, s.phi_color, s.phi_normal, s.phi_depth, filter[src],
                               filter[1 - src]);
        else
            hipLaunchKernelGGL((atrous_kernel<0, In>), grid, tile, 0, st, in, w, h, step, s.phi_color, s.phi_normal, s.phi_depth, filter[src],
                               filter[1 - src]);
        src = 1 - src;
        if (it == 0)
            hipLaunchKernelGGL((history_kernel<In, Hist>), dim3(lgrid), dim3(lin), 0, st, in, n, filter[src], illum, moments, prev_illum, prev_moments,
                               prev);
    }
    if (s.atrous_iterations == 0)
        hipLaunchKernelGGL((history_kernel<In, Hist>), dim3(lgrid), dim3(lin), 0, st, in, n, filter[src], illum, moments, prev_illum, prev_moments,
                           prev);
    hipLaunchKernelGGL((remodulate_kernel<In>), dim3(lgrid), dim3(lin), 0, st, in, n, filter[src], output);
    PH_CHECK(hipGetLastError());
}

PassPlanes planes_of(const SvgfState& f) {
    return PassPlanes{f.w, f.h, f.has_history, f.plane[0].ptr, f.plane[1].ptr, f.plane[4].ptr, f.plane[5].ptr, f.plane[6].ptr, f.plane[7].ptr};
}

// Frame-local planes of the allocation-free entry point: illumination, moments and the two a-trous
// planes of up to kStaticSvgfPixels pixels, in one file-scope device array that every GPU holds from
// the library's load until the process ends. Host code never names the array: it uses the address
// that hipGetSymbolAddress returns on each device.
__device__ __align__(256) float4 scratch[4 * size_t(PH_SVGF_STATIC_MAX_PIXELS)];

constexpr int kMaxDevices = 64;
float4* g_scratch[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_status[kMaxDevices] = {};

// The array's address on the current device, looked up once per device with that device current.
float4* static_scratch() {
    int dev = 0;
    PH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the static scratch table");
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = static_cast<float4*>(p);
    });
    if (g_scratch_status[dev] != hipSuccess || !g_scratch[dev])
        throw std::runtime_error(std::string("hipGetSymbolAddress of the SVGF scratch array failed: ") + hipGetErrorString(g_scratch_status[dev]));
    return g_scratch[dev];
}

}  // namespace

void svgf_resize(SvgfState& f, int width, int height) {
    if (width == f.w && height == f.h && f.plane[0].ptr) return;
    f.w = width;
    f.h = height;
    const size_t n = size_t(width) * size_t(height);
    for (auto& p : f.plane) {
        p.allocate(n);
        PH_CHECK(hipMemset(p.ptr, 0, n * sizeof(float4)));
    }
    f.has_history = false;
}

void svgf_run(SvgfState& f, hipStream_t st, const ph_svgf_input& input, const ph_camera_view& prev_cam, const ph_svgf_settings& s, float4* output) {
    if (f.w * f.h == 0) return;
    if (f.history_compact) f.has_history = false;  // the stored previous frame is in the other format
    f.history_compact = false;
    const RefIn in{reinterpret_cast<const float4*>(input.color_dev), reinterpret_cast<const float4*>(input.albedo_dev),
                   reinterpret_cast<const float4*>(input.normal_depth_dev), reinterpret_cast<const float4*>(input.position_dev),
                   reinterpret_cast<const float4*>(input.emission_dev)};
    run_passes(planes_of(f), st, in, RefHist{f.plane[2].ptr, f.plane[3].ptr}, prev_cam, s, output);
    f.has_history = true;
}

void svgf_run_static(hipStream_t st, int width, int height, const ph_svgf_input& input, const ph_camera_view& prev_cam, const ph_svgf_settings& s,
                     float4* history, bool has_history, float4* output) {
    const size_t n = size_t(width) * size_t(height);
    if (n == 0) return;
    if (n > size_t(PH_SVGF_STATIC_MAX_PIXELS)) throw std::runtime_error("frame larger than the static SVGF scratch");
    float4* base = static_scratch();
    // History: the caller's four planes (illumination + length, moments, normal+depth, position+material).
    const PassPlanes pl{width, height, has_history, history, history + n, base, base + n, base + 2 * n, base + 3 * n};
    const RefIn in{reinterpret_cast<const float4*>(input.color_dev), reinterpret_cast<const float4*>(input.albedo_dev),
                   reinterpret_cast<const float4*>(input.normal_depth_dev), reinterpret_cast<const float4*>(input.position_dev),
                   reinterpret_cast<const float4*>(input.emission_dev)};
    run_passes(pl, st, in, RefHist{history + 2 * n, history + 3 * n}, prev_cam, s, output);
}

void svgf_run_compact(SvgfState& f, hipStream_t st, const ph_compact_gbuffer& g, const ph_camera_view& cam, const ph_camera_view& prev_cam,
                      const ph_svgf_settings& s, float4* output) {
    const int n = f.w * f.h;
    if (n == 0) return;
    if (!f.prev_depth.ptr || f.prev_depth.count < size_t(n)) {
        f.prev_depth.allocate(size_t(n));
        f.prev_normal.allocate(size_t(n));
        f.prev_material.allocate(size_t(n));
        f.has_history = false;
    }
    if (!f.history_compact) f.has_history = false;
    f.history_compact = true;
    const CompactIn in{static_cast<const float*>(g.color_dev), static_cast<const float*>(g.albedo_dev), static_cast<const float*>(g.emission_dev),
                       static_cast<const float*>(g.depth_dev), static_cast<const uint32_t*>(g.normal_dev), static_cast<const int*>(g.material_dev),
                       cam, f.w, f.h};
    run_passes(planes_of(f), st, in, CompactHist{f.prev_depth.ptr, f.prev_normal.ptr, f.prev_material.ptr}, prev_cam, s, output);
    f.has_history = true;
}

}  // namespace phk
