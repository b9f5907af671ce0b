// SVGF denoiser on the device (Schied et al. 2017), following the reference's arithmetic and pass order
// (src/denoise/svgf.cu:19-255): temporal accumulation -> variance -> a-trous pass 0 -> history ->
// a-trous passes 1..4 -> remodulation.
//
// MI355X layout: the reference's planes are float4 per pixel (16-byte loads). The variance estimate
// stages its footprint through LDS: a 16x16 output tile reads the 22x22 halo of the G-buffer planes it
// needs once per block instead of once per tap. The a-trous passes run on lattice rows: a block takes
// 64 neighbouring pixels of 8 rows spaced one step apart, whose taps all fall on 12 rows of one lattice,
// and stages those rows whole (contiguous loads and conflict-free LDS reads, where square tiles staged
// by residue class would load strided).
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

// Edge-stopping weights. The denoiser's tolerance is 1e-3 relative, so a tap evaluates its weights as
// one product in the exponent, exp2(log2 w_l + log2 w_z + log2 w_n), with the hardware's
// single-precision exp2/log2 (a few ulp) instead of correctly rounded exp and pow.
constexpr float kLog2e = 1.44269504f;

// log2 of the normal weight max(0, d)^e: -inf where that weight is 0 (d <= 0 with e > 0), and 0 when
// e is 0.
PH_DEV float normal_log2w(float d, float e) {
    if (e == 0.0f) return 0.0f;
    return d > 0.0f ? e * __builtin_amdgcn_logf(d) : (e > 0.0f ? -__builtin_inff() : 0.0f);
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
    PH_DEV float4 ndz_w(int i) const { return nd[i]; }
    PH_DEV float mat(int i) const { return pos[i].w; }
    PH_DEV f3 position(int i, int, int) const { return xyz(pos[i]); }
    PH_DEV float4 raw_color(int i) const { return color[i]; }
};

struct RefHist {
    float4 *nd, *pos;
    PH_DEV float4 ndz(int j) const { return nd[j]; }
    PH_DEV float4 ndz_w(int j) const { return nd[j]; }
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
    // For the edge-stopping weights of the variance and a-trous passes only: the normal normalised with
    // the hardware reciprocal square root. Reprojection and the stored history decode exactly.
    PH_DEV float4 ndz_w(int i) const {
        const f3 n = oct_decode_approx(normal[i]);
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
    PH_DEV float4 ndz_w(int j) const {
        const f3 n = oct_decode_approx(normal[j]);
        return make_float4(n.x, n.y, n.z, depth[j]);
    }
    PH_DEV float mat(int j) const { return float(material[j]); }
    PH_DEV void store(const CompactIn& in, int i) const {
        depth[i] = in.depth[i];
        normal[i] = in.normal[i];
        material[i] = in.material[i];
    }
};

// Temporal accumulation (svgf.cu:37-94) on 64x4 blocks: a wave reads 64 neighbouring pixels of one row,
// so the 12-byte colour planes of the compact G-buffer stream in whole cache lines.
constexpr int kRowX = 64, kRowY = 4;

template <class In, class Hist>
__global__ __launch_bounds__(kRowX* kRowY) void temporal_kernel(In in, ph_camera_view prev_cam, int w, int h, int has_history, float alpha_c,
                                                           float alpha_m, const float4* prev_illum, const float4* prev_moments, Hist prev,
                                                           float4* out_illum, float4* out_moments) {
    const int x = blockIdx.x * kRowX + threadIdx.x, y = blockIdx.y * kRowY + threadIdx.y;
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
                const float4 pn = prev.ndz_w(j);
                if (fabsf(pn.w - dist) > 0.05f * dist + 0.01f) continue;
                // The normal test on the exactly decoded normal wherever the fast decode's few-ulp error
                // could decide it.
                float nn = dot(xyz(pn), n);
                if (fabsf(nn - 0.9f) < 1e-5f) nn = dot(xyz(prev.ndz(j)), n);
                if (nn < 0.9f) continue;
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

// The 7x7 estimate of one short-history pixel (svgf.cu:96-132): moments and illumination of the taps
// whose material matches, weighted by normal and depth. `nd_at(dx, dy)` and `mat_at(dx, dy)` read the
// neighbour at that offset (in range).
template <class NdAt, class MatAt>
PH_DEV float4 spatial_estimate(int x, int y, int w, int h, float mat, float4 ndc, NdAt nd_at, MatAt mat_at, const float4* illum, const float4* moments,
                               float phi_normal, float4 c, float var) {
    constexpr int R = 3;
    const f3 n = xyz(ndc);
    const float z = ndc.w, len = c.w;
    float wsum = 0, m1 = 0, m2 = 0;
    f3 acc = mk3(0);
    // The depth weight's exponent per unit of depth difference (svgf.cu's exp(-|dz| / (0.05 z + 1e-3))).
    const float kz = -kLog2e / (0.05f * z + 1e-3f);
#pragma unroll
    for (int dy = -R; dy <= R; dy++)
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            const int xx = x + dx, yy = y + dy;
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            if (mat_at(dx, dy) != mat) continue;
            const float4 nd = nd_at(dx, dy);
            const float wt = __builtin_amdgcn_exp2f(normal_log2w(fmaxf(0.0f, dot(n, xyz(nd))), phi_normal) + fabsf(nd.w - z) * kz);
            const int j = yy * w + xx;
            const float4 m = moments[j];
            const float4 il = illum[j];
            m1 = __builtin_fmaf(m.x, wt, m1);
            m2 = __builtin_fmaf(m.y, wt, m2);
            acc = mk3(__builtin_fmaf(il.x, wt, acc.x), __builtin_fmaf(il.y, wt, acc.y), __builtin_fmaf(il.z, wt, acc.z));
            wsum += wt;
        }
    if (wsum > 0) {
        m1 /= wsum;
        m2 /= wsum;
        var = fmaxf(0.0f, m2 - m1 * m1) * (4.0f / len);
        const f3 a3 = acc / wsum;
        c = make_float4(a3.x, a3.y, a3.z, len);
    }
    return make_float4(c.x, c.y, c.z, var);
}

// Variance pass. Pixels whose history has reached 4 frames keep their temporal estimate; the others take
// the 7x7 spatial estimate. A block with no such pixel (the common case once the camera has settled)
// only copies; any other stages the 22x22 halo of normal+depth and material in LDS once (gathering the
// taps of a block's few short-history pixels through the caches instead measured slower).

template <class In>
__global__ __launch_bounds__(TX* TY) void variance_kernel(In in, int w, int h, const float4* illum, const float4* moments, float phi_normal,
                                                           float4* out) {
    constexpr int R = 3, SX = TX + 2 * R, SY = TY + 2 * R;
    __shared__ float4 s_nd[SY][SX];
    __shared__ float s_mat[SY][SX];
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    const bool inside = x < w && y < h;
    const int i = inside ? y * w + x : 0;
    const float4 c = inside ? illum[i] : make_float4(0, 0, 0, 4.0f);
    const float mat = inside ? in.mat(i) : -1.0f;
    const bool spatial = inside && c.w < 4.0f && mat >= 0;
    if (!__syncthreads_or(spatial)) {
        if (inside) out[i] = make_float4(c.x, c.y, c.z, moments[i].z);
        return;
    }
    const int bx = blockIdx.x * TX - R, by = blockIdx.y * TY - R;
    for (int t = threadIdx.y * TX + threadIdx.x; t < SX * SY; t += TX * TY) {
        const int sy = t / SX, sx = t % SX;
        const int gx = bx + sx, gy = by + sy;
        if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
            const int j = gy * w + gx;
            s_nd[sy][sx] = in.ndz_w(j);
            s_mat[sy][sx] = in.mat(j);
        }
    }
    __syncthreads();
    if (!inside) return;
    const int ly = threadIdx.y + R, lx = threadIdx.x + R;
    float4 r = make_float4(c.x, c.y, c.z, moments[i].z);
    if (spatial)
        r = spatial_estimate(
            x, y, w, h, mat, s_nd[ly][lx], [&](int dx, int dy) { return s_nd[ly + dy][lx + dx]; }, [&](int dx, int dy) { return s_mat[ly + dy][lx + dx]; },
            illum, moments, phi_normal, c, r.w);
    out[i] = r;
}

// What a pass does with its result besides writing its output plane: the history of the next frame is
// taken after pass 0 (svgf.cu:247-248: the filtered colour with the history length, the moments, and
// the geometry), and the last pass remodulates by the albedo and adds the emission (svgf.cu:196-209:
// misses pass the raw colour through). Folding these into the passes saves a read of the plane and a
// launch each.
template <class In, class Hist>
struct PassTail {
    bool keep;     // write the pass's output plane
    bool history;  // store the next frame's history
    const float4* illum;
    const float4* moments;
    float4* prev_illum;
    float4* prev_moments;
    Hist prev;
    float4* remod;  // non-null: write the remodulated output here
};

template <class In>
PH_DEV float4 remodulated(const In& in, int i, float mat, float4 filtered) {
    if (mat < 0) return in.raw_color(i);
    const f3 c = xyz(filtered) * vmax3(in.albedo3(i), mk3(1e-3f)) + in.emission3(i);
    return make_float4(c.x, c.y, c.z, 1.0f);
}

// The a-trous filter of one pixel (svgf.cu:134-186) whose material is `mat` (>= 0). The accessors read
// what the pixel's block has staged or what lies in memory, by offset from the pixel:
//   tap_src/tap_nd/tap_mat(dx, dy): the tap at (x + dx * step, y + dy * step), in the image;
//   pre(dx, dy): src.w at (x + dx, y + dy) clamped to the image (the 3x3 variance prefilter);
//   depth(dx, dy): the depth at (x + dx, y + dy) clamped to the image (the depth gradient).
template <class TapSrc, class TapNd, class TapMat, class Pre, class Depth>
PH_DEV float4 atrous_pixel(int x, int y, int w, int h, int step, float phi_color, float phi_normal, float phi_depth, float4 center, float mat,
                           float4 ndc, TapSrc tap_src, TapNd tap_nd, TapMat tap_mat, Pre pre, Depth depth) {
    float var = 0, vw = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const float k = (dx == 0 && dy == 0) ? 0.25f : ((dx == 0 || dy == 0) ? 0.125f : 0.0625f);
            var += pre(dx, dy) * k;
            vw += k;
        }
    var /= vw;
    const f3 n = xyz(ndc);
    const float z = ndc.w;
    const float zx = depth(1, 0) - depth(-1, 0);
    const float zy = depth(0, 1) - depth(0, -1);
    const float grad = 0.5f * fmaxf(fabsf(zx), fabsf(zy));
    const float lum_c = lum(xyz(center));
    const float sigma_l = phi_color * __builtin_sqrtf(fmaxf(var, 0.0f)) + 1e-4f;
    const float kernel[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    // The luminance and depth weights (svgf.cu: exp(-|dl| / sigma_l) and exp(-|dz| / (phi_depth * grad
    // * dist + 1e-3 z + 1e-4))) as exponents per unit difference: one division per pixel for the
    // luminance and one per tap distance (dx^2 + dy^2 in {0, 1, 2, 4, 5, 8}) instead of two per tap.
    const float kl = -kLog2e / sigma_l;
    const float gz = phi_depth * grad, zb = 1e-3f * z;
    auto kz_at = [&](int d2) { return -kLog2e / (gz * (__builtin_sqrtf(float(d2)) * float(step)) + zb + 1e-4f); };
    const float kz[9] = {kz_at(0), kz_at(1), kz_at(2), 0.0f, kz_at(4), kz_at(5), 0.0f, 0.0f, kz_at(8)};
    f3 sum = mk3(0);
    float wsum = 0, vsum = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++)
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            const int xx = x + dx * step, yy = y + dy * step;
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            if (tap_mat(dx, dy) != mat) continue;
            const float4 sv = tap_src(dx, dy);
            const float4 nd = tap_nd(dx, dy);
            const float k = kernel[dx < 0 ? -dx : dx] * kernel[dy < 0 ? -dy : dy];
            const float e = fabsf(lum(xyz(sv)) - lum_c) * kl + fabsf(nd.w - z) * kz[dx * dx + dy * dy] +
                            normal_log2w(fmaxf(0.0f, dot(n, xyz(nd))), phi_normal);
            const float wt = k * __builtin_amdgcn_exp2f(e);
            sum = mk3(__builtin_fmaf(sv.x, wt, sum.x), __builtin_fmaf(sv.y, wt, sum.y), __builtin_fmaf(sv.z, wt, sum.z));
            vsum = __builtin_fmaf(wt * wt, sv.w, vsum);
            wsum += wt;
        }
    if (wsum > 0) {
        const f3 a = sum / wsum;
        return make_float4(a.x, a.y, a.z, vsum / (wsum * wsum));
    }
    return center;
}

template <class In, class Hist>
PH_DEV void finish_pass(const In& in, int i, float mat, float4 res, float4* dst, const PassTail<In, Hist>& tail) {
    if (tail.keep) dst[i] = res;
    if (tail.history) {
        tail.prev_illum[i] = make_float4(res.x, res.y, res.z, tail.illum[i].w);
        tail.prev_moments[i] = tail.moments[i];
        tail.prev.store(in, i);
    }
    if (tail.remod) tail.remod[i] = remodulated(in, i, mat, res);
}

// One a-trous pass at any step, gathering every tap through the caches: the fallback for steps beyond
// the lattice kernel's (a-trous iterations past 5).
template <class In, class Hist>
__global__ __launch_bounds__(TX* TY) void atrous_gather_kernel(In in, int w, int h, int step, float phi_color, float phi_normal, float phi_depth,
                                                                const float4* src, float4* dst, PassTail<In, Hist> tail) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    const float4 center = src[i];
    const float mat = in.mat(i);
    float4 res = center;
    if (mat >= 0) {
        auto cx = [&](int dx) { return min(max(x + dx, 0), w - 1); };
        auto cy = [&](int dy) { return min(max(y + dy, 0), h - 1); };
        res = atrous_pixel(
            x, y, w, h, step, phi_color, phi_normal, phi_depth, center, mat, in.ndz_w(i),
            [&](int dx, int dy) { return src[i + dy * step * w + dx * step]; }, [&](int dx, int dy) { return in.ndz_w(i + dy * step * w + dx * step); },
            [&](int dx, int dy) { return in.mat(i + dy * step * w + dx * step); }, [&](int dx, int dy) { return src[cy(dy) * w + cx(dx)].w; },
            [&](int dx, int dy) { return in.ndz_w(cy(dy) * w + cx(dx)).w; });
    }
    finish_pass(in, i, mat, res, dst, tail);
}

// One a-trous pass on lattice rows (steps 1 to 16): a block is 64 neighbouring pixels of each of 8 rows
// spaced `Step` apart (rows y0, y0 + Step, ...), so that all its taps fall on 12 rows of the same lattice.
// The block stages those 12 rows, 64 + 4 * Step columns wide, in LDS: coalesced loads of about 2 to 3
// texels per pixel instead of 25 gathers per pixel through the caches, and LDS reads in which a wave's
// lanes take consecutive entries. The prefilter's and gradient's neighbours in the rows above and below
// are staged rows at step 1 and are read from memory otherwise.
constexpr int kLatticeX = 64, kLatticeRows = 8;  // 4 rows per block measured 1.6 % slower

template <int Step, class In, class Hist>
__global__ __launch_bounds__(kLatticeX* kLatticeRows) void atrous_lattice_kernel(In in, int w, int h, float phi_color, float phi_normal, float phi_depth,
                                                                                 const float4* src, float4* dst, PassTail<In, Hist> tail) {
    constexpr int SW = kLatticeX + 4 * Step, SR = kLatticeRows + 4;
    __shared__ float4 s_src[SR][SW];
    __shared__ float4 s_nd[SR][SW];
    __shared__ float s_mat[SR][SW];
    // blockIdx.y = group * Step + residue: lattice rows residue + (kLatticeRows * group + r) * Step, r < kLatticeRows.
    const int residue = blockIdx.y % Step, k0 = (blockIdx.y / Step) * kLatticeRows;
    const int x0 = blockIdx.x * kLatticeX;
    for (int t = threadIdx.y * kLatticeX + threadIdx.x; t < SR * SW; t += kLatticeX * kLatticeRows) {
        const int r = t / SW, c = t % SW;
        // Clamped to the image: taps never read out-of-image entries, and the prefilter's and gradient's
        // neighbours (x +- 1, and y +- 1 at step 1) read the clamped ones, as their clamped coordinates ask.
        const int gx = min(max(x0 - 2 * Step + c, 0), w - 1), gy = min(max(residue + (k0 - 2 + r) * Step, 0), h - 1);
        const int j = gy * w + gx;
        s_src[r][c] = src[j];
        s_nd[r][c] = in.ndz_w(j);
        s_mat[r][c] = in.mat(j);
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = residue + (k0 + int(threadIdx.y)) * Step;
    if (x >= w || y >= h) return;
    const int i = y * w + x;
    const int lr = threadIdx.y + 2, lc = threadIdx.x + 2 * Step;
    const float4 center = s_src[lr][lc];
    const float mat = s_mat[lr][lc];
    float4 res = center;
    if (mat >= 0) {
        auto cx = [&](int dx) { return min(max(x + dx, 0), w - 1); };
        auto cy = [&](int dy) { return min(max(y + dy, 0), h - 1); };
        res = atrous_pixel(
            x, y, w, h, Step, phi_color, phi_normal, phi_depth, center, mat, s_nd[lr][lc],
            [&](int dx, int dy) { return s_src[lr + dy][lc + dx * Step]; }, [&](int dx, int dy) { return s_nd[lr + dy][lc + dx * Step]; },
            [&](int dx, int dy) { return s_mat[lr + dy][lc + dx * Step]; },
            [&](int dx, int dy) { return (dy == 0 || Step == 1) ? s_src[lr + dy][lc + dx].w : src[cy(dy) * w + cx(dx)].w; },
            [&](int dx, int dy) { return (dy == 0 || Step == 1) ? s_nd[lr + dy][lc + dx].w : in.ndz_w(cy(dy) * w + cx(dx)).w; });
    }
    finish_pass(in, i, mat, res, dst, tail);
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
    out[i] = remodulated(in, i, in.mat(i), filtered[i]);
}

// The planes one frame works on: the previous frame's illumination and moments (history, read when
// `has_history`, rewritten after a-trous pass 0) and four frame-local planes.
struct PassPlanes {
    int w, h;
    bool has_history;
    float4 *prev_illum, *prev_moments;
    float4 *illum, *moments, *filter0, *filter1;
};

// Grid of atrous_lattice_kernel: 64-pixel column groups by (groups of kLatticeRows lattice rows) x residues.
dim3 lattice_grid(int w, int h, int step) {
    const int rows = (h + step - 1) / step;  // lattice rows of residue 0, the most of any residue
    return dim3((w + kLatticeX - 1) / kLatticeX, ((rows + kLatticeRows - 1) / kLatticeRows) * step);
}

// The nine launches of one frame (svgf.cu:238-255) for either input format.
template <class In, class Hist>
void run_passes(const PassPlanes& pl, hipStream_t st, const In& in, const Hist& prev, const ph_camera_view& prev_cam, const ph_svgf_settings& s,
                float4* output) {
    const int w = pl.w, h = pl.h, n = w * h;
    float4 *prev_illum = pl.prev_illum, *prev_moments = pl.prev_moments;
    float4 *illum = pl.illum, *moments = pl.moments, *filter[2] = {pl.filter0, pl.filter1};
    const dim3 tile(TX, TY), grid((w + TX - 1) / TX, (h + TY - 1) / TY);
    const int lin = 256, lgrid = (n + lin - 1) / lin;
    hipLaunchKernelGGL((temporal_kernel<In, Hist>), dim3((w + kRowX - 1) / kRowX, (h + kRowY - 1) / kRowY), dim3(kRowX, kRowY), 0, st, in, prev_cam, w, h, int(pl.has_history), s.alpha_color, s.alpha_moments,
                       prev_illum, prev_moments, prev, illum, moments);
    hipLaunchKernelGGL((variance_kernel<In>), grid, tile, 0, st, in, w, h, illum, moments, s.phi_normal, filter[0]);
    int src = 0;
    const int iters = s.atrous_iterations;
    for (int it = 0; it < iters; it++) {
        const int step = 1 << it;
        const bool last = it == iters - 1;
        // The last pass's plane is needed only when it is plane 6, the variance buffer a display reads.
        PassTail<In, Hist> tail{!last || 1 - src == 0, it == 0, illum, moments, prev_illum, prev_moments, prev, last ? output : nullptr};
        if (step == 1)
            hipLaunchKernelGGL((atrous_lattice_kernel<1, In, Hist>), lattice_grid(w, h, 1), dim3(kLatticeX, kLatticeRows), 0, st, in, w, h, s.phi_color,
                               s.phi_normal, s.phi_depth, filter[src], filter[1 - src], tail);
        else if (step == 2)
            hipLaunchKernelGGL((atrous_lattice_kernel<2, In, Hist>), lattice_grid(w, h, 2), dim3(kLatticeX, kLatticeRows), 0, st, in, w, h, s.phi_color,
                               s.phi_normal, s.phi_depth, filter[src], filter[1 - src], tail);
        else if (step == 4)
            hipLaunchKernelGGL((atrous_lattice_kernel<4, In, Hist>), lattice_grid(w, h, 4), dim3(kLatticeX, kLatticeRows), 0, st, in, w, h, s.phi_color,
                               s.phi_normal, s.phi_depth, filter[src], filter[1 - src], tail);
        else if (step == 8)
            hipLaunchKernelGGL((atrous_lattice_kernel<8, In, Hist>), lattice_grid(w, h, 8), dim3(kLatticeX, kLatticeRows), 0, st, in, w, h, s.phi_color,
                               s.phi_normal, s.phi_depth, filter[src], filter[1 - src], tail);
        else if (step == 16)
            hipLaunchKernelGGL((atrous_lattice_kernel<16, In, Hist>), lattice_grid(w, h, 16), dim3(kLatticeX, kLatticeRows), 0, st, in, w, h, s.phi_color,
                               s.phi_normal, s.phi_depth, filter[src], filter[1 - src], tail);
        else
            hipLaunchKernelGGL((atrous_gather_kernel<In, Hist>), grid, tile, 0, st, in, w, h, step, s.phi_color, s.phi_normal, s.phi_depth, filter[src],
                               filter[1 - src], tail);
        src = 1 - src;
    }
    if (iters == 0) {
        hipLaunchKernelGGL((history_kernel<In, Hist>), dim3(lgrid), dim3(lin), 0, st, in, n, filter[src], illum, moments, prev_illum, prev_moments,
                           prev);
        hipLaunchKernelGGL((remodulate_kernel<In>), dim3(lgrid), dim3(lin), 0, st, in, n, filter[src], output);
    }
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
