// ORACLE — test infrastructure only (see oracle.h).
//
// Renderer drivers and batch probes over the restated integrator (oracle_render.hpp).
// or_render follows CpuRenderer::render (src/render/cpu_renderer.cpp:20-39): row-parallel workers
// pulling rows from an atomic counter, per pixel acc = sum of samples in frame order, then sum += acc.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "oracle_render.hpp"

using namespace orc;

namespace {

PathOpts opts_of(const ph_render_settings* rs) {
    PathOpts o;
    o.max_depth = rs->path.max_depth;
    o.rr_depth = rs->path.rr_depth;
    o.clamp = rs->path.clamp;
    o.nee = rs->path.nee;
    return o;
}

template <class F>
void parallel_rows(int rows, int threads, F&& body) {
    int n = threads > 0 ? threads : default_threads();
    n = std::max(1, std::min(n, rows));
    std::atomic<int> next{0};
    auto work = [&] {
        for (;;) {
            int y = next.fetch_add(1);
            if (y >= rows) break;
            body(y);
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < n; i++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
}

void put3(float* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
void put4(float* p, V3 v, float w) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
    p[3] = w;
}

MatParams params_of(const float* p) {
    MatParams m;
    m.base = v3(p[0], p[1], p[2]);
    m.metallic = p[3];
    m.roughness = p[4];
    m.transmission = p[5];
    m.ior = p[6];
    return m;
}

}  // namespace

extern "C" {

int or_render(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, int spp, uint32_t frame0, int threads, float* sum,
              float* albedo4, float* normal_depth4) {
    if (!s || !cam || !rs || !sum || spp < 0) return -1;
    const int w = rs->width, h = rs->height;
    const PathOpts opt = opts_of(rs);
    parallel_rows(h, threads, [&](int y) {
        for (int x = 0; x < w; x++) {
            V3 acc = v3s(0);
            Primary prim;
            for (int k = 0; k < spp; k++) {
                Primary* pp = (k == 0 && (albedo4 || normal_depth4)) ? &prim : nullptr;
                acc = acc + pixel_sample(*s, *cam, x, y, w, h, frame0 + uint32_t(k), rs->seed, opt, pp, nullptr);
            }
            size_t i = size_t(y) * size_t(w) + size_t(x);
            V3 cur = v3p(sum + 3 * i) + acc;
            put3(sum + 3 * i, cur);
            if (spp > 0 && albedo4) put4(albedo4 + 4 * i, prim.albedo, 1.0f);
            if (spp > 0 && normal_depth4) put4(normal_depth4 + 4 * i, prim.normal, prim.depth);
        }
    });
    return 0;
}

int or_render_samples(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, int n, const int32_t* xs, const int32_t* ys,
                      const uint32_t* frames, float* rgb, uint32_t* ray_counts) {
    if (!s || !cam || !rs || n < 0) return -1;
    // Entries are independent; chunks of them run on the host threads.
    const int chunk = 64;
    parallel_rows((n + chunk - 1) / chunk, 0, [&](int c) {
        PathOpts opt = opts_of(rs);
        for (int i = c * chunk; i < std::min(n, (c + 1) * chunk); i++) {
            uint32_t rays = 0;
            opt.ray_count = ray_counts ? &rays : nullptr;
            V3 L = pixel_sample(*s, *cam, xs[i], ys[i], rs->width, rs->height, frames[i], rs->seed, opt, nullptr, nullptr);
            if (rgb) put3(rgb + 3 * i, L);
            if (ray_counts) ray_counts[i] = rays;
        }
    });
    return 0;
}

int or_render_gbuffer(const ph_scene_desc* s, const ph_camera_view* cam, const ph_render_settings* rs, uint32_t frame, int threads, float* color4,
                      float* albedo4, float* normal_depth4, float* position4, float* emission4) {
    if (!s || !cam || !rs || !color4 || !albedo4 || !normal_depth4 || !position4 || !emission4) return -1;
    const int w = rs->width, h = rs->height;
    const PathOpts opt = opts_of(rs);
    parallel_rows(h, threads, [&](int y) {
        for (int x = 0; x < w; x++) {
            Primary prim;
            V3 L = pixel_sample(*s, *cam, x, y, w, h, frame, rs->seed, opt, &prim, nullptr);
            size_t i = size_t(y) * size_t(w) + size_t(x);
            put4(color4 + 4 * i, L, 1.0f);
            put4(albedo4 + 4 * i, prim.albedo, 1.0f);
            put4(normal_depth4 + 4 * i, prim.normal, prim.depth);
            put4(position4 + 4 * i, prim.position, float(prim.material));
            put4(emission4 + 4 * i, prim.emission, 1.0f);
        }
    });
    return 0;
}

int or_trace(const ph_scene_desc* s, const ph_ray* rays, int n, int anyhit, ph_hit* hits, int64_t* counters, int threads) {
    if (!s || !rays || !hits || n < 0) return -1;
    const int chunk = 4096;
    const int chunks = (n + chunk - 1) / chunk;
    parallel_rows(chunks, threads, [&](int c) {
        int end = std::min(n, (c + 1) * chunk);
        for (int i = c * chunk; i < end; i++) {
            Ray r;
            r.o = v3p(rays[i].o);
            r.d = v3p(rays[i].d);
            r.tmax = rays[i].tmax;
            Hit h;
            TraceCounters tc;
            TraceCounters* pc = counters ? &tc : nullptr;
            if (anyhit) {
                walk<true>(*s, r, h, pc);
            } else {
                walk<false>(*s, r, h, pc);
            }
            hits[i].t = h.t;
            hits[i].tri = h.tri;
            hits[i].u = h.u;
            hits[i].v = h.v;
            if (counters) {
                counters[3 * size_t(i)] = tc.nodes;
                counters[3 * size_t(i) + 1] = tc.tris;
                counters[3 * size_t(i) + 2] = tc.alpha;
            }
        }
    });
    return 0;
}

int or_heatmap(const ph_scene_desc* s, const ph_camera_view* cam, int width, int height, float* visits) {
    if (!s || !cam || !visits) return -1;
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            Ray r = eye_ray(*cam, float(x) + 0.5f, float(y) + 0.5f, width, height, v2(0.5f, 0.5f));
            Hit h;
            TraceCounters tc;
            closest(*s, r, h, &tc);
            visits[size_t(y) * size_t(width) + size_t(x)] = float(tc.nodes);
        }
    return 0;
}

int or_display_color(const float* sum, int n, int samples, float exposure, int tonemap, uint32_t* out) {
    if (!sum || !out) return -1;
    float inv = samples > 0 ? 1.0f / float(samples) : 0.0f;
    for (int i = 0; i < n; i++) out[i] = pack8(orc::tonemap(v3p(sum + 3 * i) * inv, exposure, tonemap));
    return 0;
}

// heat_color of gpu_renderer.cu:108-116: a blue -> green -> yellow -> red ramp over [0, 1].
static V3 heat_ramp(float t) {
    t = sat(t);
    const V3 c0 = v3(0.05f, 0.05f, 0.35f), c1 = v3(0.0f, 0.6f, 0.4f), c2 = v3(0.95f, 0.85f, 0.1f), c3 = v3(0.85f, 0.1f, 0.05f);
    auto mix = [](V3 a, V3 b, float u) { return a + (b - a) * u; };
    if (t < 1.0f / 3) return mix(c0, c1, t * 3);
    if (t < 2.0f / 3) return mix(c1, c2, t * 3 - 1);
    return mix(c2, c3, t * 3 - 2);
}

// display_kernel of gpu_renderer.cu:129-175, every mode, over the renderer's float4 planes.
int or_display(int n, int mode, float exposure, int tonemap, int samples, float max_depth, float heat_scale, int denoised,
               const float* accum4, const float* denoised4, const float* albedo4, const float* nd4, const float* variance4,
               const float* heat, uint32_t* out) {
    if (!out || mode < 0 || mode > 5) return -1;
    const float fs = float(samples);
    const float inv = samples > 0 ? 1.0f / fs : 0.0f;
    for (int i = 0; i < n; i++) {
        V3 c;
        switch (mode) {
            case 1: c = v3p(albedo4 + 4 * i); break;
            case 2: {
                const float* v = nd4 + 4 * i;
                c = v[3] > 0 ? v3p(v) * 0.5f + v3s(0.5f) : v3s(0);
                break;
            }
            case 3: {
                const float d = nd4[4 * i + 3];
                c = d > 0 ? v3s(1.0f - sat(d / max_depth)) : v3s(0);
                break;
            }
            case 4: {
                float var;
                if (denoised) {
                    var = variance4[4 * i + 3];
                } else {
                    const float* a = accum4 + 4 * i;
                    const float mean = lum(v3p(a)) * inv;
                    var = fs > 1 ? fmaxf(0.0f, a[3] * inv - mean * mean) / (fs - 1) : 0.0f;
                }
                const float t = sqrtf(var);
                c = v3s(t / (1.0f + t));
                break;
            }
            case 5: c = heat_ramp(heat[i] * heat_scale); break;
            default:
                c = denoised ? v3p(denoised4 + 4 * i) : v3p(accum4 + 4 * i) * inv;
                c = orc::tonemap(c, exposure, tonemap);
                break;
        }
        // The debug views are display values already; packing sRGB-encodes, so they are decoded first.
        if (mode != 0 && mode != 1) c = v3(srgb_decode(c.x), srgb_decode(c.y), srgb_decode(c.z));
        out[i] = pack8(c);
    }
    return 0;
}

int or_rng_u32(uint64_t seed, uint64_t stream, int n, uint32_t* out) {
    Pcg r(seed, stream);
    for (int i = 0; i < n; i++) out[i] = r.u32();
    return 0;
}

int or_rng_float(uint64_t seed, uint64_t stream, int n, float* out) {
    Pcg r(seed, stream);
    for (int i = 0; i < n; i++) out[i] = r.uni();
    return 0;
}

int or_hash_u32(const uint32_t* in, int n, uint32_t* out) {
    for (int i = 0; i < n; i++) out[i] = mix32(in[i]);
    return 0;
}

int or_warp(int kind, int n, const float* in, float* out) {
    for (int i = 0; i < n; i++) {
        switch (kind) {
            case 0: {
                V2 d = disk_concentric(v2(in[2 * i], in[2 * i + 1]));
                out[2 * i] = d.x, out[2 * i + 1] = d.y;
                break;
            }
            case 1: put3(out + 3 * i, hemi_cosine(v2(in[2 * i], in[2 * i + 1]))); break;
            case 2: put3(out + 3 * i, sphere_uniform(v2(in[2 * i], in[2 * i + 1]))); break;
            case 3: {
                V2 b = tri_uniform(v2(in[2 * i], in[2 * i + 1]));
                out[2 * i] = b.x, out[2 * i + 1] = b.y;
                break;
            }
            case 4: {
                V2 uv = dir_to_latlong(v3p(in + 3 * i));
                out[2 * i] = uv.x, out[2 * i + 1] = uv.y;
                break;
            }
            case 5: put3(out + 3 * i, latlong_to_dir(v2(in[2 * i], in[2 * i + 1]))); break;
            default: return -1;
        }
    }
    return 0;
}

int or_frame(int n, const float* normals, const float* v, float* local, float* world_of_local, float* basis9) {
    for (int i = 0; i < n; i++) {
        Onb f(v3p(normals + 3 * i));
        V3 l = f.local(v3p(v + 3 * i));
        if (local) put3(local + 3 * i, l);
        if (world_of_local) put3(world_of_local + 3 * i, f.world(l));
        if (basis9) {
            put3(basis9 + 9 * i, f.t);
            put3(basis9 + 9 * i + 3, f.b);
            put3(basis9 + 9 * i + 6, f.n);
        }
    }
    return 0;
}

int or_offset_ray(int n, const float* p, const float* nrm, float* out) {
    for (int i = 0; i < n; i++) put3(out + 3 * i, nudge(v3p(p + 3 * i), v3p(nrm + 3 * i)));
    return 0;
}

int or_bsdf_sample(const float* params, int n, const float* wo, const float* u, float* wi, float* f, float* pdf, int32_t* flags) {
    Mix b(params_of(params));
    for (int i = 0; i < n; i++) {
        Scatter s = b.sample(v3p(wo + 3 * i), u[4 * i], v2(u[4 * i + 1], u[4 * i + 2]), u[4 * i + 3]);
        put3(wi + 3 * i, s.wi);
        put3(f + 3 * i, s.f);
        pdf[i] = s.pdf;
        flags[i] = (s.valid ? 1 : 0) | (s.transmitted ? 2 : 0);
    }
    return 0;
}

int or_bsdf_eval(const float* params, int n, const float* wo, const float* wi, float* f, float* pdf) {
    Mix b(params_of(params));
    for (int i = 0; i < n; i++) {
        float p;
        put3(f + 3 * i, b.eval(v3p(wo + 3 * i), v3p(wi + 3 * i), &p));
        pdf[i] = p;
    }
    return 0;
}

int or_bsdf_info(const float* params, float* out6) {
    Mix b(params_of(params));
    out6[0] = b.alpha;
    out6[1] = b.w_metal;
    out6[2] = b.w_opaque;
    out6[3] = b.w_glass;
    out6[4] = lum(b.albedo());
    out6[5] = b.light_samplable() ? 1.0f : 0.0f;
    return 0;
}

int or_sample_light(const ph_scene_desc* s, int n, const float* p, const float* u, float* out) {
    for (int i = 0; i < n; i++) {
        LightDraw l = draw_light(*s, v3p(p + 3 * i), u[4 * i], v2(u[4 * i + 1], u[4 * i + 2]), u[4 * i + 3]);
        float* o = out + 12 * size_t(i);
        put3(o, l.wi);
        o[3] = l.dist;
        put3(o + 4, l.radiance);
        o[7] = l.pdf;
        o[8] = l.delta ? 1.0f : 0.0f;
        o[9] = float(l.tri);
        o[10] = l.bary.x;
        o[11] = l.bary.y;
    }
    return 0;
}

int or_env_eval(const ph_scene_desc* s, int n, const float* dirs, float* radiance3, float* pdf_dir) {
    for (int i = 0; i < n; i++) {
        V3 d = v3p(dirs + 3 * i);
        if (radiance3) put3(radiance3 + 3 * i, sky_radiance(*s, d));
        if (pdf_dir) pdf_dir[i] = sky_pdf_dir(*s, d);
    }
    return 0;
}

int or_env_sample(const ph_scene_desc* s, int n, const float* u, float* dirs, float* pdf, int32_t* ok) {
    for (int i = 0; i < n; i++) {
        V3 d;
        float p = 0;
        bool good = sky_sample(*s, v2(u[2 * i], u[2 * i + 1]), d, p);
        put3(dirs + 3 * i, d);
        pdf[i] = p;
        ok[i] = good ? 1 : 0;
    }
    return 0;
}

int or_emissive_pdf(const ph_scene_desc* s, int n, const int32_t* tri, const float* dirs, const float* dist, float* pdf) {
    for (int i = 0; i < n; i++) pdf[i] = emitter_pdf(*s, tri[i], v3p(dirs + 3 * i), dist[i]);
    return 0;
}

int or_hit_emission(const ph_scene_desc* s, int n, const ph_ray* rays, const ph_hit* hits, float* emission3) {
    for (int i = 0; i < n; i++) {
        if (hits[i].tri < 0) {
            put3(emission3 + 3 * i, v3s(0));
            continue;
        }
        Ray r;
        r.o = v3p(rays[i].o);
        r.d = v3p(rays[i].d);
        r.tmax = rays[i].tmax;
        Hit h;
        h.t = hits[i].t;
        h.tri = hits[i].tri;
        h.u = hits[i].u;
        h.v = hits[i].v;
        put3(emission3 + 3 * i, shade_point(*s, r, h).emission);
    }
    return 0;
}

int or_tonemap(int n, const float* rgb, float exposure, int op, float* out, uint32_t* packed) {
    for (int i = 0; i < n; i++) {
        V3 c = orc::tonemap(v3p(rgb + 3 * i), exposure, op);
        if (out) put3(out + 3 * i, c);
        if (packed) packed[i] = pack8(c);
    }
    return 0;
}

int or_srgb(int n, const float* in, int encode, float* out) {
    for (int i = 0; i < n; i++) out[i] = encode ? srgb_encode(in[i]) : srgb_decode(in[i]);
    return 0;
}

}  // extern "C"
