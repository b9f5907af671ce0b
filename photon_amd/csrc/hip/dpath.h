// One whole path per lane: the plain one-thread-per-pixel integrator (the reference's trace_path and
// render_sample, src/render/integrator.h:151-263 and src/render/pixel.h:18-29). It is the 1x baseline
// of the wavefront engine and the body of the real-time G-buffer kernel.
#pragma once

#include "dshade.h"
#include "dtrace.h"

namespace phd {

struct PathOpts {
    int max_depth, rr_depth;
    float clamp;
    int nee;
};

struct PrimaryOut {
    f3 position;
    float depth;
    f3 normal;
    int material;
    f3 albedo;
    f3 emission;
};

// A path between two vertices: what it has gathered, its throughput, the previous vertex's BSDF pdf and
// whether that vertex skipped light sampling, and the depth of the next vertex. A new path starts at
// depth 0 with nothing gathered, unit throughput and prev_delta set.
struct PathState {
    f3 L, beta;
    float prev_pdf;
    bool prev_delta;
    int depth;
};

// Rays a path traced (extension and shadow), when the caller counts them.
struct RayTally {
    unsigned ext, shadow;
};

// The path from its next extension ray (ro, rd) on, with the generator in the state it has at that
// point: the remaining vertices exactly as a path that started here would run them. trace_path is the
// whole path; the wavefront engine resumes its last paths in flight with this.
template <class Stack, bool Count>
PH_DEV f3 trace_path_from(const DScene& s, f3 ro, f3 rd, Pcg& rng, const PathOpts& opt, PathState ps, PrimaryOut* primary, Stack& stack,
                          TraceCount* cnt, RayTally* tally) {
    f3 L = ps.L, beta = ps.beta;
    float prev_pdf = ps.prev_pdf;
    bool prev_delta = ps.prev_delta;
    const bool has_lights = s.p_env + s.p_emissive + s.p_punctual > 0;
    for (int depth = ps.depth;; depth++) {
        HitRec hit{kInf, -1, 0.0f, 0.0f};
        if (tally) tally->ext++;
        if (!traverse<false, Count>(s, ro, rd, kInf, hit, stack, cnt)) {
            f3 le = sky_radiance(s, rd);
            float w = (prev_delta || s.p_env <= 0) ? 1.0f : mis_power(prev_pdf, sky_pdf(s, rd));
            L = L + clamp_contribution(beta * le * w, depth, opt.clamp);
            if (depth == 0 && primary) {
                primary->depth = 0;
                primary->material = -1;
                primary->emission = le;
                primary->albedo = mk3(1);
                primary->normal = -rd;
                primary->position = ro + rd * 1e4f;
            }
            break;
        }
        Surface it = shade_point(s, ro, rd, hit.t, hit.tri, hit.u, hit.v);
        const f3 wo_world = -rd;
        if (!black(it.emission)) {
            float w = 1.0f;
            if (!prev_delta) w = mis_power(prev_pdf, emitter_pdf(s, hit.tri, rd, hit.t));
            L = L + clamp_contribution(beta * it.emission * w, depth, opt.clamp);
        }
        const Mix bsdf = make_mix(it.params);
        if (depth == 0 && primary) {
            primary->position = it.p;
            primary->depth = hit.t;
            primary->normal = it.ns;
            primary->material = it.material;
            primary->albedo = mix_albedo(bsdf);
            primary->emission = it.emission;
        }
        if (depth + 1 >= opt.max_depth) break;

        const f3 wo = to_local(it.frame, wo_world);
        const bool nee = opt.nee && has_lights && light_samplable(bsdf);
        if (nee) {
            float ug = pcg_uniform(rng);
            f2 ul = pcg_uniform2(rng);
            float uk = pcg_uniform(rng);
            LightDraw ls = draw_light(s, it.p, ug, ul, uk);
            if (ls.pdf > 0 && !black(ls.radiance)) {
                f3 wi = to_local(it.frame, ls.wi);
                float pdf_b;
                f3 f = mix_eval(bsdf, wo, wi, &pdf_b);
                if (!black(f)) {
                    f3 so = nudge(it.p, dot(ls.wi, it.ng) > 0 ? it.ng : -it.ng);
                    float stmax = ls.dist == kInf ? kInf : ls.dist * (1.0f - 1e-4f);
                    HitRec sh{kInf, -1, 0.0f, 0.0f};
                    if (tally) tally->shadow++;
                    if (!traverse<true, Count>(s, so, ls.wi, stmax, sh, stack, cnt)) {
                        float w = ls.delta ? 1.0f : mis_power(ls.pdf, pdf_b);
                        L = L + clamp_contribution(beta * f * fabsf(wi.z) * ls.radiance * (w / ls.pdf), depth, opt.clamp);
                    }
                }
            }
        }
        float u_lobe = pcg_uniform(rng);
        f2 u = pcg_uniform2(rng);
        float u_sub = pcg_uniform(rng);
        Scatter bs = mix_sample(bsdf, wo, u_lobe, u, u_sub);
        if (!bs.valid) break;
        beta = beta * (bs.f * (fabsf(bs.wi.z) / bs.pdf));
        if (!finite3(beta) || black(beta)) break;
        prev_pdf = bs.pdf;
        prev_delta = !nee;
        f3 wi_world = to_world(it.frame, bs.wi);
        float side = dot(wi_world, it.ng) * dot(wo_world, it.ng);
        if (bs.transmitted ? side >= 0 : side <= 0) break;
        ro = nudge(it.p, dot(wi_world, it.ng) > 0 ? it.ng : -it.ng);
        rd = unit(wi_world);
        if (depth + 1 >= opt.rr_depth) {
            float q = fmaxf(0.05f, 1.0f - maxc(beta));
            if (pcg_uniform(rng) < q) break;
            beta = beta * (1.0f / (1.0f - q));
        }
    }
    return L;
}

template <class Stack, bool Count>
PH_DEV f3 trace_path(const DScene& s, f3 ro, f3 rd, Pcg& rng, const PathOpts& opt, PrimaryOut* primary, Stack& stack, TraceCount* cnt,
                     RayTally* tally = nullptr) {
    return trace_path_from<Stack, Count>(s, ro, rd, rng, opt, PathState{mk3(0), mk3(1), 0.0f, true, 0}, primary, stack, cnt, tally);
}

// render_sample: seeds the pixel's generator, draws jitter and lens, traces the path, and zeroes a
// non-finite or negative result.
template <class Stack, bool Count>
PH_DEV f3 pixel_sample(const DScene& s, const ph_camera_view& cam, int x, int y, int width, int height, uint32_t frame, uint32_t seed,
                       const PathOpts& opt, PrimaryOut* primary, Stack& stack, TraceCount* cnt, RayTally* tally = nullptr) {
    uint32_t pixel = uint32_t(y) * uint32_t(width) + uint32_t(x);
    Pcg rng = pixel_rng(pixel, frame, seed);
    f2 jitter = pcg_uniform2(rng);
    f2 lens = pcg_uniform2(rng);
    f3 o, d;
    eye_ray(cam, float(x) + jitter.x, float(y) + jitter.y, width, height, lens, o, d);
    f3 L = trace_path<Stack, Count>(s, o, d, rng, opt, primary, stack, cnt, tally);
    if (!finite3(L) || L.x < 0 || L.y < 0 || L.z < 0) L = mk3(0);
    return L;
}

}  // namespace phd
