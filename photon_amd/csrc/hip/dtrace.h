// BVH traversal over 64-byte dual-box nodes for wave64: one lane per ray.
//
// Semantics are those of the reference traversal (src/render/traverse.h:24-147), which fixes the hit
// that is reported when several triangles tie:
//   * both child boxes are slab-tested against the tmax current at node entry;
//   * leaf children are tested in place, left leaf before right leaf, shrinking tmax as they go;
//   * interior children are re-tested against the shortened tmax, nearest first (left on ties), the
//     farther one pushed; the stack holds at most 64 entries and a push beyond that is dropped;
//   * Moller-Trumbore with |det| < 1e-14 rejection and 0 < t < tmax acceptance; alpha-masked
//     triangles are skipped when the interpolated base-colour alpha is below the cutoff;
//   * any-hit returns on the first accepted triangle and reports only t and the triangle.
// The stack lives in LDS: each lane owns a column of `Depth` entries, interleaved across the lanes of
// the block so that the 64 lanes of a wave touch 64 consecutive dwords (no bank conflicts); entries
// beyond `Depth` spill to per-lane scratch in the same order, so the visit order never changes.
#pragma once

#include "dscene.h"

namespace phd {

struct HitRec {
    float t;
    int tri;
    float u, v;
};

struct TraceCount {
    uint32_t nodes, tris, alpha;
};

constexpr int kMaxStack = 64;

// Per-lane overflow entries of a LaneStack<Depth> (private memory; at least one so the array exists).
template <int Depth>
struct SpillArea {
    int v[Depth < kMaxStack ? kMaxStack - Depth : 1];
};

// The stack's handles only: the LDS column and the overflow array are separate objects, held through
// pointers of their own address spaces, so the handle stays in registers and every access is a
// ds_read/ds_write or a scratch access (never a flat-address one).
using LdsInt = __attribute__((address_space(3))) int;
using PrivateInt = __attribute__((address_space(5))) int;

template <int Depth>
struct LaneStack {
    LdsInt* column;    // this lane's first LDS entry
    int stride;        // lanes per block
    PrivateInt* spill;  // this lane's SpillArea<Depth>
    PH_DEV LaneStack(int* lds_column, int lanes, int* spill_area)
        : column((LdsInt*)lds_column), stride(lanes), spill((PrivateInt*)spill_area) {}
    PH_DEV void store(int sp, int v) {
        if (Depth >= kMaxStack || sp < Depth)
            column[sp * stride] = v;
        else
            spill[sp - Depth] = v;
    }
    PH_DEV int load(int sp) const { return (Depth >= kMaxStack || sp < Depth) ? column[sp * stride] : spill[sp - Depth]; }
};

PH_DEV bool mt_hit(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float tmax, float& t, float& u, float& v) {
    f3 p = cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-14f) return false;
    float inv = 1.0f / det;
    f3 s = o - v0;
    u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross(s, e1);
    v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * inv;
    return t > 0.0f && t < tmax;
}

PH_DEV float slab(float4 lo, float4 hi, f3 o, f3 inv, float tmax) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tnear <= tfar ? tnear : kInf;
}

PH_DEV f3 guarded_reciprocal(f3 d) {
    auto r = [](float x) { return fabsf(x) > 1e-20f ? 1.0f / x : copysignf(1e20f, x); };
    return {r(d.x), r(d.y), r(d.z)};
}

// True when the masked triangle lets the ray through at (u, v) (traverse.h:55-64).
PH_DEV bool alpha_cut(const DScene& s, int tri, int material, float u, float v) {
    const float4* mp = s.materials + 5 * material;
    float alpha = ldg(mp).w;
    float4 d = ldg(mp + 3);
    float cutoff = d.x;
    int tex = __float_as_int(d.w);
    if (tex >= 0 && s.uvs) {
        const int* idx = s.indices + 3 * tri;
        f2 uv = uv_at(s.uvs, index_at(idx, 0)) * (1 - u - v) + uv_at(s.uvs, index_at(idx, 1)) * u + uv_at(s.uvs, index_at(idx, 2)) * v;
        alpha *= tex_bilinear(s, tex, uv).w;
    }
    return alpha < cutoff;
}

// What traversal reads of a scene, held in registers: the node and triangle arrays, the whole scene
// for the (rare) alpha tests, and optionally the top `top_n` nodes staged in LDS by the block.
struct TraceScene {
    const float4* nodes;
    const float4* tris;
    const DScene* full;
    const float4* top;
    int top_n;
};

PH_DEV TraceScene trace_scene(const DScene& s) { return TraceScene{s.nodes, s.tris, &s, nullptr, 0}; }

// Copies nodes [0, n) into the block's LDS array `top` (4 float4 per node). All threads call.
PH_DEV void stage_top_nodes(float4* top, const float4* nodes, int n) {
    for (int i = threadIdx.x; i < 4 * n; i += blockDim.x) top[i] = ldg(nodes + i);
    __syncthreads();
}

// One ray's traversal position. walk_step does one unit of work per call: either it fetches the node
// at `node` and slab-tests its two children, or it tests one triangle of those children's leaves. A
// wave whose lanes stand at different points therefore runs one node fetch and one triangle test per
// step (both loads in flight together) instead of waiting for its longest leaf list, and a lane can
// start a new ray as soon as its own finishes (the persistent kernels refill lanes between steps).
// Every ray still performs exactly the reference's tests in the reference's order.
struct Walker {
    f3 o, d, inv;
    float tmax;
    int node, sp;
    bool found;
    // The fetched node whose children are pending (pos >= 0): child links and counts as stored in the
    // node, child entry distances (kInf: missed), and how many leaf triangles have been tested so far.
    int pos;
    int lidx, ridx, lcount, rcount;
    float tl, tr;
};

PH_DEV Walker walker(f3 o, f3 d, float tmax) {
    Walker w{};
    w.o = o;
    w.d = d;
    w.inv = guarded_reciprocal(d);
    w.tmax = tmax;
    w.pos = -1;
    return w;
}

// One unit of the traversal of w (see Walker). Returns false once the traversal is over (any-hit: at
// the first accepted triangle). Testing one triangle per call also for coherent rays, whose lanes reach
// their leaves together, measured faster than looping over the whole leaf list when no lane of the wave
// needs a node fetch.
template <bool AnyHit, bool Count, class Stack>
PH_DEV bool walk_step(const TraceScene& ts, Walker& w, HitRec& hit, Stack& stack, TraceCount* cnt) {
    const bool fetch = w.pos < 0;
    int tri = 0;
    if (!fetch) {
        // The next leaf triangle: the left leaf's in order, then the right leaf's.
        const int nl = (w.tl != kInf && w.lcount > 0) ? w.lcount : 0;
        tri = w.pos < nl ? w.lidx + w.pos : w.ridx + (w.pos - nl);
    }
    // One memory round trip per step for both kinds of work: every lane issues its loads (a node's four
    // float4, or a triangle's three) before any lane uses them, so a wave whose lanes do both waits once.
    float4 q0, q1, q2, q3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (fetch && w.node < ts.top_n) {
        const float4* np = ts.top + 4 * w.node;
        q0 = np[0];
        q1 = np[1];
        q2 = np[2];
        q3 = np[3];
    } else {
        const float4* p = fetch ? ts.nodes + 4 * w.node : ts.tris + 3 * tri;
        q0 = ldg(p);
        q1 = ldg(p + 1);
        q2 = ldg(p + 2);
        if (fetch) q3 = ldg(p + 3);
    }
    if (fetch) {
        // The node: test both child boxes against the tmax at node entry.
        if (Count) cnt->nodes++;
        w.tl = slab(q0, q1, w.o, w.inv, w.tmax);
        w.tr = slab(q2, q3, w.o, w.inv, w.tmax);
        w.lcount = __float_as_int(q1.w);
        w.rcount = __float_as_int(q3.w);
        w.lidx = __float_as_int(q0.w);
        w.ridx = __float_as_int(q2.w);
        w.pos = 0;
    } else {
        w.pos++;
        float th, u, v;
        if (Count) cnt->tris++;
        if (mt_hit(xyz(q0), xyz(q1), xyz(q2), w.o, w.d, w.tmax, th, u, v)) {
            bool cut = false;
            if (__float_as_int(q1.w) & PH_TRI_ALPHA_TESTED) {
                if (Count) cnt->alpha++;
                cut = alpha_cut(*ts.full, tri, __float_as_int(q0.w), u, v);
            }
            if (!cut) {
                hit.t = th;
                hit.tri = tri;
                w.found = true;
                if (AnyHit) return false;
                w.tmax = th;
                hit.u = u;
                hit.v = v;
            }
        }
    }
    const int nl = (w.tl != kInf && w.lcount > 0) ? w.lcount : 0;
    const int nr = (w.tr != kInf && w.rcount > 0) ? w.rcount : 0;
    if (w.pos < nl + nr) return true;
    // Leaves done: interior children against the shortened tmax, nearest first (left on ties), the
    // farther one pushed.
    w.pos = -1;
    const bool li = w.tl != kInf && w.lcount <= 0 && w.lidx >= 0;
    const bool ri = w.tr != kInf && w.rcount <= 0 && w.ridx >= 0;
    int next0 = li ? w.lidx : w.ridx;
    float tn0 = li ? w.tl : w.tr;
    if (li && ri) {
        int next1 = w.ridx;
        float tn1 = w.tr;
        if (tn0 > tn1) {
            next1 = next0;
            next0 = w.ridx;
            tn1 = tn0;
            tn0 = w.tr;
        }
        if (tn1 < w.tmax && w.sp < kMaxStack) stack.store(w.sp++, next1);
    }
    if ((li || ri) && tn0 < w.tmax) {
        w.node = next0;
        return true;
    }
    if (w.sp == 0) return false;
    w.node = stack.load(--w.sp);
    return true;
}

// Persistent work distribution without atomics (atomics on one address serialize): the waves take the
// list's 64-entry batches in turn (wave k takes batches k, k + W, k + 2W, ...). With `parts` = 8 the
// list is first cut into 8 contiguous parts, one per XCD (workgroups are dispatched to the 8 XCDs in
// turn, block b on XCD b % 8, and each XCD has its own L2), so the waves sharing an L2 work on
// neighbouring entries; that trades L2 locality against balance between XCDs, whose parts of an image
// can cost very differently.
struct WaveShare {
    int first;    // first batch of this XCD's part
    int batches;  // batches in the part
    int wave;     // this wave's rank among the XCD's waves
    int waves;    // waves on the XCD
};

PH_DEV WaveShare wave_share(int n, int xcd_parts) {
    const int parts = (xcd_parts == 8 && gridDim.x >= 8) ? 8 : 1;
    const int wpb = blockDim.x >> 6;
    const int part = blockIdx.x % parts;
    const int blocks_here = (int(gridDim.x) - part + parts - 1) / parts;
    const int total = (n + 63) / 64;
    const int per = (total + parts - 1) / parts;
    const int first = min(total, part * per), last = min(total, first + per);
    return WaveShare{first, last - first, int(blockIdx.x / parts) * wpb + int(threadIdx.x >> 6), blocks_here * wpb};
}

// Persistent traversal of entries [0, n) with lane refill. A lane whose ray is finished takes the next
// entry of its wave's share as soon as at least `refill` lanes of the wave are free (or all are), so a
// wave is not held by its slowest ray; the wave leaves when its share is drained and none of its lanes
// has a ray. Each ray is walked exactly as traverse() walks it: only the lane it runs on and the moment
// it starts change. `load(i, ts, w, hit)` starts entry i (and may set the lane's scene);
// `finish(ts, w, hit)` stores its result.
template <bool AnyHit, bool Count, class Stack, class Load, class Finish>
PH_DEV void trace_stream(int n, int parts, int refill, TraceScene ts, Stack& stack, TraceCount& cnt, Load load, Finish finish) {
    const WaveShare ws = wave_share(n, parts);
    int pos = 0;  // entries of this wave's share already handed to lanes
    bool drained = ws.wave >= ws.batches;
    bool live = false;
    Walker w{};
    HitRec hit{};
    for (;;) {
        if (!drained) {
            const uint64_t need = __ballot(!live);
            const int k = __popcll(need);
            if (k >= refill || k == 64) {
                if (!live) {
                    const int j = pos + __popcll(need & ((1ull << __lane_id()) - 1ull));
                    const int b = ws.wave + (j >> 6) * ws.waves;
                    const int i = (ws.first + b) * 64 + (j & 63);
                    if (b < ws.batches && i < n) {
                        load(i, ts, w, hit);
                        live = true;
                    }
                }
                pos += k;
                if (ws.wave + (pos >> 6) * ws.waves >= ws.batches) drained = true;
            }
        }
        if (__ballot(live) == 0) break;
        if (live && !walk_step<AnyHit, Count>(ts, w, hit, stack, &cnt)) {
            finish(ts, w, hit);
            live = false;
        }
    }
}

template <bool AnyHit, bool Count, class Stack>
PH_DEV bool traverse(const DScene& s, f3 o, f3 d, float tmax, HitRec& hit, Stack& stack, TraceCount* cnt) {
    const TraceScene ts = trace_scene(s);
    Walker w = walker(o, d, tmax);
    while (walk_step<AnyHit, Count>(ts, w, hit, stack, cnt)) {
    }
    return w.found;
}

}  // namespace phd
