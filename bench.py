"""Benchmark of the MI355X rendering core (one JSON line on stdout).

Headline (`value`): Msamples/s of the batched wavefront path tracer at 1920x1080, 64 spp, max depth 16,
on the 1,048,578-triangle sphere field under an HDR sky (synthetic scene, seeded). One step is one
complete 1080p/64-spp render with the scene resident in HBM. Alongside it:
  * the plain one-thread-per-pixel kernel on the same render (the 1x baseline of the wavefront form);
  * traversal Mrays/s: 4M camera rays (2048x2048) and 4M shadow rays on the 262,144-triangle sphere soup;
  * roofline of the dominant kernel (closest-hit traversal): algorithmic bytes per launch from the
    node/triangle/alpha visit counts of an instrumented pass over the same deterministic work, divided
    by the launch duration measured with HIP events on the library's stream;
  * cpu_baseline: the oracle (the CPU restatement, bit-identical to the reference CpuRenderer on the
    committed fixtures) on a bounded sample, with the host's lscpu core and thread counts; CPU baselines
    of the Cornell, traversal and SVGF legs sit in those legs' "cpu" entries.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Msamples/s at 1080p 64spp; traversal Mrays/s; request p50/p99 latency"

# Algorithmic bytes of one traversal: ray in 28, hit out 16, 64 per node visit, 48 per triangle test,
# 80 + 12 + 24 + 16 per alpha test (material, indices, uvs, 4 texels).
RAY_IO, NODE_B, TRI_B, ALPHA_B = 28 + 16, 64, 48, 80 + 12 + 24 + 16
# Algorithmic bytes of one SVGF frame per pixel with the 16-byte planes the denoiser reads: each pass
# reads its distinct inputs once and writes its outputs once. temporal: 5 current planes + 4 history
# planes in, illumination + moments out (11 x 16); variance: illumination, moments, normal, position in,
# 1 out (5 x 16); five a-trous passes: filtered, normal, position in, 1 out (4 x 16 each); history:
# filtered, illumination, moments, normal, position in, 4 out (9 x 16); remodulate: filtered, colour,
# albedo, emission, position in, 1 out (6 x 16). Total 816 bytes.
SVGF_BYTES_PER_PIXEL = 11 * 16 + 5 * 16 + 5 * 4 * 16 + 9 * 16 + 6 * 16
# The same count for the workload's compact G-buffer (12-byte colour, albedo and emission, 4-byte depth,
# normal code and material; float4 internal planes; the previous frame's geometry kept as depth, normal
# code and material). temporal: 48 current + 44 history in, 32 out (124); variance: illumination,
# moments, geometry in, 1 out (60); a-trous: filtered + geometry in, 1 out (44 each); history: filtered,
# illumination, moments, geometry in, 44 out (104); remodulate: material, filtered, albedo, emission,
# colour in, 1 out (72). Total 580 bytes.
SVGF_COMPACT_BYTES_PER_PIXEL = 124 + 60 + 5 * 44 + 104 + 72


# Algorithmic bytes of a path vertex beyond its rays (per sample: the scope table's B_sample): a shading
# vertex reads its triangle (48), indices (12), normals (36), uvs (24) and material (80); a light sample
# reads an alias entry (16), the emissive record (8), the emitter's triangle (48) and tri_emissive (4); the
# accumulation reads and writes 16 B per pixel per render call.
B_SHADE, B_NEE, B_ACC = 48 + 12 + 36 + 24 + 80, 16 + 8 + 48 + 4, 32

# The best launch of the engine's autotuner candidates for the headline workload (1080p field, 64 spp) on an
# MI355X. Neighbouring path-slot counts lie within the tuner's 2 % acceptance margin of each other, so a
# tuning run may stop at one of them; the default run uses this one so that its figures and the committed
# profiles describe the same launch. --autotune runs the tuner instead.
TUNED_LAUNCH = ("block=128,stack_depth=8,path_slots=31104000,blocks_per_cu=16,sort_materials=1,use_graph=1,graph_iterations=5,"
                "refill_lanes=16,lds_nodes=0,ray_bins=1,xcd_parts=1,sample_split=64,finish_percent=10")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the CPU baselines: the CPUs this process may run on, capped by OMP_NUM_THREADS
    (the host-core share a job on a shared box is given)."""
    n = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, omp) if omp > 0 else n


def host_cpu_info():
    """The host's cores and threads as lscpu reports them, and the CPUs this process may use."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    info["usable_cpus"] = len(os.sched_getaffinity(0))
    info["threads_used"] = cpu_threads()
    return info


def sample_bytes(cs, pixels):
    """Algorithmic bytes of one render call from its instrumented counts: every traced ray (extension and
    shadow, with their node, triangle and alpha tests), a shading vertex per extension ray (an upper
    bound: a miss reads one environment texel instead) and a light sample per shadow ray, plus the
    accumulation. The rays of the paths finished one per thread at the end of the render are not
    instrumented; they are counted at the average bytes of the others."""
    rays = traversal_bytes(cs.rays_extend, cs.node_visits, cs.tri_tests, cs.alpha_tests)
    rays += traversal_bytes(cs.rays_shadow, cs.shadow_node_visits, cs.shadow_tri_tests, cs.shadow_alpha_tests)
    b = rays + cs.rays_extend * B_SHADE + cs.rays_shadow * B_NEE
    return b * (1 + cs.rays_finish / max(1, cs.rays_extend + cs.rays_shadow)) + pixels * B_ACC


def traversal_bytes(rays, nodes, tris, alpha):
    return rays * RAY_IO + nodes * NODE_B + tris * TRI_B + alpha * ALPHA_B


def bench_render(args, gpu, host, RenderSettings):
    W, H, SPP, DEPTH = 1920, 1080, args.spp, 16
    t = time.time()
    scene = host.Scene.field(1024, seed=1, env_width=2048)
    scene.build()
    st = scene.stats()
    log(f"field scene: {scene.num_tris} triangles, {st.nodes} nodes, BVH {time.time() - t:.2f} s")
    rs = RenderSettings.make(W, H, max_depth=DEPTH, rr_depth=3, seed=1)
    cam = scene.camera_view(W / H)
    r = gpu.GpuRenderer()
    ok, err = r.upload(scene)
    assert ok, err
    r.resize(W, H)
    r.set_engine(gpu.ENGINE_WAVEFRONT)
    cfg = None
    launch = args.launch or ("" if args.autotune or args.no_autotune or SPP != 64 else TUNED_LAUNCH)
    if launch:
        kw = {k: int(v) for k, v in (item.split("=") for item in launch.split(","))}
        cfg = r.set_wavefront_config(**kw)
        log(f"launch configuration {cfg.as_dict()}")
    elif not args.no_autotune:
        t = time.time()
        cfg = r.autotune(cam, rs, args.spp)
        log(f"autotune {time.time() - t:.1f} s -> {cfg.as_dict()}")
    for i in range(args.warmup):
        r.reset()
        r.render(cam, rs, SPP)
        log(f"warmup {i}: {r.stats().ms_total:.1f} ms")
    times = []
    rays_ext = rays_sh = rays_fin = finish_paths = 0
    for i in range(args.steps):
        r.reset()
        t0 = time.perf_counter()
        r.render(cam, rs, SPP)
        times.append(time.perf_counter() - t0)
        s = r.stats()
        rays_ext, rays_sh, rays_fin, finish_paths = s.rays_extend, s.rays_shadow, s.rays_finish, s.finish_paths
        log(f"step {i}: {times[-1] * 1e3:.1f} ms, {s.iterations} iterations")
    total = sum(times)
    samples = W * H * SPP
    value = samples * args.steps / total / 1e6
    # Per-kernel durations: the same render launched directly, with HIP events around every traversal
    # launch on the library's stream (a graph replay cannot be timed per kernel).
    tuned = r.wavefront_config()
    r.set_wavefront_config(use_graph=0)
    r.reset()
    r.render(cam, rs, SPP)
    es = r.stats()
    r.set_wavefront_config(tuned)
    ext_ms, ext_launches, shadow_ms = es.ms_extend, es.extend_launches, es.ms_shadow
    log(f"event-timed render: {es.ms_total:.1f} ms, extend {ext_ms:.1f} ms over {ext_launches} launches, shadow {shadow_ms:.1f} ms")
    # Instrumented pass (same deterministic work) for the visit counts.
    r.reset()
    r.set_counting(True)
    r.render(cam, rs, SPP)
    cs = r.stats()
    r.set_counting(False)
    ext_bytes = traversal_bytes(cs.rays_extend, cs.node_visits, cs.tri_tests, cs.alpha_tests)
    sh_bytes = traversal_bytes(cs.rays_shadow, cs.shadow_node_visits, cs.shadow_tri_tests, cs.shadow_alpha_tests)
    avg_launch_ms = ext_ms / max(1, ext_launches)
    bytes_per_launch = ext_bytes / max(1, cs.extend_launches)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    b_sample = sample_bytes(cs, W * H) / samples
    out = dict(value=value, ms_per_step=total / args.steps * 1e3, samples=samples, config=cfg.as_dict() if cfg else r.wavefront_config().as_dict(),
               bytes_per_sample=b_sample, sample_achieved_gbs=b_sample * value * 1e6 / 1e9,
               rays_extend=rays_ext, rays_shadow=rays_sh, rays_finish=rays_fin, finish_paths=finish_paths,
               rays_per_sample=(rays_ext + rays_sh + rays_fin) / samples, mrays_s=(rays_ext + rays_sh + rays_fin) * args.steps / total / 1e6,
               extend_ms_total=ext_ms, shadow_ms_total=shadow_ms,
               extend_avg_launch_ms=avg_launch_ms, extend_bytes_per_launch=bytes_per_launch, achieved_gbs=achieved,
               node_visits_per_ray=cs.node_visits / max(1, cs.rays_extend), tri_tests_per_ray=cs.tri_tests / max(1, cs.rays_extend),
               shadow_bytes=sh_bytes, extend_bytes=ext_bytes)
    if not args.no_megakernel:
        r.set_engine(gpu.ENGINE_MEGAKERNEL)
        r.reset()
        r.render(cam, rs, 1)
        r.reset()
        mk_spp = max(1, SPP // max(1, args.megakernel_divisor))
        t0 = time.perf_counter()
        r.render(cam, rs, mk_spp)
        mk = time.perf_counter() - t0
        out["megakernel_msamples_s"] = W * H * mk_spp / mk / 1e6
        out["megakernel_spp_timed"] = mk_spp
        out["speedup_vs_megakernel"] = value / out["megakernel_msamples_s"]
        log(f"megakernel: {out['megakernel_msamples_s']:.1f} Msamples/s ({mk_spp} spp in {mk:.2f} s)")
    out["display"] = bench_display(gpu, r, rs, W, H)
    r.close()
    return scene, cam, rs, out


def bench_display(gpu, r, rs, W, H, reps=20):
    """The display path at 1080p: tonemap into the renderer's display buffer and copy it to the host
    (download_display, the reference's viewer path without GL interop), against tonemapping straight into
    pinned host memory (ph_host_alloc as the display target)."""
    import numpy as np
    pinned = gpu.PinnedArray(W * H * 4)
    r.display(rs)
    r.download_display()
    r.display(rs, target=pinned.ptr)
    t0 = time.perf_counter()
    for _ in range(reps):
        r.display(rs)
        img = r.download_display()
    copy_ms = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        r.display(rs, target=pinned.ptr)
    pinned_ms = (time.perf_counter() - t0) / reps * 1e3
    same = bool(np.array_equal(pinned.numpy(np.uint32, (H, W)), img))
    pinned.free()
    res = dict(width=W, height=H, display_and_copy_ms=round(copy_ms, 3), display_to_pinned_host_ms=round(pinned_ms, 3), identical=same)
    log(f"display: {res}")
    return res


def bench_traversal(gpu, host, cpu=False):
    import numpy as np
    from photon_amd.types import HIT_DTYPE
    t = time.time()
    scene = host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1).build()
    cam = scene.camera_view(1.0)
    cam_rays = host.camera_rays(cam, 2048, 2048)
    shadow = scene.shadow_rays(4194304, seed=2)
    log(f"traversal scene: {scene.num_tris} triangles, setup {time.time() - t:.2f} s")
    r = gpu.GpuRenderer()
    ok, err = r.upload(scene)
    assert ok, err
    res = {}
    for name, rays, anyhit in (("closest", cam_rays, False), ("any", shadow, True)):
        rd = gpu.DeviceArray.from_numpy(rays)
        hd = gpu.DeviceArray(len(rays) * HIT_DTYPE.itemsize)
        best = None
        for block, depth in ((64, 16), (128, 16), (128, 24), (256, 16)):
            for top in (0, 256, 512):
                kw = dict(anyhit=anyhit, block=block, stack_depth=depth, persistent=True, lds_nodes=top)
                r.trace_config(rd.ptr, len(rays), hd.ptr, **kw)
                ms = min(r.trace_config(rd.ptr, len(rays), hd.ptr, **kw) for _ in range(3))
                if best is None or ms < best[0]:
                    best = (ms, block, depth, top)
        res[name] = dict(mrays_s=len(rays) / (best[0] * 1e-3) / 1e6, ms=best[0], block=best[1], stack_depth=best[2], lds_nodes=best[3])
        # Per-ray roofline: algorithmic bytes from the instrumented traversal of the same rays.
        cd = gpu.DeviceArray(len(rays) * 12)
        r.trace(rd.ptr, len(rays), hd.ptr, anyhit=anyhit, counts_dev=cd.ptr)
        c = cd.numpy(np.uint32, (len(rays), 3)).astype(np.int64).sum(0)
        nbytes = traversal_bytes(len(rays), int(c[0]), int(c[1]), int(c[2]))
        res[name].update(bytes_per_ray=round(nbytes / len(rays), 1), achieved_gbs=round(nbytes / (best[0] * 1e-3) / 1e9, 1),
                         roofline_frac=round(nbytes / (best[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         node_visits_per_ray=round(c[0] / len(rays), 2), tri_tests_per_ray=round(c[1] / len(rays), 2))
        cd.free()
        log(f"traversal {name}: {res[name]}")
    r.close()
    if cpu:
        # The oracle's multithreaded traversal of the same rays (CPU baseline of the traversal workload).
        from oracle import oracle_py as O
        threads = cpu_threads()
        for name, rays, anyhit in (("closest", cam_rays, False), ("any", shadow, True)):
            O.trace(scene.desc(), rays[:65536], anyhit=anyhit, threads=threads)
            t0 = time.perf_counter()
            O.trace(scene.desc(), rays, anyhit=anyhit, threads=threads)
            dt = time.perf_counter() - t0
            res[name]["cpu_mrays_s"] = round(len(rays) / dt / 1e6, 2)
        res["cpu"] = dict(kind="port", threads=threads, sample="the same 4,194,304 + 4,194,304 rays, oracle traversal on all threads")
        log(f"traversal CPU: closest {res['closest']['cpu_mrays_s']} Mrays/s, any {res['any']['cpu_mrays_s']} Mrays/s on {threads} threads")
    return res


def bench_cornell(args, gpu, host, cpu):
    """The plain-method workload: the 36-triangle Cornell box at 512x512, 16 spp, depth 8, seed 1, on the
    wavefront engine and the per-pixel kernel, with its per-sample roofline; and the oracle on the same
    render on the host's threads."""
    from photon_amd.types import RenderSettings
    W = H = 512
    SPP = 16
    scene = host.Scene.builtin("cornell-c1").build()
    rs = RenderSettings.make(W, H, max_depth=8, rr_depth=3, seed=1)
    cam = scene.camera_view(1.0)
    r = gpu.GpuRenderer()
    ok, err = r.upload(scene)
    assert ok, err
    r.resize(W, H)
    res = dict(workload="Cornell box, 36 triangles, 512x512, 16 spp, max depth 8, seed 1")
    for engine, key in ((gpu.ENGINE_WAVEFRONT, "wavefront"), (gpu.ENGINE_MEGAKERNEL, "per_pixel")):
        r.set_engine(engine)
        if engine == gpu.ENGINE_WAVEFRONT and not args.no_autotune:
            r.autotune(cam, rs, SPP)
        best = None
        for _ in range(5):
            r.reset()
            t0 = time.perf_counter()
            r.render(cam, rs, SPP)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res[key + "_msamples_s"] = round(W * H * SPP / best / 1e6, 1)
    r.set_engine(gpu.ENGINE_WAVEFRONT)
    r.reset()
    r.set_counting(True)
    r.render(cam, rs, SPP)
    cs = r.stats()
    r.set_counting(False)
    r.close()
    b = sample_bytes(cs, W * H) / (W * H * SPP)
    res.update(bytes_per_sample=round(b, 1), rays_per_sample=round((cs.rays_extend + cs.rays_shadow + cs.rays_finish) / (W * H * SPP), 3),
               achieved_gbs=round(b * res["wavefront_msamples_s"] * 1e6 / 1e9, 1))
    res["roofline_frac"] = round(res["achieved_gbs"] / HBM_PEAK_GBS, 4)
    if cpu:
        from oracle import oracle_py as O
        threads = cpu_threads()
        O.render(scene.desc(), cam, rs, 1, threads=threads)  # 1-spp warm-up, as the reference's bench does
        t0 = time.perf_counter()
        O.render(scene.desc(), cam, rs, SPP, threads=threads)
        dt = time.perf_counter() - t0
        res["cpu"] = dict(value=round(W * H * SPP / dt / 1e6, 2), unit="Msamples/s", kind="port", threads=threads,
                          sample=f"the full render, {dt:.2f} s wall")
    log(f"cornell: {res}")
    return res


def bench_svgf(args, gpu, host, field, cpu=False):
    """Real-time sequence at 1600x900, camera orbiting 0.5 deg/frame: a 1-spp sample into the workload's
    compact G-buffer, then SVGF on it, per frame; then the reference's float4 G-buffer path
    (render_denoised) on the same sequence for comparison."""
    import math
    from photon_amd.types import RenderSettings, SvgfSettings
    W, H = 1600, 900
    cam = field.camera()
    target = cam["target"]
    base = [cam["position"][i] - target[i] for i in range(3)]
    rs = RenderSettings.make(W, H, max_depth=16, rr_depth=3, seed=1)
    sv = SvgfSettings.make()
    r = gpu.GpuRenderer()
    ok, err = r.upload(field)
    assert ok, err
    r.resize(W, H)

    def view(frame):
        a = math.radians(0.5 * frame)
        c, s = math.cos(a), math.sin(a)
        p = (target[0] + c * base[0] + s * base[2], target[1] + base[1], target[2] - s * base[0] + c * base[2])
        field.set_camera(p, target, cam["up"], cam["vfov"])
        return field.camera_view(W / H)

    f = gpu.Svgf()
    f.resize(W, H)
    gb = gpu.CompactGBuffer(W, H)
    out = gpu.DeviceArray(W * H * 16)
    prev = view(0)
    trace_ms, svgf_ms, wall = [], [], []
    for k in range(args.svgf_frames):
        cur = view(k)
        t0 = time.perf_counter()
        r.sample_compact(cur, rs, gb)
        trace_ms.append(r.last_frame_ms())
        f.run_compact(gb, cur, prev, sv, out.ptr)
        svgf_ms.append(f.last_ms())
        wall.append((time.perf_counter() - t0) * 1e3)
        prev = cur
    cpu_ms = None
    if cpu:
        # The oracle's denoiser (row-parallel passes) on the last frames' compact G-buffers.
        from oracle import oracle_py as O
        ora = O.SvgfOracle(W, H)
        frames, prev_c = [], view(0)
        for k in range(4):
            cur = view(k)
            r.sample_compact(cur, rs, gb)
            frames.append((cur, prev_c, gb.numpy()))
            prev_c = cur
        times = []
        for cur, prev_c, g in frames:
            t0 = time.perf_counter()
            ora.run_compact(g, cur, prev_c, sv)
            times.append((time.perf_counter() - t0) * 1e3)
        cpu_ms = dict(value=round(sum(times[1:]) / 3, 2), unit="ms/frame", kind="port", threads=cpu_threads(),
                      sample="4 frames of the orbit at 1600x900 on the device-traced compact G-buffers, the last 3 timed")
    # The reference's format: render_denoised (float4 G-buffer) over the same orbit.
    ref_ms = []
    prev = view(0)
    for k in range(min(16, args.svgf_frames)):
        cur = view(k)
        r.render_denoised(cur, prev, rs, sv)
        ref_ms.append(r.stats().ms_other)
        prev = cur
    field.set_camera(cam["position"], target, cam["up"], cam["vfov"])
    r.close()
    k = slice(1, None)  # the first frame also allocates
    mean = lambda v: sum(v[k]) / len(v[k])
    res = dict(frames=args.svgf_frames, width=W, height=H, gbuffer="compact (fp32 RGB, fp32 depth, fp16 octahedral normal, int32 material)",
               svgf_ms=round(mean(svgf_ms), 4), gbuffer_trace_ms=round(mean(trace_ms), 3), frame_ms=round(mean(wall), 3),
               svgf_bytes_per_pixel=SVGF_COMPACT_BYTES_PER_PIXEL, float4_svgf_ms=round(mean(ref_ms), 4), float4_bytes_per_pixel=SVGF_BYTES_PER_PIXEL)
    res["svgf_achieved_gbs"] = round(W * H * SVGF_COMPACT_BYTES_PER_PIXEL / (res["svgf_ms"] * 1e-3) / 1e9, 1)
    res["float4_achieved_gbs"] = round(W * H * SVGF_BYTES_PER_PIXEL / (res["float4_svgf_ms"] * 1e-3) / 1e9, 1)
    res["roofline_frac"] = round(res["svgf_achieved_gbs"] / HBM_PEAK_GBS, 4)
    if cpu_ms:
        res["cpu"] = cpu_ms
    log(f"svgf: {res}")
    return res


RESOLUTIONS = ((256, 256), (512, 512), (1280, 720), (1920, 1080))
SPPS = (4, 16, 64, 256)
DEPTHS = (4, 8, 16)


def bench_scene_build(gpu, host):
    """Scene admission costs: glTF import of the 1,048,578-triangle field (written to a temporary .gltf
    first, untimed), the threaded host BVH build of the imported scene, and the device BVH build
    (ph_bvh_build) against the host build on the 262,144-triangle sphere soup."""
    import tempfile
    import numpy as np
    out = {}
    field = host.Scene.field(1024, seed=1, env_width=64)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "field.gltf")
        field.save_gltf(path)
        size = os.path.getsize(path)
        t = time.perf_counter()
        s = host.Scene.from_gltf(path)
        load_s = time.perf_counter() - t
    t = time.perf_counter()
    s.build()
    build_s = time.perf_counter() - t
    out["gltf_import"] = dict(triangles=s.num_tris, file_mb=round(size / 1e6, 1), load_ms=round(load_s * 1e3, 1),
                              mb_s=round(size / 1e6 / load_s, 1), mtris_s=round(s.num_tris / 1e6 / load_s, 2),
                              host_bvh_ms=round(build_s * 1e3, 1), host_threads=os.cpu_count())
    log(f"glTF import: {out['gltf_import']}")
    soup = host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1)
    r = soup.raw()
    p = np.ctypeslib.as_array(r.positions, (r.num_vertices * 3,)).reshape(-1, 3).copy()
    i = np.ctypeslib.as_array(r.indices, (r.num_tris * 3,)).reshape(-1, 3).copy()
    gpu.bvh_build(p, i)  # warm-up
    dev = min(gpu.bvh_build(p, i)[2].build_ms for _ in range(3))
    soup.build()
    host_ms = soup.stats().build_ms
    out["bvh_build"] = dict(triangles=len(i), device_ms=round(dev, 2), host_ms=round(host_ms, 2), host_threads=os.cpu_count())
    log(f"BVH build: {out['bvh_build']}")
    return out


def draw_jobs(n, seed):
    """Jobs of the serving workload: resolution, spp, depth and scene drawn uniformly, seeded per job."""
    import numpy as np
    jobs = []
    for j in range(n):
        g = np.random.default_rng([seed, j])
        w, h = RESOLUTIONS[g.integers(len(RESOLUTIONS))]
        jobs.append(dict(scene=int(g.integers(3)), w=w, h=h, spp=SPPS[g.integers(len(SPPS))], depth=DEPTHS[g.integers(len(DEPTHS))],
                         seed=int(g.integers(1 << 31))))
    return jobs


def ideal_server(arrival_s, work_s):
    """Response times of the same job stream on an ideal server: one resource of the saturated capacity,
    preemptive at no cost, where a job of `work_s` seconds of work finishes after exactly that much service
    and any share of the capacity can go to any job. Returns {"fifo", "srpt", "ps"} -> response seconds
    per job, for first-come-first-served, shortest-remaining-work-first and processor sharing (every job
    present gets an equal share). Arrivals must be sorted."""
    import heapq
    import numpy as np
    a = np.asarray(arrival_s, np.float64)
    w = np.asarray(work_s, np.float64)
    n = len(a)
    out = {}
    done = np.zeros(n)
    t = 0.0
    for k in range(n):
        t = max(t, a[k]) + w[k]
        done[k] = t
    out["fifo"] = done - a

    done = np.zeros(n)
    t, i, heap = 0.0, 0, []
    while i < n or heap:
        if not heap:
            t = max(t, a[i])
        while i < n and a[i] <= t:
            heapq.heappush(heap, [w[i], i])
            i += 1
        rem, k = heap[0]
        nxt = a[i] if i < n else np.inf
        if t + rem <= nxt:
            t += rem
            heapq.heappop(heap)
            done[k] = t
        else:
            heap[0][0] = rem - (nxt - t)
            t = nxt
    out["srpt"] = done - a

    # Processor sharing: between events every present job receives 1/m of the capacity; the next event is
    # an arrival or the completion of the job with the least remaining work. Remaining work is kept as
    # attained-service offsets so that one event costs O(m log m) at most.
    done = np.zeros(n)
    t, i = 0.0, 0
    attained = 0.0  # service each present job has received since the reference point
    heap = []  # (work + attained at arrival, job): the job leaves when `attained` reaches the key
    while i < n or heap:
        if not heap:
            t = max(t, a[i])
        while i < n and a[i] <= t:
            heapq.heappush(heap, (w[i] + attained, i))
            i += 1
        m = len(heap)
        nxt = a[i] if i < n else np.inf
        need = heap[0][0] - attained
        if t + need * m <= nxt:
            t += need * m
            attained = heap[0][0]
            while heap and heap[0][0] <= attained:
                done[heapq.heappop(heap)[1]] = t
        else:
            attained += (nxt - t) / m
            t = nxt
    out["ps"] = done - a
    return out


def bench_serving(args, gpu, host, field, launch):
    """Open-loop Poisson arrivals at 80% of the saturated throughput over three resident scenes."""
    import numpy as np
    t = time.time()
    scenes = [host.Scene.field(10, seed=7, env_width=2048).build(),
              host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1).build(), field]
    log(f"serving scenes: {[s.num_tris for s in scenes]} triangles, setup {time.time() - t:.1f} s")
    srv = gpu.Server(path_slots=args.serve_slots, iterations_per_step=args.serve_iterations,
                     schedule={"fifo": gpu.SCHEDULE_FIFO, "srpt": gpu.SCHEDULE_SRPT, "hrrn": gpu.SCHEDULE_HRRN, "share": gpu.SCHEDULE_SHARE,
                               "deadline": gpu.SCHEDULE_DEADLINE}[args.serve_schedule], deadline_stretch=args.serve_stretch,
                     pass_samples=args.serve_pass_samples, min_items=args.serve_min_items,
                     express_slots=args.serve_express_slots, express_samples=args.serve_express_samples,
                     express_iterations=args.serve_express_iterations)
    cfg = gpu.WavefrontConfig(**launch)
    cfg.path_slots = 0
    srv.set_wavefront_config(cfg)
    ids = [srv.add_scene(s) for s in scenes]

    def submit(job):
        return srv.submit(ids[job["scene"]], job["w"], job["h"], job["spp"], scenes[job["scene"]].camera_view(job["w"] / job["h"]),
                          max_depth=job["depth"], rr_depth=3, seed=job["seed"])

    def drain(live, lat, t_arrival):
        for k in list(live):
            st = srv.status(live[k])
            if st.state == gpu.JOB_DONE:
                lat[k] = st.done_ms - t_arrival[k]
                srv.release(live[k])
                del live[k]

    # Saturated throughput: a burst of jobs, measured until the backlog falls to a quarter.
    burst = draw_jobs(args.serve_burst, seed=99)
    t0 = srv.now_ms()
    live = {k: submit(j) for k, j in enumerate(burst)}
    ids_burst = dict(live)
    lat = {}
    while srv.pending > args.serve_burst // 4:
        srv.step()
        drain(live, lat, {k: t0 for k in ids_burst})
    t1 = srv.now_ms()
    done_samples = sum(srv.status(j).samples_done if k in live else srv.status(j).samples for k, j in ids_burst.items())
    peak = done_samples / ((t1 - t0) * 1e-3)
    srv.run_until_idle()
    drain(live, lat, {k: t0 for k in ids_burst})
    log(f"serving peak: {peak / 1e6:.1f} Msamples/s over {(t1 - t0) / 1e3:.1f} s")

    # Open loop at 80% of peak: Poisson arrivals with the workload's mean job size.
    mean_px = np.mean([w * h for w, h in RESOLUTIONS])
    mean_samples = mean_px * np.mean(SPPS)
    rate = args.serve_load * peak / mean_samples  # jobs per second
    jobs = draw_jobs(args.serve_jobs, seed=2024)
    gaps = np.random.default_rng(7).exponential(1.0 / rate, len(jobs))
    start = srv.now_ms() + 1.0
    arrival = start + np.cumsum(gaps) * 1e3
    live, lat, t_arr = {}, {}, {}
    nxt = 0
    reported = 0
    while nxt < len(jobs) or srv.pending:
        now = srv.now_ms()
        while nxt < len(jobs) and arrival[nxt] <= now:
            live[nxt] = submit(jobs[nxt])
            t_arr[nxt] = arrival[nxt]
            nxt += 1
        if srv.pending:
            srv.step()
            drain(live, lat, t_arr)
        elif nxt < len(jobs):
            time.sleep(max(0.0, (arrival[nxt] - srv.now_ms()) * 1e-3))
        if len(lat) >= reported + 500:
            reported = len(lat) // 500 * 500
            log(f"serving: {len(lat)} of {len(jobs)} jobs done")
    end = srv.now_ms()
    # Isolated service time: each job again, alone on the idle server (submission to completion). With
    # --serve-isolated classes, each distinct (scene, resolution, spp, depth) is timed once alone, on the
    # first job of the stream in that class, and stands for every job of the class.
    def alone_ms(job):
        j = submit(job)
        while srv.status(j).state != gpu.JOB_DONE:
            srv.step()
        st = srv.status(j)
        srv.release(j)
        return st.done_ms - st.submit_ms

    iso = np.zeros(len(jobs))
    if args.serve_isolated == "classes":
        key = lambda job: (job["scene"], job["w"], job["h"], job["spp"], job["depth"])  # noqa: E731
        timed = {}
        for k, job in enumerate(jobs):
            if key(job) not in timed:
                timed[key(job)] = alone_ms(job)
                if len(timed) % 24 == 0:
                    log(f"serving: {len(timed)} job classes timed alone")
            iso[k] = timed[key(job)]
    else:
        for k, job in enumerate(jobs):
            iso[k] = alone_ms(job)
            if (k + 1) % 500 == 0:
                log(f"serving: {k + 1} of {len(jobs)} jobs timed alone")
    srv.close()
    L = np.array([lat[k] for k in range(len(jobs))])
    samples = np.array([j["w"] * j["h"] * j["spp"] for j in jobs], np.float64)
    slow = L / iso
    slow_rate = L / (samples / peak * 1e3)  # against the job's work at the saturated rate
    by_size = {}
    for lo, hi, name in ((0, 2e6, "<2M samples"), (2e6, 5e7, "2M-50M"), (5e7, 1e12, ">50M")):
        sel = (samples >= lo) & (samples < hi)
        if sel.any():
            by_size[name] = dict(jobs=int(sel.sum()), p50_ms=round(float(np.percentile(L[sel], 50)), 2),
                                 p99_ms=round(float(np.percentile(L[sel], 99)), 2))
    res = dict(jobs=len(jobs), load=args.serve_load, arrival_rate_jobs_s=round(rate, 3), peak_msamples_s=round(peak / 1e6, 2),
               msamples_s=round(samples.sum() / ((end - start) * 1e-3) / 1e6, 2), duration_s=round((end - start) * 1e-3, 2),
               latency_ms={"p50": round(float(np.percentile(L, 50)), 2), "p99": round(float(np.percentile(L, 99)), 2)},
               slowdown={"p50": round(float(np.percentile(slow, 50)), 2), "p99": round(float(np.percentile(slow, 99)), 2)},
               slowdown_vs_peak_rate={"p50": round(float(np.percentile(slow_rate, 50)), 2), "p99": round(float(np.percentile(slow_rate, 99)), 2)},
               isolated_ms={"p50": round(float(np.percentile(iso, 50)), 2), "p99": round(float(np.percentile(iso, 99)), 2)},
               by_job_size=by_size, path_slots=args.serve_slots, iterations_per_step=args.serve_iterations, schedule=args.serve_schedule,
               deadline_stretch=args.serve_stretch,
               pass_samples=args.serve_pass_samples, min_items=args.serve_min_items, express_slots=args.serve_express_slots,
               express_samples=args.serve_express_samples, express_iterations=args.serve_express_iterations,
               isolated=args.serve_isolated)
    # The same arrivals and job sizes on an ideal server of the saturated capacity (no concurrency limits,
    # free preemption): what the classical schedules reach on this stream, against the measured figure of
    # the same definition (latency over the job's samples at the saturated rate).
    arr_s = (arrival - arrival[0]) * 1e-3
    ideal = {}
    for name, resp in ideal_server(arr_s, samples / peak).items():
        s = resp / (samples / peak)
        p50, p99 = float(np.percentile(s, 50)), float(np.percentile(s, 99))
        ideal[name] = {"p50": round(p50, 2), "p99": round(p99, 2), "p99_over_p50": round(p99 / p50, 2)}
    res["offered_load"] = round(float(samples.sum() / peak / (arr_s[-1] + gaps[0])), 3)
    res["ideal_server_slowdown_vs_peak_rate"] = ideal
    worst = np.argsort(slow)[-5:][::-1]
    res["worst_slowdown"] = [dict(res=f"{jobs[k]['w']}x{jobs[k]['h']}", spp=jobs[k]["spp"], depth=jobs[k]["depth"], scene=jobs[k]["scene"],
                                  latency_ms=round(float(L[k]), 1), isolated_ms=round(float(iso[k]), 1), slowdown=round(float(slow[k]), 1))
                             for k in worst]
    res["latency_p99_over_p50"] = round(res["latency_ms"]["p99"] / res["latency_ms"]["p50"], 2)
    res["slowdown_p99_over_p50"] = round(res["slowdown"]["p99"] / res["slowdown"]["p50"], 2)
    log(f"serving: {res}")
    return res


def measured_traffic(spp, launch):
    """HBM bytes per closest-hit launch from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel calibration), or None when the
    summary was taken at another spp or launch configuration (rays per launch, and so traffic, differ)."""
    path = os.path.join(ROOT, "profiles", "extend_kernel_field_1080p_pmc.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.strip().startswith("- "):
            k, v = line.strip()[2:].split(":", 1)
            vals[k.strip()] = v.strip().split()[0]
    try:
        if int(vals.get("spp", -1)) != spp:
            return None
        if launch is None or vals.get("launch") != ",".join(f"{k}={v}" for k, v in launch.items()):
            return None
        return int(float(vals["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def cpu_baseline(scene, cam, rs, spp):
    """The oracle (the repository's CPU restatement, bit-identical to the reference CpuRenderer on the
    committed fixtures) on the full 1080p frame at `spp` samples with the host's threads: a bounded sample
    of the workload (about 20 s of CPU time at 4 spp on 16 threads), after a 1-spp warm-up as the
    reference's bench does (render_cli.cpp:319-327)."""
    from oracle import oracle_py as O
    threads = cpu_threads()
    O.render(scene.desc(), cam, rs, 1, threads=threads)
    t0 = time.perf_counter()
    O.render(scene.desc(), cam, rs, spp, threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=rs.width * rs.height * spp / dt / 1e6, unit="Msamples/s", cores=threads, kind="port",
                sample=f"{rs.width}x{rs.height} at {spp} spp, max depth {rs.path.max_depth}, same scene and sequences; {dt:.2f} s wall",
                host=host_cpu_info())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true", help="render with the engine's default launch")
    ap.add_argument("--autotune", action="store_true", help="tune the render's launch instead of using the tuned one")
    ap.add_argument("--no-megakernel", action="store_true")
    ap.add_argument("--megakernel-divisor", type=int, default=1, help="time the plain kernel on spp/divisor samples")
    ap.add_argument("--no-traversal", action="store_true")
    ap.add_argument("--no-cornell", action="store_true", help="skip the 512x512 Cornell box leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=4, help="samples per pixel of the CPU baseline's bounded sample")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-svgf", action="store_true")
    ap.add_argument("--no-scene-build", action="store_true", help="skip the glTF import and BVH build leg")
    ap.add_argument("--scene-build-only", action="store_true", help="run only the glTF import and BVH build leg")
    ap.add_argument("--svgf-frames", type=int, default=64)
    ap.add_argument("--serve-jobs", type=int, default=200, help="jobs of the open-loop serving run (the full stream is 10,000)")
    ap.add_argument("--serve-burst", type=int, default=48)
    ap.add_argument("--serve-isolated", choices=("jobs", "classes"), default="jobs",
                    help="isolated service time of every job alone, or of each (scene, size, spp, depth) class once")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving leg (needs --launch)")
    ap.add_argument("--serve-load", type=float, default=0.8)
    ap.add_argument("--serve-slots", type=int, default=1 << 21)
    ap.add_argument("--serve-iterations", type=int, default=0, help="bounces per server step (0: the server's default)")
    ap.add_argument("--serve-schedule", choices=("srpt", "fifo", "hrrn", "share", "deadline"), default="srpt")
    ap.add_argument("--serve-stretch", type=float, default=0.0, help="deadline schedule: deadline = submit + stretch x estimated service (0: 4)")
    ap.add_argument("--serve-pass-samples", type=int, default=0, help="pixel samples per job pass (0: the server's default)")
    ap.add_argument("--serve-min-items", type=int, default=0, help="split smaller jobs into sample sub-streams (0: the server's default)")
    ap.add_argument("--serve-express-slots", type=int, default=0, help="path slots of the express lane (0: the server's default, -1: none)")
    ap.add_argument("--serve-express-samples", type=int, default=0, help="largest job (samples) of the express lane (0: the server's default)")
    ap.add_argument("--serve-express-iterations", type=int, default=0, help="bounces per express-lane run (0: as the main lane)")
    ap.add_argument("--launch", default="", help="wavefront launch configuration as key=value,... (skips the autotuner)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (the known-size kernel that scales counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the renderer runs on one GPU: --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    from photon_amd import gpu, host
    from photon_amd.types import RenderSettings
    assert gpu.gpu_available(), "no HIP device"
    if args.scene_build_only:
        print(json.dumps({"scene_build": bench_scene_build(gpu, host)}), flush=True)
        return
    info = gpu.device_info(0)
    log(f"device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    if args.calibrate_copy_mb > 0:
        nbytes = args.calibrate_copy_mb << 20
        src, dst = gpu.DeviceArray(nbytes), gpu.DeviceArray(nbytes)
        src.zero()
        gpu.lib().ph_copy_device(dst.ptr, src.ptr, nbytes)
        log(f"calibration copy: {nbytes} bytes")
        src.free()
        dst.free()
    if args.serving_only:
        if not args.launch:
            sys.exit("--serving-only needs --launch (the launch configuration the render leg would tune)")
        field = host.Scene.field(1024, seed=1, env_width=2048).build()
        launch = {k: int(v) for k, v in (item.split("=") for item in args.launch.split(","))}
        print(json.dumps({"serving": bench_serving(args, gpu, host, field, launch)}), flush=True)
        return
    scene, cam, rs, rnd = bench_render(args, gpu, host, RenderSettings)
    cornell = None if args.no_cornell else bench_cornell(args, gpu, host, not args.no_cpu)
    trav = None if args.no_traversal else bench_traversal(gpu, host, not args.no_cpu)
    svgf = None if args.no_svgf else bench_svgf(args, gpu, host, scene, not args.no_cpu)
    serve = None if args.no_serve else bench_serving(args, gpu, host, scene, rnd["config"])
    build = None if args.no_scene_build else bench_scene_build(gpu, host)
    cpu = None if args.no_cpu else cpu_baseline(scene, cam, rs, args.cpu_spp)
    line = {
        "metric": METRIC,
        "value": round(rnd["value"], 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rnd["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and HDR sky)",
        "config": {"workload": "wavefront path trace, 1920x1080, 64 spp, max depth 16, 1,048,578-triangle sphere field, 2048x1024 HDR sky",
                   "engine": "wavefront", "launch": rnd["config"], "spp": args.spp},
        "roofline": {"bound": "hbm", "kernel": "extend_kernel (closest-hit traversal)", "achieved": round(rnd["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(rnd["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.spp, rnd["config"]),
                     "bytes_per_launch": int(rnd["extend_bytes_per_launch"]), "avg_launch_ms": round(rnd["extend_avg_launch_ms"], 4)},
        "cpu_baseline": cpu,
        "sample_roofline": {"bytes_per_sample": round(rnd["bytes_per_sample"], 1), "achieved": round(rnd["sample_achieved_gbs"], 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(rnd["sample_achieved_gbs"] / HBM_PEAK_GBS, 4),
                            "note": "rays, a shading vertex per extension ray (upper bound) and a light sample per shadow ray, plus accumulation"},
        "mrays_s": round(rnd["mrays_s"], 1),
        "rays_per_sample": round(rnd["rays_per_sample"], 3),
        "finish_paths": int(rnd["finish_paths"]),
        "node_visits_per_ray": round(rnd["node_visits_per_ray"], 2),
        "tri_tests_per_ray": round(rnd["tri_tests_per_ray"], 2),
        "display": rnd.get("display"),
        "megakernel_msamples_s": round(rnd.get("megakernel_msamples_s", 0.0), 3) or None,
        "speedup_vs_megakernel": round(rnd.get("speedup_vs_megakernel", 0.0), 3) or None,
        "cornell": cornell,
        "traversal": trav,
        "svgf": svgf,
        "serving": serve,
        "scene_build": build,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
