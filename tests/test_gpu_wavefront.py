"""The wavefront engine against the per-pixel kernel and the oracle.

A slot runs one pixel's samples in order, so the wavefront image must be bit-identical to the
one-thread-per-pixel kernel's for every launch shape, queue size and graph setting, and within the
workload tolerance (0.1% on 99.9% of pixels) of the CPU oracle. With the samples of each pixel split
into sub-streams, the image is deterministic for a given split and within rounding of the in-order sum.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.types import RenderSettings

from gpu_util import bitwise_fraction, radiance_agreement, uploaded

pytestmark = pytest.mark.gpu


def _render(scene, w, h, spp, engine, depth=8, seed=1, cfg=None, calls=1):
    rs = RenderSettings.make(w, h, max_depth=depth, rr_depth=3, seed=seed)
    cam = scene.camera_view(w / h)
    r = uploaded(scene, w, h, engine=engine)
    if cfg:
        r.set_wavefront_config(**cfg)
    for _ in range(calls):
        r.render(cam, rs, spp // calls)
    img = r.download_image()
    g = [gpu.copy_device_to_numpy(r.buffer_ptr(b), w * h * 16, np.float32, (h, w, 4)) for b in (gpu.BUF_ALBEDO, gpu.BUF_NORMAL_DEPTH)]
    acc = gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_ACCUM), w * h * 16, np.float32, (h, w, 4))
    st = r.stats()
    r.close()
    return img, acc, g, st


@pytest.mark.parametrize("name,w,h,spp", [("cornell-c1", 96, 96, 8), ("materials", 96, 54, 8), ("glass", 80, 60, 8), ("furnace", 48, 48, 4)])
def test_wavefront_bitwise_equals_megakernel(name, w, h, spp):
    scene = host.Scene.builtin(name).build()
    a_img, a_acc, a_g, _ = _render(scene, w, h, spp, gpu.ENGINE_MEGAKERNEL)
    b_img, b_acc, b_g, st = _render(scene, w, h, spp, gpu.ENGINE_WAVEFRONT)
    print(f"{name}: {st.iterations} iterations, accum bitwise {bitwise_fraction(a_acc, b_acc):.6f}")
    assert a_acc.tobytes() == b_acc.tobytes()
    assert a_g[0].tobytes() == b_g[0].tobytes() and a_g[1].tobytes() == b_g[1].tobytes()


@pytest.mark.parametrize("cfg", [dict(use_graph=0), dict(use_graph=1, graph_iterations=4), dict(use_graph=1, graph_iterations=3),
                                 dict(path_slots=1000), dict(refill_lanes=1), dict(refill_lanes=64), dict(sort_materials=0),
                                 dict(block=256, stack_depth=16, lds_nodes=512), dict(block=128, stack_depth=24, lds_nodes=100),
                                 dict(ray_bins=0), dict(ray_bins=1, path_slots=777), dict(xcd_parts=8), dict(xcd_parts=8, path_slots=100),
                                 dict(block=64, stack_depth=8, blocks_per_cu=2), dict(block=256, stack_depth=24, blocks_per_cu=4),
                                 dict(sample_split=4, path_slots=16384), dict(sample_split=4, path_slots=5000, graph_iterations=2),
                                 dict(sample_split=64, path_slots=12288)])
def test_wavefront_config_does_not_change_the_image(cfg):
    """Every launch shape gives the in-order accumulation; so does one sample per sub-stream (a split
    equal to spp, also when asked for more), whose partial sums are single samples added in order."""
    scene = host.Scene.builtin("cornell-c1").build()
    _, ref_acc, _, _ = _render(scene, 64, 64, 4, gpu.ENGINE_MEGAKERNEL)
    _, acc, _, _ = _render(scene, 64, 64, 4, gpu.ENGINE_WAVEFRONT, cfg=cfg)
    assert acc.tobytes() == ref_acc.tobytes()


@pytest.mark.parametrize("cfg", [dict(sample_split=2), dict(sample_split=4, path_slots=1500), dict(sample_split=8, use_graph=0),
                                 dict(sample_split=3, refill_lanes=1), dict(sample_split=8, path_slots=20000, block=256, stack_depth=16)])
def test_sample_split_is_deterministic_and_within_tolerance(cfg):
    """Sub-streams change only how each pixel's in-order sample sums are grouped: the image is within
    rounding of the in-order accumulation, and the same for any launch shape with the same split."""
    scene = host.Scene.builtin("cornell-c1").build()
    w = h = 64
    _, ref_acc, ref_g, _ = _render(scene, w, h, 8, gpu.ENGINE_MEGAKERNEL)
    _, acc, g, st = _render(scene, w, h, 8, gpu.ENGINE_WAVEFRONT, cfg=cfg)
    other = dict(cfg, path_slots=333, block=64, stack_depth=24, use_graph=1, graph_iterations=5)
    _, acc2, _, _ = _render(scene, w, h, 8, gpu.ENGINE_WAVEFRONT, cfg=other)
    assert acc.tobytes() == acc2.tobytes()
    agree = radiance_agreement(acc[..., :3], ref_acc[..., :3])
    print(f"{cfg}: {100 * agree:.4f}% within 0.1%, {100 * bitwise_fraction(acc, ref_acc):.2f}% bitwise, {st.iterations} iterations")
    assert agree >= 0.999
    assert np.allclose(acc[..., 3], ref_acc[..., 3], rtol=1e-5, atol=1e-6)
    # The first sample's G-buffer is unchanged.
    assert g[0].tobytes() == ref_g[0].tobytes() and g[1].tobytes() == ref_g[1].tobytes()


def test_sample_split_accumulates_progressively():
    scene = host.Scene.builtin("materials").build()
    _, a, _, _ = _render(scene, 48, 32, 8, gpu.ENGINE_MEGAKERNEL, calls=2)
    _, b, _, _ = _render(scene, 48, 32, 8, gpu.ENGINE_WAVEFRONT, calls=2, cfg=dict(sample_split=4))
    assert radiance_agreement(b[..., :3], a[..., :3]) >= 0.999


def test_wavefront_matches_oracle_on_the_field_scene():
    scene = host.Scene.field(96, seed=3, env_width=512).build()
    w, h, spp = 96, 64, 4
    img, _, _, st = _render(scene, w, h, spp, gpu.ENGINE_WAVEFRONT, depth=16)
    rs = RenderSettings.make(w, h, max_depth=16, rr_depth=3, seed=1)
    ref = O.render(scene.desc(), scene.camera_view(w / h), rs, spp)
    agree = radiance_agreement(img, ref)
    print(f"field: {100 * agree:.3f}% within 0.1%, {100 * bitwise_fraction(img, ref):.3f}% bitwise, {st.iterations} iterations")
    assert agree >= 0.999


def test_progressive_calls_accumulate_like_the_megakernel():
    scene = host.Scene.builtin("cornell-c1").build()
    _, a, _, _ = _render(scene, 48, 48, 8, gpu.ENGINE_MEGAKERNEL, calls=2)
    _, b, _, _ = _render(scene, 48, 48, 8, gpu.ENGINE_WAVEFRONT, calls=2)
    assert a.tobytes() == b.tobytes()


def test_counters_and_stats():
    scene = host.Scene.builtin("cornell-c1").build()
    w = h = 64
    rs = RenderSettings.make(w, h, seed=1)
    r = uploaded(scene, w, h, engine=gpu.ENGINE_WAVEFRONT)
    r.set_counting(True)
    r.render(scene.camera_view(1.0), rs, 2)
    st = r.stats()
    assert st.samples == w * h * 2 and st.node_visits > 0 and st.tri_tests > 0 and st.iterations > 0


def test_autotune_keeps_the_image():
    scene = host.Scene.builtin("cornell-c1").build()
    w = h = 64
    rs = RenderSettings.make(w, h, seed=1)
    cam = scene.camera_view(1.0)
    r = uploaded(scene, w, h, engine=gpu.ENGINE_WAVEFRONT)
    cfg = r.autotune(cam, rs, 1)
    print("autotuned:", cfg.as_dict())
    r.reset()
    r.render(cam, rs, 4)
    acc = gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_ACCUM), w * h * 16, np.float32, (h, w, 4))
    _, ref, _, _ = _render(scene, w, h, 4, gpu.ENGINE_MEGAKERNEL)
    assert acc.tobytes() == ref.tobytes()


@pytest.mark.parametrize("name,w,h,spp,cfg", [
    ("materials", 96, 54, 8, dict(finish_percent=10)),
    ("materials", 96, 54, 8, dict(finish_percent=100)),
    ("cornell-c1", 64, 64, 8, dict(finish_percent=100, path_slots=1000)),
    ("glass", 80, 60, 8, dict(finish_percent=40, path_slots=3000, use_graph=0)),
    ("cornell-c1", 64, 64, 16, dict(finish_percent=30, sample_split=4, path_slots=2048)),
    ("materials", 64, 48, 8, dict(finish_percent=20, sample_split=8, path_slots=5000)),
])
def test_finishing_the_last_paths_one_per_thread_keeps_the_image(name, w, h, spp, cfg):
    """The paths left once every work item is handed out, finished one per thread (finish_percent),
    give the same accumulation and G-buffer as the engine without it, bit for bit, and so the per-pixel
    kernel's image where the split keeps the in-order sum. With 100 they are finished after the first
    bounce, with new items' first samples (sample 0, depth 0) among them when the slots are fewer than
    the pixels."""
    scene = host.Scene.builtin(name).build()
    _, off_acc, off_g, off_st = _render(scene, w, h, spp, gpu.ENGINE_WAVEFRONT, cfg=dict(cfg, finish_percent=0))
    _, on_acc, on_g, on_st = _render(scene, w, h, spp, gpu.ENGINE_WAVEFRONT, cfg=cfg)
    print(f"{name} {cfg}: finished one per thread {on_st.finish_paths} paths, {on_st.rays_finish} rays; "
          f"iterations {off_st.iterations} -> {on_st.iterations}")
    assert off_st.finish_paths == 0 and off_st.rays_finish == 0
    assert on_st.finish_paths > 0 and on_st.rays_finish > 0
    assert on_acc.tobytes() == off_acc.tobytes()
    assert on_g[0].tobytes() == off_g[0].tobytes() and on_g[1].tobytes() == off_g[1].tobytes()
    if cfg.get("sample_split", 1) in (1, spp):
        _, ref_acc, ref_g, _ = _render(scene, w, h, spp, gpu.ENGINE_MEGAKERNEL)
        assert on_acc.tobytes() == ref_acc.tobytes()
        assert on_g[0].tobytes() == ref_g[0].tobytes() and on_g[1].tobytes() == ref_g[1].tobytes()
