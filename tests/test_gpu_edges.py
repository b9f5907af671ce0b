"""Edge cases of the device path against the oracle: empty ray buffers, a scene without triangles (the
builder's single root with two empty children, as bvh.cpp:172-189 makes it), a single triangle, the
alpha-masked quad of the reference's BVH test (test_bvh.cpp:150-170: the masked quad is skipped and the
opaque one behind it is hit at t = 5), 1x1 and ragged image sizes, one-bounce paths, and the denoiser on
images smaller than its 16x16 tiles."""
import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd import gpu, host
from photon_amd.types import ALPHA_MASK, RAY_DTYPE, Material, RenderSettings, SvgfSettings

from gpu_util import radiance_agreement, trace_on_gpu, uploaded

pytestmark = pytest.mark.gpu


def _rays(o, d, tmax=1e30):
    r = np.zeros(len(o), RAY_DTYPE)
    r["o"], r["d"], r["tmax"] = o, d, tmax
    return r


def test_empty_ray_buffers():
    scene = host.Scene.builtin("cornell").build()
    r = uploaded(scene)
    for anyhit in (False, True):
        hits = trace_on_gpu(r, np.zeros(0, RAY_DTYPE), anyhit=anyhit)
        assert len(hits) == 0
    r.close()


def test_scene_without_triangles():
    s = host.Scene()
    s.set_camera((0, 0, 5), (0, 0, 0))
    s.build()
    assert s.num_tris == 0 and len(s.nodes()) == 1
    r = uploaded(s, 16, 12)
    rays = _rays([[0, 0, 0], [1, 2, 3]], [[0, 0, -1], [1, 0, 0]])
    for anyhit in (False, True):
        assert trace_on_gpu(r, rays, anyhit=anyhit)["tri"].tolist() == [-1, -1]
    rs = RenderSettings.make(16, 12, max_depth=4, seed=2)
    cam = s.camera_view(16 / 12)
    for engine in (gpu.ENGINE_MEGAKERNEL, gpu.ENGINE_WAVEFRONT):
        r.set_engine(engine)
        r.reset()
        r.render(cam, rs, 2)
        assert np.array_equal(r.download_image(), O.render(s.desc(), cam, rs, 2))
    r.close()


def test_single_triangle_and_alpha_quad():
    s = host.Scene()
    s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.int32), s.add_material(Material.default()))
    s.build()
    r = uploaded(s)
    rays = _rays([[0.2, 0.2, 1], [5, 5, 1], [0.2, 0.2, 1]], [[0, 0, -1], [0, 0, -1], [0, 0, 1]])
    for anyhit in (False, True):
        got = trace_on_gpu(r, rays, anyhit=anyhit)
        want = O.trace(s.desc(), rays, anyhit=anyhit)
        assert got["tri"].tolist() == want["tri"].tolist() == [0, -1, -1]
        assert got["t"][0] == want["t"][0]
    r.close()

    q = host.Scene()
    masked = Material.default()
    masked.base_color[:] = (1, 1, 1, 0)
    masked.alpha_mode = ALPHA_MASK
    m0 = q.add_material(masked)
    m1 = q.add_material(Material.default())
    quad = lambda z: np.array([[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]], np.float32)
    q.add_mesh(quad(1), np.array([0, 1, 2, 0, 2, 3], np.int32), m0)
    q.add_mesh(quad(0), np.array([0, 1, 2, 0, 2, 3], np.int32), m1)
    q.build()
    r = uploaded(q)
    rays = _rays([[0.1, 0.2, 5], [-0.7, 0.4, 5]], [[0, 0, -1], [0, 0, -1]])
    for anyhit in (False, True):
        got = trace_on_gpu(r, rays, anyhit=anyhit)
        want = O.trace(q.desc(), rays, anyhit=anyhit)
        assert got["tri"].tolist() == want["tri"].tolist()
        assert np.all(got["tri"] >= 0) and np.allclose(got["t"], 5.0, atol=1e-4)
        assert np.array_equal(got["t"], want["t"])
    r.close()


@pytest.mark.parametrize("w,h,spp,depth", [(1, 1, 1, 1), (1, 1, 4, 8), (17, 13, 3, 1), (17, 13, 6, 8), (3, 29, 2, 16)])
def test_tiny_and_ragged_images(w, h, spp, depth):
    scene = host.Scene.builtin("cornell-c1").build()
    rs = RenderSettings.make(w, h, max_depth=depth, rr_depth=3, seed=4)
    cam = scene.camera_view(w / h)
    ref = O.render(scene.desc(), cam, rs, spp)
    accs = []
    for engine, cfg in ((gpu.ENGINE_MEGAKERNEL, None), (gpu.ENGINE_WAVEFRONT, None), (gpu.ENGINE_WAVEFRONT, dict(sample_split=spp, path_slots=7))):
        r = uploaded(scene, w, h, engine=engine)
        if cfg:
            r.set_wavefront_config(**cfg)
        r.render(cam, rs, spp)
        img = r.download_image()
        accs.append(gpu.copy_device_to_numpy(r.buffer_ptr(gpu.BUF_ACCUM), w * h * 16, np.float32, (h, w, 4)))
        r.close()
        assert radiance_agreement(img, ref) >= 0.999
    # In-order accumulation whatever the engine, also one sample per sub-stream on fewer slots than pixels.
    assert accs[0].tobytes() == accs[1].tobytes() == accs[2].tobytes()


@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (17, 13), (33, 5)])
def test_svgf_on_images_smaller_than_a_tile(w, h):
    scene = host.Scene.builtin("cornell-c1").build()
    rs = RenderSettings.make(w, h, seed=9)
    cam = scene.camera_view(w / h)
    sv = SvgfSettings.make()
    g = O.render_gbuffer(scene.desc(), cam, rs, 0)
    f = gpu.Svgf()
    f.resize(w, h)
    out = gpu.DeviceArray(w * h * 16)
    ora = O.SvgfOracle(w, h)
    for frame in range(2):
        dev = [gpu.DeviceArray.from_numpy(p) for p in g]
        f.run(*[d.ptr for d in dev], cam, sv, out.ptr)
        a = out.numpy(np.float32, (h, w, 4))
        b = ora.run(*g, cam, sv)
        d = np.abs(a - b)
        assert np.all((d <= 1e-3 * np.abs(b)) | (d <= 1e-6)), f"frame {frame}"
    c = O.compact_gbuffer(*g)
    gb = gpu.CompactGBuffer.from_numpy(c, w, h)
    ora_c = O.SvgfOracle(w, h)
    f.run_compact(gb, cam, cam, sv, out.ptr)
    a = out.numpy(np.float32, (h, w, 4))
    b = ora_c.run_compact(c, cam, cam, sv)
    d = np.abs(a - b)
    assert np.all((d <= 1e-3 * np.abs(b)) | (d <= 1e-6))


def test_serving_a_one_pixel_job():
    scene = host.Scene.builtin("cornell-c1").build()
    srv = gpu.Server(path_slots=1024, min_items=1)
    sid = srv.add_scene(scene)
    j = srv.submit(sid, 1, 1, 5, scene.camera_view(1.0), max_depth=8, seed=3)
    srv.run_until_idle()
    got = srv.result(j)
    srv.close()
    r = uploaded(scene, 1, 1, engine=gpu.ENGINE_WAVEFRONT)
    r.render(scene.camera_view(1.0), RenderSettings.make(1, 1, max_depth=8, rr_depth=3, seed=3), 5)
    assert got.tobytes() == r.download_image().tobytes()
    r.close()
