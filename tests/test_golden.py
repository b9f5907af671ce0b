"""The host producers and the oracle against the golden vectors written by the reference's own CPU
build (tests/golden/make_golden.py): BVH nodes, BVH order, triangle arrays and light tables must be
bit-identical, and so must the oracle's hits and per-pixel radiance. A second group repeats the
comparison live against oracle/_ref where that build exists."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from photon_amd.types import CameraView, RenderSettings

from golden_scenes import GOLDEN, RENDERS, make_scene, soup_scene

BVH_SCENES = ["cornell-c1", "cornell", "furnace", "soup", "field-small", "showcase"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def built(name):
    s = make_scene(name)
    s.build(threads=4)
    return s


@pytest.mark.parametrize("name", BVH_SCENES)
def test_host_producers_match_reference_fixtures(name):
    f = np.load(os.path.join(GOLDEN, f"bvh_{name}.npz"))
    s = built(name)
    a = O.desc_arrays(s.desc())
    a["order"] = s.order()
    a["nodes"] = a["nodes"].view(np.uint8)
    want = dict(d.split(":") for d in f["digests"])
    got = {k: sha(v) for k, v in a.items()}
    mismatched = [k for k in want if want[k] != got.get(k)]
    assert not mismatched, mismatched
    for k in f.files:
        if k != "digests":
            assert np.array_equal(f[k], a[k]), k


def test_oracle_hits_match_reference_fixture():
    f = np.load(os.path.join(GOLDEN, "hits_soup.npz"))
    s = soup_scene().build(threads=4)
    assert O.trace(s.desc(), f["rays"]).tobytes() == f["closest"].tobytes()
    assert O.trace(s.desc(), f["rays"], anyhit=True).tobytes() == f["any"].tobytes()
    hit = f["closest"]["tri"] >= 0
    assert 0.1 < hit.mean() < 0.95  # the fixture exercises both hits and misses


@pytest.mark.parametrize("name,w,h,spp,depth,seed", RENDERS)
def test_oracle_render_matches_reference_fixture(name, w, h, spp, depth, seed):
    f = np.load(os.path.join(GOLDEN, f"render_{name}.npz"))
    s = built(name)
    cam = CameraView.from_buffer_copy(f["camera"].tobytes())
    rs = RenderSettings.make(w, h, max_depth=depth, rr_depth=3, seed=seed)
    img = O.render(s.desc(), cam, rs, spp, threads=8)
    assert img.tobytes() == f["image"].tobytes()
    g = np.stack(O.render_gbuffer(s.desc(), cam, rs, 0, threads=8))
    assert g.tobytes() == f["gbuffer"].tobytes()


ref_only = pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present")


@ref_only
@pytest.mark.parametrize("name", ["cornell", "materials", "glass", "furnace"])
def test_builtin_scenes_match_the_reference_builtins(name):
    mine = make_scene(name).build(threads=4)
    theirs = O.RefScene(builtin=name).build(threads=4)
    a, b = O.desc_arrays(mine.desc()), O.desc_arrays(theirs.desc())
    assert [k for k in a if a[k].tobytes() != b[k].tobytes()] == []
    assert np.array_equal(mine.order(), theirs.order())


@ref_only
def test_threaded_build_matches_reference_on_a_large_soup():
    from photon_amd import host
    s = host.Scene.spheres(64, alpha_fraction=0.1, seed=9).build(threads=8)
    r = O.RefScene(raw=s.raw()).build(threads=8)
    a, b = O.desc_arrays(s.desc()), O.desc_arrays(r.desc())
    assert [k for k in a if a[k].tobytes() != b[k].tobytes()] == []
    st, rt = s.stats(), r.stats()
    assert (st.nodes, st.leaves, st.max_depth, st.sah_cost) == (rt.nodes, rt.leaves, rt.max_depth, rt.sah_cost)


@ref_only
def test_oracle_matches_reference_live():
    from photon_amd import host
    s = host.Scene.spheres(32, alpha_fraction=0.2, seed=4).build()
    r = O.RefScene(raw=s.raw()).build()
    cam = s.camera_view(1.0)
    rays = host.camera_rays(cam, 128, 128)
    assert O.trace(s.desc(), rays).tobytes() == r.trace(rays).tobytes()
    sh = s.shadow_rays(8192, seed=3)
    assert O.trace(s.desc(), sh, anyhit=True).tobytes() == r.trace(sh, anyhit=True).tobytes()
    f = host.Scene.field(24, seed=2, env_width=256).build()
    rf = O.RefScene(raw=f.raw()).build()
    rs = RenderSettings.make(40, 30, max_depth=16, seed=5, clamp=3.0)
    cam = f.camera_view(40 / 30)
    assert O.render(f.desc(), cam, rs, 3).tobytes() == rf.render(cam, rs, 3).tobytes()
