"""The device BVH build's own code (photon_amd/csrc/hip/bvh_build.hip: host driver, large-node passes,
subtree kernel, flattening) executed on the CPU under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/hipcpu: every thread a fiber, wave intrinsics and barriers honoured, global memory as exact-size
heap allocations). Every access beyond a buffer, every barrier that not all threads reach and every
shuffle from a finished lane stops the run; the result must be the host builder's, bit for bit (nodes,
triangle order, statistics), as on the device. The cases cross the large-node threshold (2,048
triangles) and the subtree bound, and include coincident centroids and a flat layer."""
import os
import sys

import numpy as np
import pytest

from photon_amd import gpu, host
from photon_amd.types import Material

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "hipcpu"))

import hipcpu_build  # noqa: E402
import hipcpu_run  # noqa: E402


@pytest.fixture(scope="module")
def exe():
    return hipcpu_build.build()


def _soup(tris):
    s = host.Scene()
    m = s.add_material(Material.default())
    p = np.asarray(tris, np.float32).reshape(-1, 3)
    s.add_mesh(p, np.arange(len(p), dtype=np.int32), m)
    return s


def _check(exe, scene):
    scene.build()
    r = scene.raw()
    p = np.ctypeslib.as_array(r.positions, (r.num_vertices * 3,)).reshape(-1, 3).copy() if r.num_vertices else np.zeros((0, 3), np.float32)
    i = np.ctypeslib.as_array(r.indices, (r.num_tris * 3,)).reshape(-1, 3).copy() if r.num_tris else np.zeros((0, 3), np.int32)
    nodes, order, stats, _ = hipcpu_run.run(exe, p, i, timeout=600)
    st = gpu.BvhBuildStats.from_buffer_copy(stats)
    hs = scene.stats()
    assert nodes.tobytes() == scene.nodes().tobytes()
    if len(i):
        assert np.array_equal(order, scene.order())
    assert (st.nodes, st.leaves, st.max_depth) == (hs.nodes, hs.leaves, hs.max_depth)
    assert np.float32(st.sah_cost).tobytes() == np.float32(hs.sah_cost).tobytes()


def test_builtin_scene(exe):
    _check(exe, host.Scene.builtin("cornell-c1"))


def test_random_soups_across_the_thresholds(exe):
    g = np.random.default_rng(11)
    for n in (1, 2, 3, 9, 2047, 2048, 2049, 6000):
        c = g.uniform(-10, 10, (n, 1, 3))
        _check(exe, _soup(c + g.normal(0, 0.3, (n, 3, 3))))


def test_degenerate_layouts(exe):
    g = np.random.default_rng(12)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    _check(exe, _soup(np.repeat(tri[None], 3000, 0)))  # coincident centroids: forced middle splits
    pts = np.repeat(g.uniform(-1, 1, (50, 1, 3)), 60, 0) + g.normal(0, 1e-3, (3000, 3, 3))
    pts[:1500, :, 1] = 0.0  # duplicate clusters and a flat layer
    _check(exe, _soup(pts))
