"""The device BVH build (ph_bvh_build) against the host builder, which the golden fixtures pin to the
reference's build_bvh (src/bvh/bvh.cpp:51-255): the node array, the triangle order and the statistics
must be bit-identical, on scenes that take the large-node passes (more than 2048 triangles per node)
and the one-block subtrees, and on the cases that stress the rules: coincident centroids (forced
middle splits), duplicates, +0/-0 coordinates (the last-occurrence rule of the node-box folds), one
triangle and none."""
import numpy as np
import pytest

from photon_amd import gpu, host

pytestmark = pytest.mark.gpu


def _mesh(scene):
    r = scene.raw()
    p = np.ctypeslib.as_array(r.positions, (r.num_vertices * 3,)).reshape(-1, 3).copy() if r.num_vertices else np.zeros((0, 3), np.float32)
    i = np.ctypeslib.as_array(r.indices, (r.num_tris * 3,)).reshape(-1, 3).copy() if r.num_tris else np.zeros((0, 3), np.int32)
    return p, i


def _compare(scene):
    scene.build()
    p, i = _mesh(scene)
    nodes, order, st = gpu.bvh_build(p, i)
    want = scene.nodes()
    hs = scene.stats()
    print(f"{len(i)} triangles: {len(want)} nodes, device build {st.build_ms:.1f} ms, host {hs.build_ms:.1f} ms")
    assert len(nodes) == len(want)
    assert nodes.tobytes() == want.tobytes()
    if len(i):
        assert np.array_equal(order, scene.order())
    assert (st.nodes, st.leaves, st.max_depth) == (hs.nodes, hs.leaves, hs.max_depth)
    assert np.float32(st.sah_cost).tobytes() == np.float32(hs.sah_cost).tobytes()
    d = scene.desc()
    if len(i):
        assert np.array(st.bounds_min, np.float32).tobytes() == np.array(d.bounds_min, np.float32).tobytes()
        assert np.array(st.bounds_max, np.float32).tobytes() == np.array(d.bounds_max, np.float32).tobytes()


@pytest.mark.parametrize("name", ["cornell-c1", "materials", "glass", "furnace"])
def test_builtin_scenes(name):
    _compare(host.Scene.builtin(name))


def test_sphere_soup_and_field():
    _compare(host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1))
    _compare(host.Scene.field(64, seed=3, env_width=64))


@pytest.mark.skip(reason="a 1M-triangle device build with an earlier subtree kernel made a GPU memory fault whose cause was not "
                         "found (a CPU emulation of that build keeps every index in bounds on this scene); ph_bvh_build refuses "
                         "inputs above PH_BVH_BUILD_MAX_TRIS = 262144 triangles and this case is not run until a cause is shown")
def test_million_triangle_field():
    _compare(host.Scene.field(1024, seed=1, env_width=64))


def _soup(tris, mat=0):
    s = host.Scene()
    m = s.add_material(__import__("photon_amd.types", fromlist=["Material"]).Material.default())
    p = np.asarray(tris, np.float32).reshape(-1, 3)
    s.add_mesh(p, np.arange(len(p), dtype=np.int32), m)
    return s


def test_random_soups_and_degenerate_cases():
    g = np.random.default_rng(11)
    # Random soups across the large-node threshold.
    for n in (1, 2, 3, 9, 100, 2047, 2048, 2049, 5000, 40000):
        c = g.uniform(-10, 10, (n, 1, 3))
        _compare(_soup(c + g.normal(0, 0.3, (n, 3, 3))))
    # Coincident centroids: identical triangles, split in the middle.
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    for n in (9, 3000):
        _compare(_soup(np.repeat(tri[None], n, 0)))
    # Clusters of duplicates and a flat layer: zero extents on one axis.
    pts = np.repeat(g.uniform(-1, 1, (50, 1, 3)), 60, 0) + g.normal(0, 1e-3, (3000, 3, 3))
    pts[:1500, :, 1] = 0.0
    _compare(_soup(pts))


def test_signed_zero_coordinates():
    """Coordinates of +0 and -0 mixed in one node: the node boxes keep the sign of the last occurrence,
    as the host's sequential fmin/fmax fold does."""
    g = np.random.default_rng(5)
    for n in (40, 6000):
        t = g.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
        z = g.uniform(size=(n, 3, 3)) < 0.3
        t[z] = 0.0
        t[z & (g.uniform(size=(n, 3, 3)) < 0.5)] = -0.0
        _compare(_soup(t))


def test_empty_scene():
    s = host.Scene()
    s.build()
    nodes, order, st = gpu.bvh_build(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    assert nodes.tobytes() == s.nodes().tobytes() and len(order) == 0 and st.nodes == 1


def test_rejects_bad_input():
    with pytest.raises(RuntimeError, match="out of range"):
        gpu.bvh_build(np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.int32))
    with pytest.raises(RuntimeError, match="bins"):
        gpu.bvh_build(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.int32), bins=65)
