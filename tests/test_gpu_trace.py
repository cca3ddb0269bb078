"""Closest-hit parity: HIP traversal vs the oracle's restatement of
SBVH::traverse (acceleration/sbvh.cpp:478-510) on identical rays. IDs and t
must be bit-exact; count mode must reproduce the reference's counters."""
import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu


def _setup(device, sc):
    bvh = api.Bvh(sc.verts)
    device.set_scene(bvh, sc.normals, sc.mat_ids)
    orc = O.Scene(sc.verts, sc.normals, sc.mat_ids)
    return bvh, orc


def _check_closest(device, orc, rays):
    t_g, p_g = device.trace_host(rays)
    t_c, p_c, _, _ = orc.trace(rays)
    mism = np.nonzero(p_g != p_c)[0]
    assert len(mism) == 0, f"{len(mism)} id mismatches, first {mism[:5]} gpu {p_g[mism[:5]]} cpu {p_c[mism[:5]]}"
    hit = p_c != L.SPT_MISS
    assert np.array_equal(t_g[hit].view(np.uint32), t_c[hit].view(np.uint32))
    assert np.all(np.isinf(t_g[~hit]))
    return hit.mean()


def test_trace_cornell_incoherent(device):
    sc = scenes.cornell_icosphere()
    _, orc = _setup(device, sc)
    rays = scenes.incoherent_rays(sc.verts, 200_000, seed=11)
    frac = _check_closest(device, orc, rays)
    assert frac > 0.3


def test_trace_count_mode_matches_reference_counters(device):
    sc = scenes.cornell_icosphere(subdiv=3)
    _, orc = _setup(device, sc)
    rays = scenes.incoherent_rays(sc.verts, 50_000, seed=12)
    t_g, p_g, tests_g, isects_g = device.trace_host(rays, L.SPT_TRACE_COUNT)
    t_c, p_c, tests_c, isects_c = orc.trace(rays)
    assert np.array_equal(p_g, p_c)
    assert np.array_equal(tests_g, tests_c)
    assert np.array_equal(isects_g, isects_c)


def test_trace_displaced_mesh_subset(device):
    m = scenes.displaced_mesh(rings=128, segments=256)
    _, orc = _setup(device, m)
    rays = scenes.incoherent_rays(m.verts, 100_000, seed=13)
    _check_closest(device, orc, rays)


def test_trace_unpruned_equals_pruned(device):
    sc = scenes.cornell_icosphere(subdiv=3)
    _setup(device, sc)
    rays = scenes.incoherent_rays(sc.verts, 50_000, seed=14)
    t1, p1 = device.trace_host(rays, L.SPT_TRACE_CLOSEST)
    t2, p2 = device.trace_host(rays, L.SPT_TRACE_UNPRUNED)
    assert np.array_equal(p1, p2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))


def test_trace_edge_cases(device):
    """Axis-parallel directions (+-0 components -> +-inf reciprocals and NaN
    slab products), origins on box faces, rays behind everything, n = 0."""
    sc = scenes.cornell_icosphere(subdiv=2)
    _, orc = _setup(device, sc)
    rng = np.random.default_rng(5)
    rays = []
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (-0.0, 1, 0), (0.0, -0.0, -1.0), (0.6, 0.8, 0.0), (0, 0.6, -0.8)]
    origins = [(0, 0, 0), (-1, 0, 0), (1, 1, 1), (0.25, 0.999, 0.25), (0, -1, 0), (0, 0, 5), (3, 3, 3),
               (-1, -1, -1), (0.5, 0.5, -1.0)]
    for o in origins:
        for d in dirs:
            rays.append(list(o) + [0.0] + list(d) + [0.0])
    for _ in range(2000):  # origins exactly on the box planes, axis-parallel dirs
        o = rng.choice([-1.0, 1.0, 0.0, 0.5], 3)
        d = np.zeros(3)
        d[rng.integers(3)] = rng.choice([-1.0, 1.0])
        rays.append(list(o) + [0.0] + list(d) + [0.0])
    rays = np.array(rays, np.float32)
    _check_closest(device, orc, rays)
    t_g, p_g, te_g, is_g = device.trace_host(rays, L.SPT_TRACE_COUNT)
    _, _, te_c, is_c = orc.trace(rays)
    assert np.array_equal(te_g, te_c) and np.array_equal(is_g, is_c)
    empty = np.zeros((0, 8), np.float32)
    t0, p0 = device.trace_host(empty)
    assert len(t0) == 0


def test_trace_non_finite_rays_miss(device):
    """NaN or infinite ray components: every t comparison of the reference
    fails (no hit); the persistent traversal rejects such rays at the root
    instead of visiting every box, and agrees with the oracle's result."""
    sc = scenes.displaced_mesh(64, 128)
    _, orc = _setup(device, sc)
    nan, inf = np.float32("nan"), np.float32("inf")
    rays = np.array([[0, 0, 3, 0, 0, 0, -1, 0], [0, 0, 3, 0, nan, 0, -1, 0], [nan, 0, 3, 0, 0, 0, -1, 0],
                     [0, 0, 3, 0, inf, 0, -1, 0], [0, 0, inf, 0, 0, 0, -1, 0], [0, 0, 3, 0, 0, nan, nan, 0]],
                    np.float32)
    rays = np.tile(rays, (4096, 1))
    t_g, p_g = device.trace_host(rays)
    t_c, p_c, _, _ = orc.trace(rays[:6])
    assert p_g[0] != L.SPT_MISS and p_c[0] == p_g[0]
    assert np.all(p_g.reshape(-1, 6)[:, 1:] == L.SPT_MISS) and np.all(p_c[1:] == L.SPT_MISS)


def test_trace_root_leaf_scene(device):
    """A scene small enough that the root is a leaf (no split beats N)."""
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)
    bvh = api.Bvh(tri)
    assert bvh.info["n_pair_nodes"] == 0
    device.set_scene(bvh)
    orc = O.Scene(tri)
    rng = np.random.default_rng(1)
    rays = np.zeros((4096, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-0.5, 1.5, (4096, 3))
    d = rng.normal(size=(4096, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    _check_closest(device, orc, rays)


def test_raycasters_match_reference(device):
    sc = scenes.cornell_icosphere(subdiv=4, width=96, height=72)
    _, orc = _setup(device, sc)
    cam = sc.camera
    n_g = device.raycast(cam, L.SPT_RAYCAST_NORMAL)
    n_c = orc.raycast(cam, 0)
    assert np.array_equal(n_g.view(np.uint32), n_c.view(np.uint32))
    d_g = device.raycast(cam, L.SPT_RAYCAST_DEPTH)
    d_c = orc.raycast(cam, 1)
    assert np.array_equal(d_g.view(np.uint32), d_c.view(np.uint32))
    c_g = device.raycast(cam, L.SPT_RAYCAST_COUNT)
    c_c = orc.raycast(cam, 2)
    assert np.array_equal(c_g, c_c)


def test_raycast_smooth_normals_orthographic(device):
    m = scenes.displaced_mesh(rings=64, segments=128)
    _, orc = _setup(device, m)
    cam = api.orthographic(64, 64, -1.3, 1.3, -1.3, 1.3, position=(0.3, 0.2, 4))
    n_g = device.raycast(cam, L.SPT_RAYCAST_NORMAL)
    n_c = orc.raycast(cam, 0)
    assert np.array_equal(n_g.view(np.uint32), n_c.view(np.uint32))
