"""Sphere primitives on the device (Sphere::intersect, primitives/sphere.cpp:6-38,
normal = normalize(hit - centre)): closest-hit ids and t bit-exact against the
oracle on a scene of triangles and spheres interleaved in scene order, the
reference's counters, the three ray casters, and the reference's own furnace
scene (samples/furnace_test.lua: two spheres) with each of its six materials,
per path within 1e-5 of the oracle."""
import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _mixed(seed=5, n_spheres=60):
    rng = np.random.default_rng(seed)
    box = scenes.cornell_icosphere(subdiv=3)
    sp = np.concatenate([rng.uniform(-0.8, 0.8, (n_spheres, 3)), rng.uniform(0.02, 0.3, (n_spheres, 1))], 1)
    parts = [scenes.concat([box], []), scenes.spheres(sp[: n_spheres // 2], np.arange(n_spheres // 2) % 5),
             scenes.icosphere(2, (0.3, 0.2, -0.1), 0.2, mat=2, smooth=True),
             scenes.spheres(sp[n_spheres // 2:], 4)]
    sc = scenes.concat(parts, box.materials)
    sc.camera, sc.integrator = box.camera, box.integrator
    return sc


def _setup(device, sc):
    device.set_scene(api.Bvh(sc.verts, kinds=sc.kinds), sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    return O.Scene(sc.verts, sc.normals, sc.mat_ids, kinds=sc.kinds, materials=sc.materials)


def _random_rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(-0.95, 0.95, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    return r


def _rel_ok(a, b):
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(a), np.abs(b)) + 1e-30, axis=1)


def test_trace_spheres_and_triangles(device):
    sc = _mixed()
    orc = _setup(device, sc)
    rays = _random_rays(200_000, 3)
    t_g, p_g = device.trace_host(rays)
    t_c, p_c, te_c, is_c = orc.trace(rays)
    assert np.array_equal(p_g, p_c), f"{np.sum(p_g != p_c)} id mismatches"
    hit = p_c != L.SPT_MISS
    assert np.array_equal(t_g[hit].view(np.uint32), t_c[hit].view(np.uint32))
    on_sphere = hit & (sc.kinds[np.minimum(p_c, sc.n_tris - 1)] == L.SPT_PRIM_SPHERE)
    assert on_sphere.mean() > 0.1  # spheres are actually hit
    _, p2, te_g, is_g = device.trace_host(rays[:20_000], L.SPT_TRACE_COUNT)
    assert np.array_equal(p2, p_c[:20_000])
    assert np.array_equal(te_g, te_c[:20_000]) and np.array_equal(is_g, is_c[:20_000])
    t3, p3 = device.trace_host(rays, L.SPT_TRACE_UNPRUNED)
    assert np.array_equal(p3, p_g) and np.array_equal(t3.view(np.uint32), t_g.view(np.uint32))


def test_raycasters_with_spheres(device):
    sc = _mixed()
    orc = _setup(device, sc)
    cam = api.pinhole(96, 72, position=(0, 0, 3.6), look_at=(0, 0, 0), fov=40.0)
    for kind in (L.SPT_RAYCAST_NORMAL, L.SPT_RAYCAST_DEPTH, L.SPT_RAYCAST_COUNT):
        g, c = device.raycast(cam, kind), orc.raycast(cam, kind)
        assert np.array_equal(np.ascontiguousarray(g).view(np.uint32), np.ascontiguousarray(c).view(np.uint32)), kind


def test_mixed_scene_path_replay(device):
    sc = _mixed()
    sc.integrator = api.Integrator(spp=16, pixel_sampler="jittered", termination="russian-roulette", path_length=3,
                                   max_depth=12, seed=4)
    sc.camera = api.pinhole(48, 40, position=(0, 0, 3.6), look_at=(0, 0, 0), fov=40.0)
    orc = _setup(device, sc)
    _, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    ok = _rel_ok(paths, paths_c)
    assert ok.all(), f"{np.sum(~ok)} paths differ"


@pytest.mark.parametrize("material", sorted(scenes.reference_furnace_materials()))
def test_reference_furnace_scene(device, material):
    """samples/furnace_test.lua with each of its materials, rendered as the
    script describes it (two spheres, orthographic view, RR from length 3)."""
    sc = scenes.reference_furnace(material, width=64, height=64, spp=16, seed=9)
    orc = _setup(device, sc)
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    ok = _rel_ok(paths, paths_c)
    print(f"furnace {material}: {ok.mean() * 100:.4f}% of {len(ok)} paths within {RTOL}")
    assert ok.all(), f"{np.sum(~ok)} paths differ"
    if material == "lambertian":  # every path returns 0.5 (emitter) or 0.5 * 0.999 (sphere)
        v = paths[:, 0]
        assert np.all(np.isclose(v, 0.5, rtol=1e-6) | np.isclose(v, 0.5 * 0.999, rtol=2e-6))
        assert np.mean(np.isclose(v, 0.5 * 0.999, rtol=2e-6)) > 0.4
    # energy: no material of the furnace reflects more than it receives
    assert img.mean() <= 0.5 * (1 + 1e-3)


def test_batch_with_sphere_scenes_equals_single_renders(device):
    scs = [scenes.reference_furnace(m, width=32, height=32, spp=4, seed=i)
           for i, m in enumerate(["coating", "lambertian", "smooth_dielectric"])]
    scs.append(scenes.cornell_icosphere(subdiv=2, width=24, height=24, spp=4))
    b = api.Batch(device, scs, per_path=True)
    b.render()
    for i, sc in enumerate(scs):
        _setup(device, sc)
        _, ps = device.render_image(sc.camera, sc.integrator, per_path=True)
        assert np.array_equal(b.path_radiance(i).view(np.uint32), ps.view(np.uint32)), i
    b.close()
