"""Per-path replay parity for every BSDF, BSDF sampler and pixel sampler of
the reference that the device implements, and for the FP64 film.

Each case renders the box scene with the object's material (or the pixel
sampler) swapped and compares every path's radiance with the oracle's
restatement of PathTracer::traceRay (integrators/pathtracer.cpp:12-57) on the
same counter-based draws: 100% of paths within 1e-5 relative.
"""
import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel_ok(a, b):
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(a), np.abs(b)) + 1e-30, axis=1)


def _setup(device, sc):
    device.set_scene(api.Bvh(sc.verts), sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    return O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)


def _box(obj_material=None, pixel_sampler="jittered", spp=16, termination="russian-roulette", path_length=3,
         max_depth=12):
    sc = scenes.cornell_icosphere(subdiv=3, width=40, height=32, spp=spp)
    if obj_material is not None:
        sc.materials[4] = obj_material
    sc.integrator = api.Integrator(spp=spp, pixel_sampler=pixel_sampler, termination=termination,
                                   path_length=path_length, max_depth=max_depth, seed=5)
    return sc


MATERIALS = {
    # materials/cook_torrance.cpp + bsdf_samplers/cook_torrance_bsdf_sampler.cpp
    "cook_torrance": api.material([api.cook_torrance(0.3, (0.9, 0.6, 0.3))]),
    # the reference furnace's default (samples/furnace_test.lua:94-103)
    "cook_torrance_furnace_default": api.material([api.cook_torrance(0.2, (0.999, 0.999, 0.999))]),
    # bsdf_samplers/uniform_bsdf_sampler.cpp with each BSDF that takes it
    "cook_torrance_uniform": api.material([api.cook_torrance(0.4, (0.8, 0.8, 0.8), sampler="uniform")]),
    "lambertian_uniform": api.material([api.lambertian((0.7, 0.5, 0.3), sampler="uniform")]),
    # materials/smooth_conductor.cpp with Schlick r0 < 1 (fresnel_schlick.cpp:11-18)
    "smooth_conductor": api.material([api.smooth_conductor((0.5, 0.7, 0.9))]),
    "smooth_conductor_uniform": api.material([api.smooth_conductor((0.5, 0.7, 0.9), sampler="uniform")]),
    # materials/smooth_dielectric.cpp alone (refraction through the object)
    "smooth_dielectric": api.material([api.smooth_dielectric(1.49)]),
    # the reference's 'plastic' and 'coating' stacks (samples/furnace_test.lua:105-133)
    "plastic": api.material([api.smooth_dielectric(1.49), api.lambertian((0.8, 0.5, 0.2))]),
    "coated_cook_torrance": api.material([api.smooth_dielectric(1.49), api.cook_torrance(0.2, (0.9, 0.7, 0.5))]),
    "coated_ggx": api.material([api.smooth_dielectric(1.5), api.ggx_conductor(0.3)]),
}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_material_replay_parity(device, name):
    sc = _box(MATERIALS[name])
    orc = _setup(device, sc)
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    ok = _rel_ok(paths, paths_c)
    print(f"{name}: {len(ok)} paths, {ok.mean() * 100:.4f}% within {RTOL}")
    assert ok.all(), f"{np.sum(~ok)} paths differ, first {np.nonzero(~ok)[0][:5]}"
    assert np.any(paths != 0.0)


@pytest.mark.parametrize("sampler,spp", [("regular", 16), ("regular", 10), ("uniform", 7), ("jittered", 5)])
def test_pixel_sampler_replay_parity(device, sampler, spp):
    """RegularPixelSampler's n x n stratum centres with spp > 1 (spp rounded up
    to n^2, pixel_samplers/regular_pixel_sampler.cpp:3-14) and
    UniformPixelSampler (pixel_samplers/uniform_pixel_sampler.cpp:7-10)."""
    sc = _box(pixel_sampler=sampler, spp=spp)
    orc = _setup(device, sc)
    _, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    n_expect = 40 * 32 * sc.integrator.effective_spp()
    assert paths.shape[0] == n_expect and paths_c.shape[0] == n_expect
    ok = _rel_ok(paths, paths_c)
    assert ok.all(), f"{np.sum(~ok)} paths differ"


def test_fp64_film_matches_oracle_fp64_image(device):
    """SPT_RENDER_FILM64: per-pixel FP64 sums of the FP32 path radiances. The
    device adds its own per-path values with FP64 atomics in any order, so the
    film equals the float64 sum of its per-path output up to double rounding
    (1e-12 relative). Against the oracle's FP64 image the bound is the
    per-path one (1e-5 relative): the paths agree to 1e-5, not bit for bit."""
    sc = _box(spp=25)
    orc = _setup(device, sc)
    img64, paths = device.render_image(sc.camera, sc.integrator, film64=True, per_path=True)
    _, img64_c, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), rgb64=True, per_path=True)
    assert img64.dtype == np.float64
    assert _rel_ok(paths, paths_c).all()
    # the same paths summed from the device's own per-path output in float64
    spp = sc.integrator.effective_spp()
    sums = paths.astype(np.float64).reshape(32, 40, spp, 3).sum(2) / spp
    np.testing.assert_allclose(img64, sums, rtol=1e-12, atol=1e-300)
    # the oracle's FP64 film of its own paths: within the per-path tolerance
    np.testing.assert_allclose(img64, img64_c, rtol=RTOL, atol=1e-30)
    assert np.abs(img64 - img64_c).max() > 0 or np.array_equal(paths, paths_c)
    img32, _ = device.render_image(sc.camera, sc.integrator)
    np.testing.assert_allclose(img32, img64, rtol=2e-5, atol=1e-7)


def test_rr_from_depth_zero_furnace(device):
    """Russian roulette from the first vertex (pathtracer.cpp:43-49 with
    path_length 0): in the Lambertian furnace a path survives the object with
    probability k = max(kd) and then returns 0.5 kd / k, else 0; the estimator's
    mean stays 0.5 kd."""
    kd = np.array([0.9, 0.5, 0.25], np.float32)
    sc = scenes.furnace(albedo=tuple(kd), spp=16)
    sc.integrator = api.Integrator(spp=16, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=0, seed=11)
    orc = _setup(device, sc)
    _, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    assert _rel_ok(paths, paths_c).all()
    k = np.float32(0.9)
    light = np.all(paths == np.float32(0.5), axis=1)  # camera rays that saw the emitter
    surv = np.all(np.abs(paths - 0.5 * kd / k) <= 2e-6 * 0.5, axis=1)
    dead = np.all(paths == 0.0, axis=1)
    # the emitter (kd = 0) is itself rouletted away after adding its emission
    assert np.all(light | surv | dead)
    n_obj = np.sum(surv | dead)
    assert n_obj > 10000
    p = np.sum(surv) / n_obj
    se = np.sqrt(0.9 * 0.1 / n_obj)
    assert abs(p - 0.9) < 5 * se, (p, se)
