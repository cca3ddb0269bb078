"""Pins of the oracle's shading and sampling (CPU only), from closed forms:
the published SplitMix64 output, sampler moments and normalisation, BSDF
reciprocity and energy bounds, conductor Fresnel special cases, the white
furnace (samples/furnace_test.lua) and the pinhole camera's geometry."""
import math

import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L


def test_rng_matches_published_splitmix64():
    """key(seed 0, path 0) = mix(0) = 0, so u(k) are the top 24 bits of the
    SplitMix64 stream seeded with 0 (Steele, Lea, Flood 2014; Vigna's
    splitmix64.c): 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    for k, word in enumerate([0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]):
        assert O.rng_u(0, 0, k) == np.float32((word >> 40) / 2.0 ** 24)
    u = np.array([O.rng_u(7, p, k) for p in range(200) for k in range(50)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def _sample_many(mat, wo, n=20000, seed=3):
    return np.array([O.material_sample(mat, wo, seed, p) for p in range(n)])


def test_cosine_sampler_moments_and_pdf():
    m = api.material([api.lambertian((0.5, 0.6, 0.7))])
    s = _sample_many(m, [0.0, 1.0, 0.0])
    wi, pdf, f = s[:, 0:3], s[:, 3], s[:, 4:7]
    assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-6)
    assert abs(wi[:, 1].mean() - 2.0 / 3.0) < 0.01        # E[cos] = 2/3
    np.testing.assert_allclose(pdf, wi[:, 1] / math.pi, rtol=1e-6)
    est = f * wi[:, 1:2] / pdf[:, None]                  # f cos / pdf = kd
    np.testing.assert_allclose(est, np.broadcast_to([0.5, 0.6, 0.7], est.shape), rtol=1e-5)


def test_uniform_sampler_moments():
    m = api.material([api.lambertian((0.5, 0.5, 0.5), sampler="uniform")])
    s = _sample_many(m, [0.0, 1.0, 0.0])
    assert abs(s[:, 1].mean() - 0.5) < 0.01
    np.testing.assert_allclose(s[:, 3], 1.0 / (2.0 * math.pi), rtol=1e-6)


def _hemisphere_integral(fn, n_t=256, n_p=256):
    th = (np.arange(n_t) + 0.5) / n_t * (math.pi / 2)
    ph = (np.arange(n_p) + 0.5) / n_p * (2 * math.pi)
    total = 0.0
    for t in th:
        for p in ph[::1]:
            wi = [math.sin(t) * math.cos(p), math.cos(t), math.sin(t) * math.sin(p)]
            total += fn(wi) * math.sin(t)
    return total * (math.pi / 2 / n_t) * (2 * math.pi / n_p)


@pytest.mark.parametrize("layer,mass", [
    # Beckmann: P(tan^2 theta_m > 1) = exp(-1 / m^2), negligible at m = 0.2
    (api.cook_torrance(0.2, (1, 1, 1)), 1.0),
    # GGX: P(tan^2 theta_m > 1) = a^2 / (1 + a^2); those half vectors send wi below the horizon
    (api.ggx_conductor(0.2), 1.0 - 0.04 / 1.04),
])
def test_microfacet_pdf_integrates_to_upper_hemisphere_mass(layer, mass):
    """Normal incidence: wi = reflect(wo, m) lies above the horizon iff
    theta_m < 45 degrees, so the density of wi over the upper hemisphere
    integrates to the distribution's mass of tan^2 theta_m < 1."""
    wo = [0.0, 1.0, 0.0]
    val = _hemisphere_integral(lambda wi: O.layer_eval(layer, wo, wi)[3], 200, 64)
    assert abs(val - mass) < 0.01


def test_sampled_pdf_equals_evaluated_pdf():
    for layer in (api.cook_torrance(0.3, (0.9, 0.9, 0.9)), api.ggx_conductor(0.3)):
        m = api.material([layer])
        wo = np.array([0.3, 0.8, -0.2], np.float32)
        wo /= np.linalg.norm(wo)
        s = _sample_many(m, wo, 500)
        for row in s:
            if row[3] > 0 and row[1] > 0 and row[4] > 0:
                ev = O.layer_eval(layer, wo, row[0:3])
                assert abs(ev[3] - row[3]) <= 1e-5 * row[3]


def test_ggx_reciprocity_and_energy():
    layer = api.ggx_conductor(0.3, eta=(1.0, 1.0, 1.0), k=(1e4, 1e4, 1e4))  # F ~ 1
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[1], b[1] = abs(a[1]), abs(b[1])
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        f1 = O.layer_eval(layer, a, b)[0]
        f2 = O.layer_eval(layer, b, a)[0]
        assert abs(f1 - f2) <= 1e-5 * max(f1, f2) + 1e-12
    for cos_o in (1.0, 0.7, 0.3):
        wo = [math.sqrt(1 - cos_o ** 2), cos_o, 0.0]
        albedo = _hemisphere_integral(lambda wi: O.layer_eval(layer, wo, wi)[0] * wi[1], 200, 64)
        assert albedo <= 1.0 + 1e-3
        assert albedo > 0.7  # single scattering loses some energy, most at grazing wo


def test_conductor_fresnel_special_cases():
    for eta, k in [(1.5, 0.0), (0.143, 3.98), (0.375, 2.39), (1.44, 1.60)]:
        r0 = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)  # normal incidence
        assert abs(O.fresnel_conductor(1.0, eta, k) - r0) < 1e-6
        assert abs(O.fresnel_conductor(0.0, eta, k) - 1.0) < 1e-6  # grazing
    # k = 0 is the unpolarised dielectric Fresnel term
    eta, ci = 1.5, 0.6
    st = math.sqrt(1 - ci * ci) / eta
    ct = math.sqrt(1 - st * st)
    rs = ((ci - eta * ct) / (ci + eta * ct)) ** 2
    rp = ((eta * ci - ct) / (eta * ci + ct)) ** 2
    assert abs(O.fresnel_conductor(ci, eta, 0.0) - 0.5 * (rs + rp)) < 1e-6


def test_layered_bsdf_conserves_energy():
    """Dielectric over Lambertian (samples/furnace_test.lua 'plastic'): the
    estimator f * cos / pdf never exceeds 1 on average."""
    m = api.material([api.smooth_dielectric(1.49), api.lambertian((0.999, 0.999, 0.999))])
    s = _sample_many(m, [0.0, 1.0, 0.0], 20000)
    est = s[:, 4] * np.abs(s[:, 1]) / s[:, 3]
    assert est.mean() <= 1.0 + 5e-3
    assert np.all(s[:, 10] >= 1)  # at least one draw (the dielectric's Fresnel choice)


def test_white_furnace_oracle():
    sc = scenes.furnace(albedo=(0.999, 0.5, 0.25))
    o = O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)
    _, _, paths, _ = o.render(sc.camera, sc.integrator.params(), per_path=True)
    obj = np.abs(paths[:, 1] - 0.5) > 0.1
    assert obj.mean() > 0.3
    np.testing.assert_allclose(paths[obj], np.broadcast_to([0.4995, 0.25, 0.125], paths[obj].shape), rtol=2e-6)
    np.testing.assert_allclose(paths[~obj], 0.5, rtol=2e-6)


def test_pinhole_camera_closed_form():
    w, h, fov = 65, 49, 40.0
    cam = api.pinhole(w, h, position=(0, 0, 3), look_at=(0, 0, 0), fov=fov)
    r = O.camera_ray(cam, (w - 1) / 2 + 0.5, (h - 1) / 2 + 0.5)  # image centre
    assert np.array_equal(r[4:7], np.array([0, 0, -1], np.float32))
    d = (h * 0.5) / math.tan(math.radians(fov) / 2)
    r = O.camera_ray(cam, 0.5, 0.5)  # first pixel centre
    v = np.array([-w / 2 + 0.5, h / 2 - 0.5, -d])
    np.testing.assert_allclose(r[4:7], v / np.linalg.norm(v), atol=2e-7)
    assert np.array_equal(r[0:3], np.array([0, 0, 3], np.float32))


def test_jittered_samples_stay_in_their_strata():
    spp = 16
    n = 4
    for sidx in range(spp):
        for p in range(20):
            sx, sy, draws = O.pixel_sample(L.SPT_PIXEL_JITTERED, spp, sidx, 5, p)
            i, j = sidx // n, sidx % n
            assert i / n <= sx < (i + 1) / n and j / n <= sy < (j + 1) / n
            assert draws == 4  # getSample is called twice, two draws each
    sx, sy, draws = O.pixel_sample(L.SPT_PIXEL_REGULAR, 1, 0, 5, 0)
    assert (sx, sy, draws) == (0.5, 0.5, 0)


# ---------------------------------------------------------------------------
# Closed forms of the smooth dielectric, Schlick, Beckmann, the conductor and
# layered furnaces and Russian roulette (each fails on a plausible mistake:
# a refraction sent upward, a wrong IOR ratio or exponent, a dropped term).
# ---------------------------------------------------------------------------
import _furnace_geom as FG


def _first_draw_path(seed, want_below, thresh, k=0):
    """A path index whose draw k is below (or at/above) thresh."""
    for p in range(100000):
        u = O.rng_u(seed, p, k)
        if (u < thresh) == want_below and abs(u - thresh) > 1e-4:
            return p
    raise AssertionError("no such path")


@pytest.mark.parametrize("cos_o", [1.0, 0.8, 0.5, 0.2, -1.0, -0.9, -0.75])
def test_dielectric_snell_refraction_and_mirror(cos_o):
    """materials/smooth_dielectric.cpp:30-51 and bsdf_samplers/refraction_bsdf_sampler.cpp:5-24:
    refraction keeps the plane of incidence, flips the side, and scales the
    tangential part by eta_i / eta_t (Snell); the Fresnel choice reflects about
    the normal; pdf 1, f = 1 / wi.y."""
    ior = 1.5
    m = api.material([api.smooth_dielectric(ior)])
    phi = 0.7
    s_o = math.sqrt(1 - cos_o * cos_o)
    wo = np.array([s_o * math.cos(phi), cos_o, s_o * math.sin(phi)], np.float32)
    eta_i, eta_t = (1.0, ior) if cos_o > 0 else (ior, 1.0)
    F = O.schlick_ior(1.0, ior, float(wo[1]))
    # refraction: the Fresnel draw at or above F
    p = _first_draw_path(5, False, F)
    r = O.material_sample(m, wo, 5, p)
    wi, pdf, f, draws = r[0:3], r[3], r[4:7], r[10]
    assert draws == 1 and pdf == 1.0
    assert np.sign(wi[1]) == -np.sign(wo[1])
    assert abs(np.linalg.norm(wi) - 1) < 1e-6
    t_expect = -(eta_i / eta_t) * np.array([wo[0], wo[2]], np.float64)
    np.testing.assert_allclose([wi[0], wi[2]], t_expect, atol=2e-6)
    cos_t = math.sqrt(1 - (eta_i / eta_t) ** 2 * (1 - cos_o * cos_o))
    assert abs(abs(wi[1]) - cos_t) < 2e-6
    np.testing.assert_allclose(f, 1.0 / wi[1], rtol=1e-6)
    # mirror: the draw below F
    if F > 1e-3:
        p = _first_draw_path(5, True, F)
        r = O.material_sample(m, wo, 5, p)
        np.testing.assert_array_equal(r[0:3], np.array([-wo[0], wo[1], -wo[2]], np.float32))
        assert r[10] == 1


def test_dielectric_total_internal_reflection_draws_nothing():
    """Inside the dielectric beyond the critical angle (eta^2 (1 - cos^2) > 1):
    mirror reflection without a Fresnel draw (smooth_dielectric.cpp:38-40)."""
    m = api.material([api.smooth_dielectric(1.5)])
    crit = math.sqrt(1 - 1 / 1.5 ** 2)  # |cos| below this: TIR
    for c in (-0.7, -0.5, -0.1):
        assert abs(c) < crit
        s = math.sqrt(1 - c * c)
        wo = np.array([s, c, 0.0], np.float32)
        for p in range(20):
            r = O.material_sample(m, wo, 3, p)
            np.testing.assert_array_equal(r[0:3], np.array([-wo[0], wo[1], -wo[2]], np.float32))
            assert r[10] == 0


def test_schlick_ior_values():
    """FresnelSchlick(ior_i, ior_t) (materials/fresnel_schlick.cpp:20-38):
    R0 = ((n_i - n_t) / (n_i + n_t))^2 + (1 - R0)(1 - cos)^5; for cos < 0 the
    IORs swap and the transmitted cosine sqrt(1 - eta^2 (1 - cos^2)) replaces
    it."""
    r0 = (0.5 / 2.5) ** 2
    for c in (1.0, 0.9, 0.5, 0.1, 0.0):
        assert abs(O.schlick_ior(1.0, 1.5, c) - (r0 + (1 - r0) * (1 - c) ** 5)) < 1e-7
    for c in (-1.0, -0.95, -0.9, -0.8):
        ct = math.sqrt(1 - 1.5 ** 2 * (1 - c * c))
        assert abs(O.schlick_ior(1.0, 1.5, c) - (r0 + (1 - r0) * (1 - ct) ** 5)) < 1e-6
    assert abs(O.schlick_ior(1.0, 1.0, 0.3) - 0.7 ** 5) < 1e-7  # matched media: R0 = 0
    assert abs(O.schlick_ior(1.5, 1.0, 0.6) - O.schlick_ior(1.0, 1.5, 0.6)) < 1e-7  # R0 symmetric


@pytest.mark.parametrize("wo,wi", [
    ([0.0, 1.0, 0.0], [0.0, 1.0, 0.0]),
    ([0.3, 0.9, 0.1], [-0.2, 0.8, -0.3]),
    ([0.9, 0.2, 0.0], [-0.8, 0.3, 0.2]),      # grazing: the V-cavity G clips
    ([0.0, 0.6, 0.8], [0.5, 0.5, -0.7]),
    ([0.6, 0.1, 0.0], [-0.6, 0.1, 0.05]),     # G = 2 nh nv / vh < 1
])
def test_beckmann_value_closed_form(wo, wi):
    """CookTorrance::f (materials/cook_torrance.cpp:9-44): F D G / (4 nv nl)
    with D = exp((nh^2 - 1) / (m^2 nh^2)) / (pi m^2 nh^4), G = min(1, 2 nh nv /
    vh, 2 nh nl / vh), F = Schlick(R0, l.h); and the sampler's density
    D nh / (4 wi.h) (cook_torrance_bsdf_sampler.cpp:22-42)."""
    m_, r0 = 0.3, np.array([0.9, 0.6, 0.3])
    layer = api.cook_torrance(m_, tuple(r0))
    a = np.array(wo, np.float64)
    b = np.array(wi, np.float64)
    a /= np.linalg.norm(a)
    b /= np.linalg.norm(b)
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    a, b = a32.astype(np.float64), b32.astype(np.float64)
    h = (a + b) / np.linalg.norm(a + b)
    nv, nl, vh, nh, lh = a[1], b[1], a @ h, h[1], b @ h
    D = math.exp((nh * nh - 1) / (m_ * m_ * nh * nh)) / (math.pi * m_ * m_ * nh ** 4)
    G = min(1.0, 2 * nh * nv / vh, 2 * nh * nl / vh)
    F = r0 + (1 - r0) * (1 - lh) ** 5
    f_expect = F * D * G / (4 * nv * nl)
    out = O.layer_eval(layer, a32, b32)
    np.testing.assert_allclose(out[0:3], f_expect, rtol=2e-5)
    np.testing.assert_allclose(out[3], D * nh / (4 * lh), rtol=2e-5)


def _furnace_orc(sc):
    return O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)


def test_conductor_furnace_closed_form():
    """Smooth conductor with r0 = (0.5, 0.2, 0.8) in the furnace: the camera
    ray mirrors off the object into the emitter, so every object path returns
    exactly 0.5 F_schlick(cos), cos between the normal and the camera ray
    (smooth_conductor.cpp:6-20, fresnel_schlick.cpp:11-18); emitter paths 0.5."""
    r0 = (0.5, 0.2, 0.8)
    sc = scenes.furnace(layer=api.smooth_conductor(r0), width=24, height=24, spp=4)
    orc = _furnace_orc(sc)
    _, _, paths, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    c, obj = FG.first_vertex(sc, orc)
    assert obj.mean() > 0.3
    np.testing.assert_allclose(paths[obj], 0.5 * FG.schlick_r0(r0, c[obj]), rtol=1e-5)
    np.testing.assert_allclose(paths[~obj], 0.5, rtol=1e-6)


def test_layered_dielectric_over_black_returns_the_coat_fresnel():
    """Smooth dielectric (IOR 1.5) over a kd = 0 Lambertian in the furnace: the
    layered walk (layered_bsdf.cpp:10-74) reflects off the coat with
    probability F_schlick(cos) (its first draw, after the pixel sampler's four,
    below F) and then returns the emitter's 0.5; a refracted path meets the
    black base and returns 0. Checked path by path against the draw."""
    sc = scenes.furnace(width=24, height=24, spp=4)
    sc.materials[1] = api.material([api.smooth_dielectric(1.5), api.lambertian((0.0, 0.0, 0.0))])
    orc = _furnace_orc(sc)
    p = sc.integrator.params()
    _, _, paths, _ = orc.render(sc.camera, p, per_path=True)
    c, obj = FG.first_vertex(sc, orc)
    F = FG.schlick_ior(1.0, 1.5, c)
    u = np.array([O.rng_u(p.seed, k, 4) for k in range(len(paths))])
    clear = obj & (np.abs(u - F) > 1e-5)
    assert clear.sum() > 1000
    refl = u < F
    np.testing.assert_allclose(paths[clear & refl], 0.5, rtol=1e-6)
    assert np.all(paths[clear & ~refl] == 0.0)
    assert np.sum(clear & refl) > 20  # both outcomes occur


def test_russian_roulette_from_depth_zero_keeps_the_mean():
    """pathtracer.cpp:43-49 with RR from the first vertex: an object path
    survives with probability k = max(kd) and then carries kd / k into the
    emitter (0.5 kd / k), else returns 0; the survival fraction is k and the
    estimator's mean 0.5 kd."""
    kd = np.array([0.9, 0.5, 0.25])
    sc = scenes.furnace(albedo=tuple(kd), width=48, height=48, spp=16)
    sc.integrator = api.Integrator(spp=16, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=0, seed=11)
    orc = _furnace_orc(sc)
    _, _, paths, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    light = np.all(paths == np.float32(0.5), axis=1)
    surv = np.all(np.abs(paths - 0.5 * kd / 0.9) <= 1e-6, axis=1)
    dead = np.all(paths == 0.0, axis=1)
    assert np.all(light | surv | dead)
    n = np.sum(surv | dead)
    assert n > 10000
    frac = np.sum(surv) / n
    assert abs(frac - 0.9) < 5 * math.sqrt(0.09 / n)
    mean = paths[surv | dead].astype(np.float64).mean(0)
    se = paths[surv | dead].astype(np.float64).std(0) / math.sqrt(n)
    assert np.all(np.abs(mean - 0.5 * kd) < 5 * se)
