"""Pins of the float-only transcendentals (swpathtracer_amd/csrc/spt_fmath.h)
that the device and the oracle share in place of the C library's cosf, sinf,
expf, logf and atanf (bsdf_samplers/cosine_bsdf_sampler.cpp:13,
bsdf_samplers/cook_torrance_bsdf_sampler.cpp:11-14, materials/cook_torrance.cpp:33).

Every result is compared with the float64 value of the same float input: at
most one ulp everywhere on the domains the samplers use (and well beyond),
and the correctly rounded float in at least 85-95% of cases, so a dropped
correction term, a wrong quadrant or a mis-split constant fails."""
import numpy as np
import pytest

import pyoracle as O

ULP_MAX = 1.0


def _ulps(got, exact):
    e = np.abs(exact).astype(np.float32)
    u = np.spacing(np.where(e == 0, np.float32(1e-45), e)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / u


def _windows(centres, half=1 << 14):
    """Every float within +-half ulps of each centre (positive centres)."""
    out = []
    for c in centres:
        b = int(np.float32(c).view(np.uint32))
        out.append(np.arange(max(b - half, 0), b + half, dtype=np.uint32).view(np.float32))
    return np.concatenate(out)


def _check(name, x, ref_fn, cr_min=0.9):
    y = O.fmath(name, x)
    ref = ref_fn(x.astype(np.float64))
    ok = np.isfinite(ref) & (np.abs(ref) >= np.finfo(np.float32).tiny)
    err = _ulps(y[ok], ref[ok])
    worst = np.argmax(err)
    print(f"{name}: {ok.sum()} inputs, max {err.max():.3f} ulp at {x[ok][worst]!r}, "
          f"correctly rounded {np.mean(y[ok] == ref[ok].astype(np.float32)) * 100:.2f}%")
    assert err.max() <= ULP_MAX, (name, x[ok][worst], y[ok][worst], ref[ok][worst])
    assert np.mean(y[ok] == ref[ok].astype(np.float32)) >= cr_min
    return y


@pytest.mark.parametrize("name,fn", [("sin", np.sin), ("cos", np.cos)])
def test_sin_cos(name, fn):
    rng = np.random.default_rng(1)
    two_pi = np.float32(2 * np.pi)
    x = np.concatenate([
        rng.random(2_000_000, dtype=np.float32) * two_pi,           # phi = 2 pi u (all samplers)
        rng.random(500_000, dtype=np.float32) * np.float32(np.pi / 2),  # Beckmann theta
        np.linspace(-100, 100, 1_000_001, dtype=np.float32),
        _windows([np.pi / 4, np.pi / 2, 3 * np.pi / 4, np.pi, 3 * np.pi / 2, 2 * np.pi, 1.0, 1e-3]),
        -_windows([np.pi / 2, np.pi]),
    ])
    _check(name, x, fn, cr_min=0.95)
    s, c = O.fmath("sin", np.float32([0, np.pi / 2])), O.fmath("cos", np.float32([0, np.pi]))
    assert s[0] == 0 and s[1] == 1 and c[0] == 1 and c[1] == -1
    tiny = np.float32([1e-30, -1e-20, 1e-8])
    assert np.array_equal(O.fmath("sin", tiny), tiny)  # sin x = x below 2^-12


def test_exp():
    rng = np.random.default_rng(2)
    x = np.concatenate([
        -rng.random(2_000_000, dtype=np.float32) * np.float32(30),   # Beckmann exponent (<= 0)
        rng.uniform(-87.3, 88.7, 1_000_000).astype(np.float32),
        np.linspace(-1, 1, 1_000_001, dtype=np.float32),
        _windows([0.5, 1.0, 10.0, 80.0]),
        -_windows([0.5, 1.0, 10.0, 80.0]),
    ])
    _check("exp", x, np.exp, cr_min=0.85)
    # results in the subnormal range: within one subnormal step
    xs = np.linspace(-103.9, -87.4, 200_001, dtype=np.float32)
    ys = O.fmath("exp", xs).astype(np.float64)
    assert np.max(np.abs(ys - np.exp(xs.astype(np.float64)))) <= 2.0 ** -149
    sp = O.fmath("exp", np.float32([0, -0.0, -np.inf, np.inf, 89.0, -104.0]))
    assert sp[0] == 1 and sp[1] == 1 and sp[2] == 0 and np.isinf(sp[3]) and np.isinf(sp[4]) and sp[5] == 0


def test_log():
    rng = np.random.default_rng(3)
    u = rng.random(2_000_000, dtype=np.float32)
    x = np.concatenate([
        np.float32(1) - u,                                             # log(1 - r1), r1 in [0, 1)
        np.float32(1) - np.arange(1, 1 << 16, dtype=np.float32) * np.float32(2.0 ** -24),
        np.exp(rng.uniform(-100, 88, 1_000_000)).astype(np.float32),
        _windows([1.0, np.sqrt(0.5), np.sqrt(2.0), 2.0, 0.5, 1e-30, 3e38]),
        np.float32([1e-40, 1e-45]),
    ])
    _check("log", x[x > 0], np.log)
    sp = O.fmath("log", np.float32([1, 0, np.inf, -1]))
    assert sp[0] == 0 and sp[1] == -np.inf and sp[2] == np.inf and np.isnan(sp[3])


def test_atan():
    rng = np.random.default_rng(4)
    x = np.concatenate([
        np.sqrt(-np.float32(0.2) ** 2 * np.log1p(-rng.random(1_000_000))).astype(np.float32),  # Beckmann
        rng.random(1_000_000, dtype=np.float32) * np.float32(4),
        np.exp(rng.uniform(-20, 40, 1_000_000)).astype(np.float32),
        _windows([0.4375, 0.6875, 1.1875, 2.4375, 0.5, 1.0, 1.5]),
    ])
    x = np.concatenate([x, -x[:100_000]])
    _check("atan", x, np.arctan)
    sp = O.fmath("atan", np.float32([0, 1, np.inf, -np.inf]))
    assert sp[0] == 0 and sp[1] == np.float32(np.pi / 4)
    assert sp[2] == np.float32(np.pi / 2) and sp[3] == -np.float32(np.pi / 2)
