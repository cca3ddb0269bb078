"""Path-tracing parity: the wavefront HIP tracer vs the oracle's recursive
restatement of PathTracer::traceRay (integrators/pathtracer.cpp:12-57) on the
same draws. Per path radiance within 1e-5 relative (FP32)."""
import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _setup(device, sc):
    bvh = api.Bvh(sc.verts)
    device.set_scene(bvh, sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    return O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)


def _rel_ok(a, b):
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(a), np.abs(b)) + 1e-30, axis=1)


def test_replay_counter_rng_cornell(device):
    sc = scenes.cornell_icosphere(width=64, height=64, spp=16)
    orc = _setup(device, sc)
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    img_c, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    ok = _rel_ok(paths, paths_c)
    frac = ok.mean()
    print(f"paths {len(ok)} within {RTOL}: {frac * 100:.4f}%  mismatched {np.sum(~ok)}")
    assert ok.all(), np.nonzero(~ok)[0][:10]
    assert np.allclose(img, img_c, rtol=1e-4, atol=1e-6)


def test_replay_host_stream(device):
    """Replay mode: draws come from a host-generated stream, same for both."""
    sc = scenes.cornell_icosphere(subdiv=3, width=32, height=32, spp=4)
    orc = _setup(device, sc)
    n_paths = 32 * 32 * 4
    stream = np.random.default_rng(9).random((n_paths, 64), dtype=np.float32)
    stream = np.minimum(stream, np.float32(0.99999994))
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True, stream=stream)
    _, _, paths_c, ov = orc.render(sc.camera, sc.integrator.params(), per_path=True, stream=stream)
    assert ov == 0
    assert _rel_ok(paths, paths_c).all()




@pytest.mark.parametrize("layer", ["lambertian", "conductor"])
def test_white_furnace(device, layer):
    """Convex object of albedo rho inside a kd=0 emitter of radiance 0.5:
    every path returns 0.5 * rho (samples/furnace_test.lua)."""
    if layer == "lambertian":
        sc = scenes.furnace(albedo=(0.999, 0.5, 0.25))
        expect = 0.5 * np.array([0.999, 0.5, 0.25], np.float32)
    else:
        sc = scenes.furnace(layer=api.smooth_conductor((1.0, 1.0, 1.0)))
        expect = None
    _setup(device, sc)
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    if expect is not None:
        hit_obj = np.abs(paths[:, 1] - 0.5) > 0.1  # paths that saw the object, not the emitter
        assert hit_obj.mean() > 0.3
        np.testing.assert_allclose(paths[hit_obj], np.broadcast_to(expect, paths[hit_obj].shape), rtol=2e-6)
    else:
        np.testing.assert_allclose(paths, 0.5, rtol=2e-6)


def test_layered_gold_mesh_replay(device):
    """Layered dielectric-over-GGX-gold mesh in the box, RR from depth 3, cap 16."""
    sc = scenes.mesh_in_box(rings=64, segments=128, width=48, height=36, spp=16)
    orc = _setup(device, sc)
    img, paths = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, _, paths_c, _ = orc.render(sc.camera, sc.integrator.params(), per_path=True)
    ok = _rel_ok(paths, paths_c)
    print(f"paths {len(ok)} within {RTOL}: {ok.mean() * 100:.4f}%  mismatched {np.sum(~ok)}")
    assert ok.all()


def test_render_sort_invariance(device):
    """Material sorting permutes the shading order only: per-path output is
    identical with and without it."""
    sc = scenes.cornell_icosphere(subdiv=3, width=32, height=32, spp=4)
    _setup(device, sc)
    p = sc.integrator.params()
    n = 32 * 32 * 4
    outs = []
    for flags in (0, L.SPT_RENDER_NO_SORT):
        d = device.alloc(n * 12)
        p.flags = L.SPT_RENDER_PER_PATH | flags
        device.render(sc.camera, p, None, None, d.ptr)
        device.sync()
        outs.append(d.download(np.float32, (n, 3)))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
