"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must reproduce them exactly (CPU), and the HIP core must match
them -- closest-hit ids, t and counters bit-exact, per-path radiance within
1e-5 relative (GPU)."""
import os

import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _trace():
    return np.load(os.path.join(G, "trace_cornell_332.npz"))


def _paths():
    return np.load(os.path.join(G, "paths_cornell_16x16x4.npz"))


def _scene_of(g):
    sc = scenes.cornell_icosphere(subdiv=2, width=16, height=16, spp=4)
    assert np.array_equal(sc.verts, g["verts"])
    c = g["camera"]
    cam = api.pinhole(int(c[0]), int(c[1]), position=c[2:5], look_at=c[5:8], up=c[8:11], fov=float(c[11]))
    spp, max_depth, rr_depth, seed = [int(x) for x in g["integrator"]]
    integ = api.Integrator(spp=spp, termination="max-length", path_length=max_depth, seed=seed)
    return sc, cam, integ


def test_oracle_reproduces_trace_golden():
    g = _trace()
    o = O.Scene(g["verts"])
    t, prim, tests, isects = o.trace(g["rays"])
    assert np.array_equal(t.view(np.uint32), g["t"].view(np.uint32))
    assert np.array_equal(prim, g["prim"]) and np.array_equal(tests, g["tests"]) and np.array_equal(isects, g["isects"])


def test_oracle_reproduces_path_golden():
    g = _paths()
    sc, cam, integ = _scene_of(g)
    o = O.Scene(g["verts"], None, g["mat_ids"], materials=sc.materials)
    _, _, paths, _ = o.render(cam, integ.params(), per_path=True)
    assert np.array_equal(paths.view(np.uint32), g["paths"].view(np.uint32))


@pytest.mark.gpu
def test_device_matches_trace_golden(device):
    g = _trace()
    device.set_scene(api.Bvh(g["verts"]))
    t, prim, tests, isects = device.trace_host(g["rays"], L.SPT_TRACE_COUNT)
    assert np.array_equal(prim, g["prim"]) and np.array_equal(tests, g["tests"]) and np.array_equal(isects, g["isects"])
    t2, p2 = device.trace_host(g["rays"], L.SPT_TRACE_CLOSEST)
    assert np.array_equal(p2, g["prim"]) and np.array_equal(t2.view(np.uint32), g["t"].view(np.uint32))


@pytest.mark.gpu
def test_device_matches_path_golden(device):
    g = _paths()
    sc, cam, integ = _scene_of(g)
    device.set_scene(api.Bvh(g["verts"]), None, g["mat_ids"])
    device.set_materials(sc.materials)
    _, paths = device.render_image(cam, integ, per_path=True)
    ref = g["paths"]
    assert np.all(np.abs(paths - ref) <= 1e-5 * np.maximum(np.abs(paths), np.abs(ref)) + 1e-30)
