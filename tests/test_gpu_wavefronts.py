"""Concurrent wavefronts: a render call split into 1, 2 or 3 independent
wavefronts (own streams, queues and path ranges) gives every path the same
radiance bit for bit, plain and guided; a guided training call writes the same
number of records through the shared record counter."""
import numpy as np
import pytest

from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(autouse=True)
def default_tuning(device):
    device.set_tuning()
    yield
    device.set_tuning()


@pytest.fixture(scope="module")
def box(device):
    # 256 x 256 x 144 = 9.4M paths: at least 2^21 paths per wavefront for
    # three wavefronts, so the call really splits
    sc = scenes.cornell_icosphere(width=256, height=256, spp=144, subdiv=3)
    device.set_scene(api.Bvh(sc.verts), sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    return sc


@pytest.mark.parametrize("guided", [False, True])
def test_per_path_results_do_not_depend_on_the_wavefronts(device, box, guided):
    flags = 0
    if guided:
        device.guide_init(seed=3, batch=65536)
        flags = L.SPT_RENDER_GUIDED
    out = {}
    for k in (1, 2, 3):
        device.set_tuning(wavefronts=k)
        img, paths = device.render_image(box.camera, box.integrator, per_path=True, extra_flags=flags)
        st = device.render_stats()
        assert st["paths"] == 256 * 256 * 144
        out[k] = (img, paths, st["rays"])
    for k in (2, 3):
        assert _same_bits(out[1][1], out[k][1]), f"per-path radiance differs with {k} wavefronts"
        assert out[1][2] == out[k][2]  # the same extension rays
        np.testing.assert_allclose(out[k][0], out[1][0], rtol=1e-5, atol=1e-6)  # film: atomics' order


def test_training_records_are_shared_across_wavefronts(device, box):
    counts = {}
    for k in (1, 3):
        device.guide_init(seed=3, batch=65536)
        device.set_tuning(wavefronts=k)
        device.render_image(box.camera, box.integrator, extra_flags=L.SPT_RENDER_GUIDED | L.SPT_RENDER_GUIDE_TRAIN)
        st = device.guide_stats()
        counts[k] = (st["records"], st["candidates"])
        rec = device.guide_records()
        assert np.all(np.isfinite(rec[:, :13]))
    assert counts[1] == counts[3] and counts[1][0] > 1000


def test_launch_shapes_follow_the_wavefront_count(device):
    # zero fields resolve by wavefront count; fields the caller sets are kept
    device.set_tuning()
    t = device.get_tuning()
    assert (t["wavefronts"], t["shade_blocks_per_cu"], t["refill_lanes"]) == (3, 64, 20)
    device.set_tuning(wavefronts=1)
    t = device.get_tuning()
    assert (t["wavefronts"], t["shade_blocks_per_cu"], t["refill_lanes"]) == (1, 128, 48)
    device.set_tuning(wavefronts=2, refill_lanes=30)
    t = device.get_tuning()
    assert (t["wavefronts"], t["shade_blocks_per_cu"], t["refill_lanes"]) == (2, 64, 30)
