"""Batched multi-scene rendering, graph-captured bounce loops and launch
tuning: none of them may change a single path's result. Per-path radiance of
every scene in a batch is bit-identical to a single-scene render of that scene
with the same parameters; the same holds for every tuning the autotuner can
pick (queue capacity, grids, tiles, graph replay)."""
import numpy as np
import pytest

from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu


def _small_batch(n=6, seed=11):
    return scenes.batch_scenes(n=n, seed=seed, min_tris=200, max_tris=6000, min_side=16, max_side=40,
                               spp_choices=(1, 4, 9), max_depth=8)


def _single(device, sc):
    bvh = api.Bvh(sc.verts)
    device.set_scene(bvh, sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    return device.render_image(sc.camera, sc.integrator, per_path=True)


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(autouse=True)
def default_tuning(device):
    device.set_tuning()
    yield
    device.set_tuning()


def test_batch_equals_single_scene_renders(device):
    scs = _small_batch()
    b = api.Batch(device, scs, per_path=True)
    b.render()
    st, ms = b.stats()
    assert st["scenes"] == len(scs)
    assert st["paths"] == sum(s.camera.width * s.camera.height * s.integrator.effective_spp() for s in scs)
    assert np.all(ms > 0) and np.all(ms <= st["ms"] * 1.05 + 1.0)
    for i, sc in enumerate(scs):
        img, paths = _single(device, sc)
        pb = b.path_radiance(i)
        assert _same_bits(paths, pb), f"scene {i}: {np.sum(np.any(paths != pb, 1))} paths differ"
        spp = sc.integrator.effective_spp()
        np.testing.assert_allclose(b.image(i) / spp, img, rtol=1e-5, atol=1e-6)
    b.close()


@pytest.mark.parametrize("tuning", [
    dict(path_capacity=1000, tile_pixels=37, tile_samples=3),            # many generation chunks, ragged tiles
    dict(use_graph=1, check_every=2),
    dict(use_graph=1, path_capacity=4096, tile_pixels=1024, tile_samples=64, extend_blocks_per_cu=2),
])
def test_batch_tuning_invariance(device, tuning):
    scs = _small_batch(n=4, seed=12)
    ref = api.Batch(device, scs, per_path=True)
    ref.render()
    device.set_tuning(**tuning)
    b = api.Batch(device, scs, per_path=True)
    b.render()
    for i in range(len(scs)):
        assert _same_bits(ref.path_radiance(i), b.path_radiance(i)), (tuning, i)
    ref.close()
    b.close()


def test_graph_replays_match_eager(device):
    """The captured two-bounce graph is replayed many times per render; three
    renders, with eager work and synchronisation between them, each equal the
    eager launches bit for bit."""
    sc = scenes.cornell_icosphere(subdiv=3, width=32, height=32, spp=16)
    device.set_tuning(use_graph=2)
    _, eager = _single(device, sc)
    device.set_tuning(use_graph=1)
    rays = scenes.incoherent_rays(sc.verts, 4096, seed=5)
    for _ in range(3):
        _, g = device.render_image(sc.camera, sc.integrator, per_path=True)
        assert _same_bits(g, eager)
        device.trace_host(rays)  # eager work between replays
        device.sync()


def test_autotune_single_scene(device):
    sc = scenes.cornell_icosphere(subdiv=3, width=64, height=64, spp=16)
    _, ref = _single(device, sc)
    t = device.autotune(sc.camera, sc.integrator, sample_paths=1 << 16)
    assert t["path_capacity"] in (1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27)
    assert t["use_graph"] in (1, 2)
    assert device.get_tuning() == t
    _, after = device.render_image(sc.camera, sc.integrator, per_path=True)
    assert _same_bits(ref, after)


def test_batch_autotune(device):
    scs = _small_batch(n=4, seed=13)
    b = api.Batch(device, scs, per_path=True)
    b.render()
    before = [b.path_radiance(i) for i in range(len(scs))]
    t = b.autotune(sample_div=4)
    assert t["tile_pixels"] in (1024, 4096, 16384, 65536) and t["tile_samples"] in (4, 16, 64)
    b.clear()
    b.render()
    for i in range(len(scs)):
        assert _same_bits(before[i], b.path_radiance(i))
    b.close()


def test_batch_rejects_unsupported(device):
    sc = _small_batch(n=1, seed=14)[0]
    sc.integrator = api.Integrator(spp=1, termination="max-length", path_length=0)
    with pytest.raises(api.SptError):
        api.Batch(device, [sc])


@pytest.mark.parametrize("tuning", [
    dict(path_capacity=777),                         # many refills, ragged last one
    dict(path_capacity=4096, use_graph=1, check_every=2),
    dict(path_capacity=1 << 20, refill_lanes=8),
])
def test_single_scene_regeneration_invariance(device, tuning):
    """Path regeneration: with a queue much smaller than the render, paths are
    generated into the survivors' buffer after every bounce and run in other
    slots and orders; every path's radiance is bit-identical, and so is the
    replay of a host stream (per-path output index pixel * spp + sample)."""
    sc = scenes.cornell_icosphere(subdiv=3, width=40, height=24, spp=16)
    _, ref = _single(device, sc)
    rng = np.random.default_rng(3)
    stream = np.minimum(rng.random((40 * 24 * 16, 256), dtype=np.float32), np.float32(0.99999994))
    _, ref_s = device.render_image(sc.camera, sc.integrator, per_path=True, stream=stream)
    device.set_tuning(**tuning)
    _, got = device.render_image(sc.camera, sc.integrator, per_path=True)
    _, got_s = device.render_image(sc.camera, sc.integrator, per_path=True, stream=stream)
    st = device.render_stats()
    assert st["paths"] == 40 * 24 * 16
    assert _same_bits(ref, got)
    assert _same_bits(ref_s, got_s)
