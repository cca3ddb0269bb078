"""What the trained guide gains. On the slit-lit scene (800 x 600, the
1,048,576-triangle layered mesh, light entering through a slit of 2% of the
wall), three 64-spp images are compared by per-pixel RMSE against an FP64
film image of 4,096 spp with an independent seed:
- plain BSDF sampling (integrators/pathtracer.cpp:36-54);
- the one-sample mixture with an untrained guide (initial parameters);
- the same mixture after the guide has been trained online on the records
  of a few training renders by the device's forward/backward kernels.
The trained guide must lower the RMSE clearly below both; the untrained one
shows that the gain comes from training, not from the mixture itself."""
import dataclasses

import numpy as np
import pytest

from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu


def _rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))


def test_trained_guide_lowers_rmse_at_equal_samples(device):
    sc = scenes.slit_box()
    device.set_scene(api.Bvh(sc.verts, n_threads=16), sc.normals, sc.mat_ids)
    device.set_materials(sc.materials)
    cam = sc.camera
    ref, _ = device.render_image(cam, dataclasses.replace(sc.integrator, spp=4096, seed=0x5EED), film64=True)
    print(f"reference: 4096 spp, FP64 film, mean {ref.reshape(-1, 3).mean(0)}", flush=True)
    i64 = dataclasses.replace(sc.integrator, spp=64, seed=11)
    plain, _ = device.render_image(cam, i64)
    device.guide_init(seed=1, lr=1e-2, batch=262144, steps_per_pass=24)
    untrained, _ = device.render_image(cam, i64, extra_flags=L.SPT_RENDER_GUIDED)
    train = L.SPT_RENDER_GUIDED | L.SPT_RENDER_GUIDE_TRAIN
    for k, spp in enumerate([4, 4, 4, 4, 64, 64, 64, 64]):
        device.render_image(cam, dataclasses.replace(sc.integrator, spp=spp, seed=100 + k), extra_flags=train)
        device.guide_train_collected()
    st = device.guide_stats()
    trained, _ = device.render_image(cam, i64, extra_flags=L.SPT_RENDER_GUIDED)
    r_plain, r_un, r_tr = _rmse(plain, ref), _rmse(untrained, ref), _rmse(trained, ref)
    print(f"64 spp RMSE: plain {r_plain:.5f}, untrained guide {r_un:.5f}, trained guide {r_tr:.5f} "
          f"({st['steps']} Adam steps, {st['trained_records']} records)", flush=True)
    assert st["steps"] >= 100 and st["skipped_steps"] < st["steps"]
    assert np.all(np.isfinite(trained))
    assert r_tr < 0.9 * r_plain
    assert r_tr < 0.9 * r_un
    # unbiased: the three images estimate the same mean
    for img in (plain, untrained, trained):
        assert np.allclose(img.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0), rtol=0.03)
