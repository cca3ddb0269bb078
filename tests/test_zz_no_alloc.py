"""spt_trace_noalloc (closest hit with no device allocation, scratch in a
static device array) against spt_trace on the same rays, through the C ABI:
ids and t bit for bit, two calls in a row with the second batch smaller, on
the cornell scene and on the 65,536-triangle displaced mesh; sizes past the
limit are rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from swpathtracer_amd import api, scenes
import swpathtracer_amd._lib as L

pytestmark = pytest.mark.gpu


def _noalloc(dev, rays):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    dr = dev.to_device(rays)
    dh = dev.alloc(8 * n)
    dev._check(dev.lib.spt_trace_noalloc(dev.h, dr.ptr, n, dh.ptr))
    dev.sync()
    h = dh.download(np.uint32, (n, 2))
    return h[:, 0].view(np.float32).copy(), h[:, 1].copy()


@pytest.mark.parametrize("scene", ["cornell", "displaced_65k"])
def test_noalloc_trace_equals_trace(device, scene):
    if scene == "cornell":
        sc = scenes.cornell_icosphere(subdiv=3)
    else:
        sc = scenes.displaced_sphere(128, 256, 0.1, 4.0, 2)
    device.set_scene(api.Bvh(sc.verts), sc.normals, sc.mat_ids)
    rays = scenes.incoherent_rays(sc.verts, 300_000, seed=11)
    print(f"{scene}: {sc.n_tris} triangles, calls of 300000 then 77777 rays", flush=True)
    for n in (300_000, 77_777):  # the second call reuses the scratch with a smaller input
        r = rays[:n] if n == 300_000 else rays[-n:]
        t_ref, p_ref = device.trace_host(r)
        t_na, p_na = _noalloc(device, r)
        assert np.array_equal(p_na, p_ref), f"{np.sum(p_na != p_ref)} ids differ"
        assert np.array_equal(t_na.view(np.uint32), t_ref.view(np.uint32))
        assert np.mean(p_ref != L.SPT_MISS) > 0.2


def test_noalloc_trace_rejects_oversize_and_empty(device):
    sc = scenes.cornell_icosphere(subdiv=1)
    device.set_scene(api.Bvh(sc.verts), sc.normals, sc.mat_ids)
    buf = device.alloc(64)
    s = device.lib.spt_trace_noalloc(device.h, buf.ptr, L.SPT_NOALLOC_MAX_RAYS + 1, buf.ptr)
    assert s == L.SPT_ERR_INVALID
    assert b"SPT_NOALLOC_MAX_RAYS" in device.lib.spt_last_error(device.h)
    assert device.lib.spt_trace_noalloc(device.h, None, 0, None) == L.SPT_OK
    device.sync()
