"""Python face of the C ABI: tree build, device context, ray casting and path
tracing. Mirrors the reference's host interfaces: ``Bvh`` stands for
SBVH::build (acceleration/sbvh.cpp:12-37), ``Device.trace`` for a batch of
AccelStructure::intersect calls (acceleration/accel_structure.h:29-30),
``Device.raycast`` for RayCaster<T>::render (integrators/raycaster.h:61-99)
and ``Device.render`` for PathTracer::render (integrators/pathtracer.cpp:59-147).
Errors raise ``SptError`` (the reference throws std::runtime_error).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


class SptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spt error {code}: {msg}")
        self.code = code


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def _f3(v) -> L.F3:
    return L.F3(*[float(x) for x in v])


# ----------------------------------------------------------------------------
# materials / cameras
# ----------------------------------------------------------------------------

def lambertian(kd=(0.75, 0.75, 0.75), sampler: str = "importance") -> L.Layer:
    """Lambertian{kd, bsdf_sampler} (lua_core/bsdfs.lua:11-27 defaults)."""
    return L.Layer(type=L.SPT_BSDF_LAMBERTIAN, sampler=_sampler(sampler), kd=_f3(kd))


def cook_torrance(roughness=0.2, r0=(0.99, 0.99, 0.99), sampler: str = "importance") -> L.Layer:
    """CookTorrance{roughness, reflectance_at_normal_incidence} (bsdfs.lua:61-78)."""
    return L.Layer(type=L.SPT_BSDF_COOK_TORRANCE, sampler=_sampler(sampler), roughness=roughness, r0=_f3(r0))


def smooth_conductor(r0=(0.99, 0.99, 0.99), sampler: str = "importance") -> L.Layer:
    """SmoothConductor{reflectance_at_normal_incidence} (bsdfs.lua:29-43)."""
    return L.Layer(type=L.SPT_BSDF_SMOOTH_CONDUCTOR, sampler=_sampler(sampler), r0=_f3(r0))


def smooth_dielectric(ior=1.5) -> L.Layer:
    """SmoothDielectric{ior} (bsdfs.lua:45-59)."""
    return L.Layer(type=L.SPT_BSDF_SMOOTH_DIELECTRIC, sampler=0, ior=ior)


GOLD_ETA = (0.143119, 0.374957, 1.44248)  # Au, complex IOR at ~650/550/450 nm
GOLD_K = (3.98316, 2.38572, 1.60322)


def ggx_conductor(alpha=0.2, eta=GOLD_ETA, k=GOLD_K, sampler: str = "importance") -> L.Layer:
    """GGX microfacet conductor with exact complex-IOR Fresnel (no reference counterpart)."""
    return L.Layer(type=L.SPT_BSDF_GGX_CONDUCTOR, sampler=_sampler(sampler), roughness=alpha, eta=_f3(eta),
                   k=_f3(k))


def _sampler(name: str) -> int:
    if name not in ("importance", "uniform"):
        raise ValueError("bsdf_sampler must be 'importance' or 'uniform'")
    return L.SPT_SAMPLER_IMPORTANCE if name == "importance" else L.SPT_SAMPLER_UNIFORM


def material(layers: Sequence[L.Layer], emission=(0.0, 0.0, 0.0)) -> L.Material:
    """Material{bsdf = LayeredBSDF{...}, emission = Emission{...}}; layers[0] is the top."""
    if not 1 <= len(layers) <= L.SPT_MAX_LAYERS:
        raise ValueError("1..4 layers")
    m = L.Material(n_layers=len(layers), emission=_f3(emission))
    for i, l in enumerate(layers):
        m.layers[i] = l
    return m


def pinhole(width, height, position=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), fov=55.0) -> L.Camera:
    """Camera{type='pinhole'} (lua_core/cameras.lua defaults: position {0,0,5}, fov 55)."""
    return L.Camera(type=L.SPT_CAMERA_PINHOLE, width=width, height=height, position=_f3(position),
                    look_at=_f3(look_at), up=_f3(up), fov_deg=fov)


def orthographic(width, height, min_x, max_x, min_y, max_y, position=(0, 0, 5), look_at=(0, 0, 0),
                 up=(0, 1, 0)) -> L.Camera:
    return L.Camera(type=L.SPT_CAMERA_ORTHOGRAPHIC, width=width, height=height, position=_f3(position),
                    look_at=_f3(look_at), up=_f3(up), min_x=min_x, max_x=max_x, min_y=min_y, max_y=max_y)


@dataclass
class Integrator:
    """Integrator{type='path-tracer'} settings plus the capped RR mode.

    termination 'max-length': max_depth = path_length; 'russian-roulette':
    rr_depth = path_length (integrators.lua:8-17 defaults: RR, 5). max_depth
    may be set together with RR for a hard cap.
    """
    spp: int = 16
    pixel_sampler: str = "jittered"
    termination: str = "russian-roulette"
    path_length: int = 5
    max_depth: Optional[int] = None
    seed: int = 0
    background: tuple = (0.0, 0.0, 0.0)

    def params(self) -> L.RenderParams:
        p = L.RenderParams()
        p.pixel_sampler = {"regular": L.SPT_PIXEL_REGULAR, "jittered": L.SPT_PIXEL_JITTERED,
                           "uniform": L.SPT_PIXEL_UNIFORM}[self.pixel_sampler]
        p.spp = self.spp
        if self.termination == "max-length":
            p.max_depth, p.rr_depth = self.path_length, -1
        elif self.termination == "russian-roulette":
            p.max_depth, p.rr_depth = -1, self.path_length
        else:
            raise ValueError("termination must be 'max-length' or 'russian-roulette'")
        if self.max_depth is not None:
            p.max_depth = self.max_depth
        p.seed = self.seed
        p.background = _f3(self.background)
        return p

    def effective_spp(self) -> int:
        if self.pixel_sampler == "uniform":
            return self.spp
        n = int(np.ceil(np.sqrt(self.spp)))
        return n * n


# ----------------------------------------------------------------------------
# tree
# ----------------------------------------------------------------------------

NODE_DTYPE = np.dtype([("lmin", "<f4", 3), ("lmax", "<f4", 3), ("rmin", "<f4", 3), ("rmax", "<f4", 3),
                       ("lref", "<u4"), ("lcnt", "<u4"), ("rref", "<u4"), ("rcnt", "<u4")])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("e1", "<f4", 3), ("e2", "<f4", 3), ("prim_id", "<u4"),
                      ("pad", "<u4", 2)])
assert NODE_DTYPE.itemsize == 64 and TRI_DTYPE.itemsize == 48


class Bvh:
    """Host SAH build over primitives (n, 9) float32 in scene order: triangles
    (v0 v1 v2), or, where kinds[i] == SPT_PRIM_SPHERE, spheres (centre xyz,
    radius, 5 unused floats)."""

    def __init__(self, verts: np.ndarray, n_threads: int = 8, kinds: Optional[np.ndarray] = None):
        self.lib = L.load()
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
        self.verts = v
        self.kinds = None if kinds is None else np.ascontiguousarray(kinds, np.uint8).reshape(-1)
        if self.kinds is not None and self.kinds.shape[0] != v.shape[0]:
            raise ValueError("one kind per primitive")
        h = C.c_void_p()
        s = self.lib.spt_bvh_build_prims(v.ctypes.data, _ptr(self.kinds), v.shape[0], n_threads, C.byref(h))
        if s != L.SPT_OK:
            raise SptError(s, "spt_bvh_build failed")
        self.h = h

    @property
    def info(self) -> dict:
        inf = L.BvhInfo()
        self.lib.spt_bvh_get_info(self.h, C.byref(inf))
        return {f: (list(getattr(inf, f)) if f in ("root_min", "root_max") else getattr(inf, f))
                for f, _ in L.BvhInfo._fields_}

    def export(self):
        inf = self.info
        nodes = np.zeros(inf["n_pair_nodes"], NODE_DTYPE)
        tris = np.zeros(inf["n_prims"], TRI_DTYPE)
        leaf = np.zeros(inf["n_prims"], np.uint32)
        box = np.zeros(6, np.float32)
        s = self.lib.spt_bvh_export(self.h, _ptr(nodes) if len(nodes) else None, _ptr(tris), _ptr(leaf), _ptr(box))
        if s != L.SPT_OK:
            raise SptError(s, "spt_bvh_export failed")
        return nodes, tris, leaf, box

    def wide(self):
        """The compressed four-wide tree the traversal reads (spt_bvh_wide):
        nodes (n, 16) uint32 and leaf records (m, 8) uint32 (float fields as
        their bits)."""
        nn, nl = C.c_uint64(0), C.c_uint64(0)
        s = self.lib.spt_bvh_wide(self.h, None, C.byref(nn), None, C.byref(nl))
        nodes = np.zeros((max(nn.value, 1), 16), np.uint32)
        leaves = np.zeros((max(nl.value, 1), 8), np.uint32)
        if s == L.SPT_OK:
            s = self.lib.spt_bvh_wide(self.h, nodes.ctypes.data, C.byref(nn), leaves.ctypes.data, C.byref(nl))
        if s != L.SPT_OK:
            raise SptError(s, "spt_bvh_wide failed")
        return nodes[:nn.value], leaves[:nl.value]

    def dump(self, path: str):
        s = self.lib.spt_bvh_dump(self.h, path.encode())
        if s != L.SPT_OK:
            raise SptError(s, "spt_bvh_dump failed")

    def close(self):
        if self.h:
            self.lib.spt_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------
# device
# ----------------------------------------------------------------------------

class DeviceArray:
    """Device allocation made through the C ABI (spt_malloc / spt_free)."""

    def __init__(self, dev: "Device", nbytes: int):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        dev._check(dev.lib.spt_malloc(dev.h, max(self.nbytes, 16), C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.dev._check(self.dev.lib.spt_memcpy_h2d(self.dev.h, self.ptr, a.ctypes.data, a.nbytes))
        self.dev.sync()

    def download(self, dtype, shape, out: Optional[np.ndarray] = None) -> np.ndarray:
        """Copies to a new host array, or into `out` (reusing one host buffer
        avoids mapping fresh pages for every copy in loops)."""
        if out is None:
            out = np.empty(shape, dtype)
        assert out.dtype == np.dtype(dtype) and out.flags.c_contiguous
        assert out.nbytes <= self.nbytes
        self.dev._check(self.dev.lib.spt_memcpy_d2h(self.dev.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        self.upload(np.zeros(self.nbytes, np.uint8))

    def free(self):
        if self.ptr:
            self.dev.lib.spt_free(self.dev.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Device:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        s = self.lib.spt_create(device, C.byref(h))
        if s != L.SPT_OK:
            raise SptError(s, f"spt_create({device}) failed (no HIP device?)")
        self.h = h

    def _check(self, s: int):
        if s != L.SPT_OK:
            raise SptError(s, self.lib.spt_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.spt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self.lib.spt_stream(self.h)

    def sync(self):
        self._check(self.lib.spt_sync(self.h))

    def event_record(self, slot: int):
        self._check(self.lib.spt_event_record(self.h, slot))

    def event_elapsed(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(self.lib.spt_event_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    def alloc(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        d = DeviceArray(self, a.nbytes)
        d.upload(a)
        return d

    # scene -------------------------------------------------------------
    def set_scene(self, bvh: Bvh, normals: Optional[np.ndarray] = None, mat_ids: Optional[np.ndarray] = None):
        n_prims = bvh.verts.shape[0]
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        m = None if mat_ids is None else np.ascontiguousarray(mat_ids, np.uint32).reshape(-1)
        if n is not None and n.shape[0] != n_prims:
            raise ValueError(f"normals has {n.shape[0]} rows for {n_prims} primitives")
        if m is not None and m.shape[0] != n_prims:
            raise ValueError(f"mat_ids has {m.shape[0]} entries for {n_prims} primitives")
        self._check(self.lib.spt_set_scene(self.h, bvh.h, bvh.verts.ctypes.data, _ptr(n), _ptr(m)))

    def set_materials(self, mats: Sequence[L.Material]):
        arr = (L.Material * len(mats))(*mats)
        self._check(self.lib.spt_set_materials(self.h, arr, len(mats)))

    # ray casting -------------------------------------------------------
    def trace(self, d_rays: int, n: int, d_hits: int, mode: int = L.SPT_TRACE_CLOSEST,
              d_tests: Optional[int] = None, d_isects: Optional[int] = None):
        self._check(self.lib.spt_trace(self.h, d_rays, n, d_hits, mode, d_tests, d_isects))

    def trace_host(self, rays: np.ndarray, mode: int = L.SPT_TRACE_CLOSEST):
        """rays (n, 8) float32 -> (t, prim_id[, tests, isects]) as numpy."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr = self.to_device(rays)
        dh = self.alloc(8 * n)
        counted = mode in (L.SPT_TRACE_COUNT, L.SPT_TRACE_FETCH)
        dt = self.alloc(4 * n) if counted else None
        di = self.alloc(4 * n) if counted else None
        self.trace(dr.ptr, n, dh.ptr, mode, dt.ptr if dt else None, di.ptr if di else None)
        self.sync()
        h = dh.download(np.uint32, (n, 2))
        t = h[:, 0].view(np.float32).copy()
        prim = h[:, 1].copy()
        if counted:
            return t, prim, dt.download(np.uint32, n), di.download(np.uint32, n)
        return t, prim

    def camera_rays(self, cam: L.Camera) -> np.ndarray:
        """Pixel-centre primary rays (h*w, 8) float32, computed on the device."""
        n = cam.width * cam.height
        d = self.alloc(32 * n)
        self._check(self.lib.spt_camera_rays(self.h, C.byref(cam), d.ptr))
        self.sync()
        out = d.download(np.float32, (n, 8))
        d.free()
        return out

    def raycast(self, cam: L.Camera, kind: int) -> np.ndarray:
        n = cam.width * cam.height
        comps = 3 if kind == L.SPT_RAYCAST_NORMAL else 1
        d = self.alloc(4 * n * comps)
        self._check(self.lib.spt_raycast(self.h, C.byref(cam), kind, d.ptr))
        self.sync()
        dt = np.uint32 if kind == L.SPT_RAYCAST_COUNT else np.float32
        out = d.download(dt, (cam.height, cam.width, comps) if comps == 3 else (cam.height, cam.width))
        d.free()
        return out

    # path tracing ------------------------------------------------------
    @staticmethod
    def call_paths(cam: L.Camera, params: L.RenderParams) -> int:
        """Paths of one spt_render call: viewport pixels x sample count, where
        the regular and jittered samplers round spp up to n^2, n = ceil(sqrt(spp))
        (jittered_pixel_sampler.cpp:6-8)."""
        spp = params.spp
        if params.pixel_sampler != L.SPT_PIXEL_UNIFORM:
            spp = int(np.ceil(np.sqrt(spp))) ** 2
        cnt = params.spp_count or max(spp - params.spp_begin, 0)
        vw, vh = params.viewport[2], params.viewport[3]
        if vw == 0:
            vw, vh = cam.width, cam.height
        return vw * vh * cnt

    def render(self, cam: L.Camera, params: L.RenderParams, d_rgb32=None, d_rgb64=None, d_path=None):
        """spt_render. Outputs are DeviceArrays (sizes checked here, since the
        C ABI takes bare pointers) or raw device addresses (unchecked)."""
        npix = cam.width * cam.height
        need = ((d_rgb32, npix * 12, "FP32 film"), (d_rgb64, npix * 24, "FP64 film"),
                (d_path, self.call_paths(cam, params) * 12, "per-path output"))
        ptrs = []
        for buf, nbytes, what in need:
            if isinstance(buf, DeviceArray):
                if buf.nbytes < nbytes:
                    raise ValueError(f"{what} needs {nbytes} bytes, buffer has {buf.nbytes}")
                buf = buf.ptr
            ptrs.append(buf)
        self._check(self.lib.spt_render(self.h, C.byref(cam), C.byref(params), *ptrs))

    # tuning ------------------------------------------------------------
    def set_tuning(self, **kw):
        t = L.Tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(self.lib.spt_set_tuning(self.h, C.byref(t)))

    def get_tuning(self) -> dict:
        t = L.Tuning()
        self._check(self.lib.spt_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    def autotune(self, cam: L.Camera, integ: "Integrator", sample_paths: int = 8 << 20) -> dict:
        p = integ.params()
        t = L.Tuning()
        self._check(self.lib.spt_autotune(self.h, C.byref(cam), C.byref(p), sample_paths, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    # learned guide -----------------------------------------------------
    def guide_init(self, seed: int = 1, lr: float = 5e-3, loss_scale: float = 65536.0, bsdf_fraction: float = 0.5,
                   batch: int = 262144, steps_per_pass: int = 1, flags: int = 0, train_every: int = 0):
        cfg = L.GuideConfig(seed=seed, learning_rate=lr, loss_scale=loss_scale, bsdf_fraction=bsdf_fraction,
                            batch=batch, steps_per_pass=steps_per_pass, flags=flags, train_every=train_every)
        self._check(self.lib.spt_guide_init(self.h, C.byref(cfg)))

    def guide_set_params(self, theta: np.ndarray):
        t = np.ascontiguousarray(theta, np.float32)
        assert t.size == L.SPT_GUIDE_PARAMS
        self._check(self.lib.spt_guide_set_params(self.h, t.ctypes.data))

    def guide_get_params(self) -> np.ndarray:
        t = np.zeros(L.SPT_GUIDE_PARAMS, np.float32)
        self._check(self.lib.spt_guide_get_params(self.h, t.ctypes.data))
        return t

    def guide_eval(self, queries: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        dq = self.to_device(q)
        do = self.alloc(q.shape[0] * 40 * 4)
        self._check(self.lib.spt_guide_eval(self.h, dq.ptr, q.shape[0], do.ptr))
        self.sync()
        return do.download(np.float32, (q.shape[0], 40))

    def _records(self, records: np.ndarray) -> DeviceArray:
        r = np.ascontiguousarray(records, np.float32).reshape(-1, L.SPT_GUIDE_RECORD_FLOATS)
        return self.to_device(r), r.shape[0]

    def guide_gradient(self, records: np.ndarray):
        dr, n = self._records(records)
        g = np.zeros(L.SPT_GUIDE_PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.lib.spt_guide_gradient(self.h, dr.ptr, n, g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def guide_train(self, records: np.ndarray) -> float:
        dr, n = self._records(records)
        loss = C.c_float()
        self._check(self.lib.spt_guide_train(self.h, dr.ptr, n, C.byref(loss)))
        return loss.value

    def guide_train_collected(self, wait: bool = True) -> Optional[float]:
        """Queues the training steps; with wait, returns the last step's loss."""
        if not wait:
            self._check(self.lib.spt_guide_train_collected(self.h, None))
            return None
        loss = C.c_float()
        self._check(self.lib.spt_guide_train_collected(self.h, C.byref(loss)))
        return loss.value

    def guide_stats(self) -> dict:
        st = L.GuideStats()
        self._check(self.lib.spt_guide_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in L.GuideStats._fields_}

    def guide_records(self) -> np.ndarray:
        p, n = C.c_void_p(), C.c_uint32()
        self._check(self.lib.spt_guide_records(self.h, C.byref(p), C.byref(n)))
        out = np.empty((n.value, L.SPT_GUIDE_RECORD_FLOATS), np.float32)
        if n.value:
            self._check(self.lib.spt_memcpy_d2h(self.h, out.ctypes.data, p.value, out.nbytes))
        return out

    def render_stats(self) -> dict:
        st = L.RenderStats()
        self._check(self.lib.spt_render_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in L.RenderStats._fields_}

    def render_image(self, cam: L.Camera, integ: Integrator, film64: bool = False, per_path: bool = False,
                     spp_begin: int = 0, spp_count: int = 0, viewport=None, stream: Optional[np.ndarray] = None,
                     extra_flags: int = 0):
        """Renders and returns (image (h, w, 3) as the reference's pixel values,
        per-path radiance or None). Host convenience for tests and tools."""
        p = integ.params()
        spp = integ.effective_spp()
        cnt = spp_count or (spp - spp_begin)
        p.spp_begin, p.spp_count = spp_begin, cnt
        if viewport is not None:
            p.viewport = (C.c_uint32 * 4)(*viewport)
            vw, vh = viewport[2], viewport[3]
        else:
            vw, vh = cam.width, cam.height
        npix = cam.width * cam.height
        film = self.alloc(npix * 3 * (8 if film64 else 4))
        film.zero()
        p.flags = (L.SPT_RENDER_FILM64 if film64 else L.SPT_RENDER_FILM32) | extra_flags
        dpath = None
        if per_path:
            dpath = self.alloc(vw * vh * cnt * 12)
            p.flags |= L.SPT_RENDER_PER_PATH
        dstream = None
        if stream is not None:
            stream = np.ascontiguousarray(stream, np.float32)
            if stream.shape[0] < vw * vh * cnt:
                raise ValueError("replay stream needs one row per path")
            dstream = self.to_device(stream)
            p.rng = L.SPT_RNG_STREAM
            p.stream_stride = stream.shape[1]
            p.d_stream = dstream.ptr
        self.render(cam, p, None if film64 else film, film if film64 else None, dpath)
        self.sync()
        img = film.download(np.float64 if film64 else np.float32, (cam.height, cam.width, 3))
        img = img / (np.float64(cnt) if film64 else np.float32(cnt))
        paths = dpath.download(np.float32, (vw * vh * cnt, 3)) if dpath else None
        return img, paths


class Batch:
    """Many scenes rendered by one wavefront (spt_batch_*). `scenes` are
    SceneData-like objects (verts, normals, mat_ids, materials, camera,
    integrator). Each scene gets an FP32 film, and per-path radiance when
    per_path is True or, given a collection of scene indices, for those
    scenes only."""

    def __init__(self, dev: Device, scenes, per_path=False, film64: bool = False, n_threads: int = 8):
        self.dev, self.lib = dev, dev.lib
        self.scenes = list(scenes)
        self.bvhs, self.keep, self.films, self.films64, self.paths = [], [], [], [], []
        arr = (L.BatchScene * len(self.scenes))()
        for i, sc in enumerate(self.scenes):
            bvh = Bvh(sc.verts, n_threads=n_threads, kinds=getattr(sc, "kinds", None))
            self.bvhs.append(bvh)
            nrm = None if sc.normals is None else np.ascontiguousarray(sc.normals, np.float32).reshape(-1, 9)
            mid = np.ascontiguousarray(sc.mat_ids, np.uint32).reshape(-1)
            if mid.shape[0] != bvh.verts.shape[0] or (nrm is not None and nrm.shape[0] != bvh.verts.shape[0]):
                raise ValueError(f"scene {i}: mat_ids/normals do not have one row per primitive")
            mats = (L.Material * len(sc.materials))(*sc.materials)
            self.keep += [nrm, mid, mats]
            cam, p = sc.camera, sc.integrator.params()
            npix = cam.width * cam.height
            film = dev.alloc(npix * 12)
            film.zero()
            self.films.append(film)
            p.flags = L.SPT_RENDER_FILM32
            f64 = None
            if film64:
                f64 = dev.alloc(npix * 24)
                f64.zero()
                p.flags |= L.SPT_RENDER_FILM64
            self.films64.append(f64)
            pp = None
            if per_path is True or (not isinstance(per_path, bool) and i in per_path):
                pp = dev.alloc(npix * sc.integrator.effective_spp() * 12)
                p.flags |= L.SPT_RENDER_PER_PATH
            self.paths.append(pp)
            e = arr[i]
            e.bvh, e.verts, e.normals, e.mat_ids = bvh.h, bvh.verts.ctypes.data, _ptr(nrm), mid.ctypes.data
            e.mats, e.n_mats = mats, len(sc.materials)
            e.camera, e.params = cam, p
            e.d_rgb32 = film.ptr
            e.d_rgb64 = f64.ptr if f64 is not None else None
            e.d_path_rgb = pp.ptr if pp is not None else None
        h = C.c_void_p()
        dev._check(self.lib.spt_batch_create(dev.h, arr, len(self.scenes), C.byref(h)))
        self.h = h

    def render(self):
        self.dev._check(self.lib.spt_batch_render(self.dev.h, self.h))

    def autotune(self, sample_div: int = 16) -> dict:
        t = L.Tuning()
        self.dev._check(self.lib.spt_batch_autotune(self.dev.h, self.h, sample_div, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    def stats(self):
        st = L.BatchStats()
        ms = np.zeros(len(self.scenes), np.float64)
        self.dev._check(self.lib.spt_batch_get_stats(self.dev.h, self.h, C.byref(st), ms.ctypes.data))
        return {f: getattr(st, f) for f, _ in L.BatchStats._fields_}, ms

    def clear(self):
        for f in self.films + [f for f in self.films64 if f is not None]:
            f.zero()

    def image(self, i: int) -> np.ndarray:
        cam = self.scenes[i].camera
        return self.films[i].download(np.float32, (cam.height, cam.width, 3))

    def path_radiance(self, i: int) -> np.ndarray:
        sc = self.scenes[i]
        n = sc.camera.width * sc.camera.height * sc.integrator.effective_spp()
        return self.paths[i].download(np.float32, (n, 3))

    def close(self):
        if getattr(self, "h", None):
            self.lib.spt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
