"""The reference's Lua scene scripts through spt_lua_load (the library's own
interpreter of the scene-script subset of Lua 5.3 plus lua_bind's extraction).

``load(script, workdir)`` stands for ``lb::LuaBind::runScript`` and the
scene assembly of ``SwpathtracerApp`` (lua_bind/lua_bind.cpp:18-31,
core/swpathtracer_app.cpp:40-262): it returns the primitives in scene order
(meshes expanded as Scene::loadMesh adds them), the material table, the
camera, the integrator settings and the image buffer's viewport and file
name. Errors raise ``LuaSceneError`` with the reference's message (its
validation prints "Lua ERROR: ..." and exits; lua_bind throws).
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib as L
from . import api
from .scenes import SceneData


class LuaSceneError(RuntimeError):
    pass


INTEGRATOR_NAMES = {L.SPT_LUA_NONE: "none", L.SPT_LUA_PATHTRACER: "path-tracer",
                    L.SPT_LUA_RAYCAST_NORMAL: "normal", L.SPT_LUA_RAYCAST_DEPTH: "depth",
                    L.SPT_LUA_RAYCAST_COUNT: "intersection-test-count"}
RAYCAST_KIND = {L.SPT_LUA_RAYCAST_NORMAL: L.SPT_RAYCAST_NORMAL, L.SPT_LUA_RAYCAST_DEPTH: L.SPT_RAYCAST_DEPTH,
                L.SPT_LUA_RAYCAST_COUNT: L.SPT_RAYCAST_COUNT}


@dataclass
class LuaScene:
    scene: SceneData               # verts/normals/mat_ids/kinds/materials/camera/integrator
    integrator_kind: str           # "path-tracer", "normal", "depth", "intersection-test-count"
    viewport: Tuple[int, int, int, int]  # top, left, width, height
    file_name: str
    accel: int                     # SPT_ACCEL_*
    value_range: Tuple[float, float]  # depth / count ray caster range
    registry: dict                 # _G.__RENDERER_ELEMENTS__ after the script ran
    printed: str
    info: L.LuaInfo

    @property
    def raycast_kind(self) -> Optional[int]:
        return RAYCAST_KIND.get(self.info.integrator)


def _text(lib, fn, h) -> str:
    n = C.c_uint64(0)
    fn(h, None, C.byref(n))
    buf = C.create_string_buffer(n.value)
    n2 = C.c_uint64(n.value)
    if fn(h, buf, C.byref(n2)) != L.SPT_OK:
        raise LuaSceneError("could not read back the scene text")
    return buf.value.decode("utf-8", "replace")


def load(script: str, workdir: Optional[str] = None, builtin_core: bool = False) -> LuaScene:
    lib = L.load()
    h = C.c_void_p()
    err = C.create_string_buffer(4096)
    flags = L.SPT_LUA_BUILTIN_CORE if builtin_core else 0
    s = lib.spt_lua_load(script.encode(), workdir.encode() if workdir else None, flags, C.byref(h), err, len(err))
    if s != L.SPT_OK:
        raise LuaSceneError(err.value.decode("utf-8", "replace"))
    try:
        info = L.LuaInfo()
        lib.spt_lua_get_info(h, C.byref(info))
        n = info.n_prims
        prims = np.zeros((n, 9), np.float32)
        kinds = np.zeros(n, np.uint8)
        normals = np.zeros((n, 9), np.float32)
        mat_ids = np.zeros(n, np.uint32)
        if n:
            lib.spt_lua_get_prims(h, prims.ctypes.data, kinds.ctypes.data, normals.ctypes.data, mat_ids.ctypes.data)
        mats = (L.Material * max(1, info.n_materials))()
        lib.spt_lua_get_materials(h, mats)
        registry = json.loads(_text(lib, lib.spt_lua_registry_json, h))
        printed = _text(lib, lib.spt_lua_printed, h)
    finally:
        lib.spt_lua_destroy(h)
    p = info.params
    integ = None
    if info.integrator == L.SPT_LUA_PATHTRACER:
        sampler = {L.SPT_PIXEL_REGULAR: "regular", L.SPT_PIXEL_JITTERED: "jittered",
                   L.SPT_PIXEL_UNIFORM: "uniform"}[p.pixel_sampler]
        integ = api.Integrator(spp=int(p.spp), pixel_sampler=sampler,
                               termination="max-length" if p.max_depth >= 0 else "russian-roulette",
                               path_length=int(info.path_length), seed=0, background=tuple(p.background))
    sc = SceneData(prims, normals if np.any(normals) else None, mat_ids, list(mats)[: info.n_materials],
                   camera=info.camera, integrator=integ, name=info.file_name.decode(),
                   kinds=kinds if np.any(kinds) else None)
    return LuaScene(scene=sc, integrator_kind=INTEGRATOR_NAMES.get(info.integrator, "none"),
                    viewport=tuple(int(v) for v in p.viewport), file_name=info.file_name.decode(),
                    accel=int(info.accel), value_range=(float(info.min_value), float(info.max_value)),
                    registry=registry, printed=printed, info=info)
