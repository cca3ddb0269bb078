"""Scene scripts without liblua (spt_lua_load; lua_bind/lua_bind.cpp,
core/swpathtracer_app.cpp:40-262, lua_core/*.lua).

Pins:
- the reference's four sample scripts run unmodified, once through the
  reference's own lua_core files (interpreted) and once through the built-in
  restatement of those files: the two registries (_G.__RENDERER_ELEMENTS__)
  and the extracted scenes are identical, so the built-in core carries the
  same defaults, validation and registry layout;
- what the samples describe, read off the scripts by hand (camera, sampler,
  integrator, primitive and material counts, the furnace's spheres);
- math.random() equals glibc random() / 2^31 from srandom(1), which is Lua
  5.3's math.random on Linux with no seed (samples/procedural.lua draws the
  sphere colours with it);
- the monkey mesh (samples/meshes/monkey.obj) expands to the same triangles
  and normals through the library's OBJ reader as through host_io.load_obj;
- error behaviour: each lua_core validation message, lua_bind's errors;
- interpreter semantics on a script whose printed output is worked out by
  hand from the Lua 5.3 manual.
"""
import ctypes
import os

import numpy as np
import pytest

from swpathtracer_amd import host_io
from swpathtracer_amd import lua_scene as LS
import swpathtracer_amd._lib as L

REF = "/root/reference"
HAVE_REF = os.path.isdir(os.path.join(REF, "lua_core"))
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="the reference tree is not present")
BUILD = os.path.join(REF, "build")  # the reference's app runs from build/ (README.md:74)
SAMPLES = ["furnace_test", "procedural", "materials_test", "monkey_in_a_box"]
HERE = os.path.dirname(os.path.abspath(__file__))


def _write(tmp_path, text, name="scene.lua"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@needs_ref
@pytest.mark.parametrize("name", SAMPLES)
def test_samples_interpreted_core_equals_builtin_core(name):
    a = LS.load(f"../samples/{name}.lua", workdir=BUILD)
    b = LS.load(f"../samples/{name}.lua", workdir=BUILD, builtin_core=True)
    assert a.registry == b.registry
    assert bytes(a.info) == bytes(b.info)
    for f in ("verts", "mat_ids"):
        assert np.array_equal(getattr(a.scene, f), getattr(b.scene, f))
    assert (a.scene.kinds is None) == (b.scene.kinds is None)
    assert [bytes(m) for m in a.scene.materials] == [bytes(m) for m in b.scene.materials]
    assert a.printed == b.printed == ""


@needs_ref
def test_furnace_sample_contents():
    s = LS.load("../samples/furnace_test.lua", workdir=BUILD)
    cam = s.scene.camera
    assert cam.type == L.SPT_CAMERA_ORTHOGRAPHIC and (cam.width, cam.height) == (512, 512)
    assert (cam.min_x, cam.max_x, cam.min_y, cam.max_y) == (-1.25, 1.25, -1.25, 1.25)
    assert list(cam.position) == [0, 0, 6] and list(cam.up) == [0, 1, 0]
    it = s.scene.integrator
    assert (it.spp, it.pixel_sampler, it.termination, it.path_length) == (256, "jittered", "russian-roulette", 3)
    assert s.viewport == (0, 0, 512, 512) and s.file_name == "furnace_test.ppm"
    assert s.accel == L.SPT_ACCEL_BVH_SAH
    # two spheres: the radius-10 light (material 1) and the unit sphere with
    # mat_cook_torrance, the fifth Material{} of the script
    assert s.scene.n_tris == 2 and list(s.scene.kinds) == [L.SPT_PRIM_SPHERE] * 2
    assert s.scene.verts[0, 3] == 10.0 and s.scene.verts[1, 3] == 1.0
    assert list(s.scene.mat_ids) == [0, 4]
    mats = s.scene.materials
    assert len(mats) == 7
    assert list(mats[0].emission) == pytest.approx([0.5, 0.5, 0.5])
    ct = mats[4].layers[0]
    assert mats[4].n_layers == 1 and ct.type == L.SPT_BSDF_COOK_TORRANCE
    assert ct.roughness == pytest.approx(0.2) and list(ct.r0) == pytest.approx([0.999] * 3)
    coat = mats[6]  # mat_coating: dielectric over Cook-Torrance
    assert coat.n_layers == 2 and coat.layers[0].type == L.SPT_BSDF_SMOOTH_DIELECTRIC
    assert coat.layers[0].ior == pytest.approx(1.49) and coat.layers[1].type == L.SPT_BSDF_COOK_TORRANCE


@needs_ref
def test_procedural_sample_draws_lua53_math_random():
    s = LS.load("../samples/procedural.lua", workdir=BUILD)
    # 2 light triangles, then 10 x 10 spheres with one material each
    assert s.scene.n_tris == 102 and len(s.scene.materials) == 101
    assert list(s.scene.kinds[:2]) == [0, 0] and np.all(s.scene.kinds[2:] == L.SPT_PRIM_SPHERE)
    libc = ctypes.CDLL("libc.so.6")
    libc.random.restype = ctypes.c_long
    libc.srandom(1)
    draws = np.array([libc.random() for _ in range(300)], np.float64) / 2.0 ** 31
    kd = np.array([list(m.layers[1].kd) for m in s.scene.materials[1:]], np.float64)
    assert np.array_equal(kd, draws.reshape(100, 3).astype(np.float32).astype(np.float64))
    # the registry keeps the doubles themselves
    b3 = s.registry["__BSDFS__"]["#000000000003"]["kd"]
    assert [b3["#000000000001"], b3["#000000000002"], b3["#000000000003"]] == list(draws[:3])
    # centres from the nested numeric for loops
    c = s.scene.verts[2:, :3]
    assert np.allclose(c[:, 0], np.repeat(np.arange(1, 11) - 5.5, 10))
    assert np.allclose(c[:, 2], np.tile(np.arange(1, 11) - 5.5, 10))


@needs_ref
def test_monkey_mesh_expansion_matches_host_reader():
    s = LS.load("../samples/monkey_in_a_box.lua", workdir=BUILD)
    assert s.scene.n_tris == 12 + 968
    mesh = host_io.load_obj(os.path.join(REF, "samples", "meshes", "monkey.obj"), default_material_id=0)
    assert mesh.n_tris == 968
    assert np.array_equal(s.scene.verts[12:], mesh.verts)
    assert np.array_equal(s.scene.normals[12:], mesh.normals)
    assert np.all(s.scene.mat_ids[12:] == 0)  # mat_white
    # every monkey face carries its vn normals (s off: per-face normals)
    assert np.all(np.any(s.scene.normals[12:] != 0, axis=1))
    assert np.all(s.scene.normals[:12] == 0)


def test_builtin_core_scene_from_this_repository():
    s = LS.load("spheres_in_box.lua", workdir=os.path.join(HERE, "scenes"))
    assert s.scene.n_tris == 12 + 3 + 1 + 5
    assert list(s.scene.kinds) == [0] * 12 + [1, 1, 1, 0] + [1] * 5
    assert list(s.scene.mat_ids[-5:]) == [5, 3, 5, 3, 5]  # (i % 2 == 0) and copper or plastic
    cam = s.scene.camera
    assert cam.type == L.SPT_CAMERA_PINHOLE and cam.fov_deg == 45 and list(cam.up) == [0, 1, 0]
    assert s.viewport == (0, 0, 48, 36)
    pl = s.scene.materials[5]
    assert pl.n_layers == 2 and pl.layers[1].sampler == L.SPT_SAMPLER_UNIFORM
    assert s.scene.normals is not None and np.any(s.scene.normals[15] != 0)


BASE = """
require "core"
Camera{ type = "pinhole" }
ImageBuffer{ width = 8, height = 8, viewport_width = 8, viewport_height = 8 }
GlobalSettings{}
Integrator{ type = "path-tracer" }
PixelSampler{}
local m = Material{ bsdf = LayeredBSDF{ Lambertian{} }, emission = Emission{0, 0, 0} }
Triangle{ v1 = {0, 0, 0}, v2 = {1, 0, 0}, v3 = {0, 1, 0}, material = m }
"""


def test_lua_core_defaults(tmp_path):
    s = LS.load(_write(tmp_path, BASE))
    cam = s.scene.camera
    assert list(cam.position) == [0, 0, 5] and list(cam.look_at) == [0, 0, 0] and cam.fov_deg == 55
    it = s.scene.integrator
    assert (it.spp, it.pixel_sampler, it.termination, it.path_length) == (16, "jittered", "russian-roulette", 5)
    lam = s.scene.materials[0].layers[0]
    assert list(lam.kd) == pytest.approx([0.75] * 3) and lam.sampler == L.SPT_SAMPLER_IMPORTANCE
    assert s.file_name == "image.ppm" and s.viewport == (0, 0, 8, 8)
    r = s.registry
    assert r["__GLOBAL_SETTINGS__"]["acceleration_structure"] == "bvh-sah"
    assert r["__GLOBAL_SETTINGS__"]["overlap_threshold"] == "inf"
    assert r["__OBJECT_COUNT__"] == 1 and r["__BSDF_COUNT__"] == 1 and r["__LAYERED_BSDF_COUNT__"] == 1
    assert r["__LAYERED_BSDFS__"]["#000000000001"]["num_layers"] == 1


@pytest.mark.parametrize("edit, message", [
    ('Camera{ type = "fisheye" }', "Lua ERROR: Invalid camera type. Exiting..."),
    ('Camera{ position = {1, 2} }', "Lua ERROR: Invalid camera position vector. Exiting..."),
    ('Camera{ type = "orthographic" }', "Lua ERROR: Invalid camera min_x. Exiting..."),
    ('PixelSampler{ type = "stratified" }', "Lua ERROR: Invalid pixel sampler type. Exiting..."),
    ('ImageBuffer{ width = 4 }', "Lua ERROR: Invalid image buffer 'viewport_width'. Exiting..."),
    ('GlobalSettings{ acceleration_structure = "sbvh-sah" }', "Lua ERROR: Invalid acceleration structure. Exiting..."),
    ('Integrator{ type = "path-tracer", path_termination = "never" }',
     "Lua ERROR: Invalid integrator path termination criterion. Exiting..."),
    ('Lambertian{ kd = {1, 1, "x"} }', "Lua ERROR: Invalid lambertian reflectance (kd field). Exiting..."),
    ('CookTorrance{ bsdf_sampler = "mis" }', "Lua ERROR: Invalid Cook-Torrance BSDF sampler. Exiting..."),
    ('LayeredBSDF{ 7 }', "Lua ERROR: Invalid layered BSDF layer: index out of range. Exiting..."),
    ('Emission{ 1, 2 }', "Lua ERROR: Invalid emission value. Exiting..."),
    ('Material{ bsdf = 1, emission = 9 }', "Lua ERROR: Invalid material emission: index out of range. Exiting..."),
    ('Sphere{ center = {0, 0, 0}, radius = "big" }', "Lua ERROR: Invalid sphere radius. Exiting..."),
    ('Mesh{ material = 1 }', "Lua ERROR: Invalid mesh file name. Exiting..."),
])
def test_validation_messages(tmp_path, edit, message):
    with pytest.raises(LS.LuaSceneError) as e:
        LS.load(_write(tmp_path, BASE + edit + "\n"))
    assert str(e.value) == message


def test_lua_bind_errors(tmp_path):
    # no ImageBuffer{}: lua_bind.cpp:49-50 throws
    text = BASE.replace("ImageBuffer{ width = 8, height = 8, viewport_width = 8, viewport_height = 8 }", "")
    with pytest.raises(LS.LuaSceneError, match="Invalid image buffer definition"):
        LS.load(_write(tmp_path, text))
    # a Mesh{} without render_submeshes (objects.lua does not default it): lua_bind.cpp:425-426
    with pytest.raises(LS.LuaSceneError, match="Invalid object definiton"):
        LS.load(_write(tmp_path, BASE + 'Mesh{ filename = "x.obj", material = 1 }\n'))
    # syntax errors carry chunk and line
    with pytest.raises(LS.LuaSceneError, match=r"scene.lua:3:"):
        LS.load(_write(tmp_path, "local a = 1\nlocal b = 2\nlocal = 3\n"))
    # a script that is not there
    with pytest.raises(LS.LuaSceneError, match="cannot open"):
        LS.load(str(tmp_path / "missing.lua"))


def test_mesh_with_sub_meshes(tmp_path):
    (tmp_path / "two.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvn 0 0 1\n"
        "o first\nf 1//1 2//1 3//1\no second\nf 2//1 4//1 3//1\nf 1 2 4 3\n")
    text = BASE + """
local m2 = Material{ bsdf = LayeredBSDF{ Lambertian{ kd = {0.1, 0.2, 0.3} } }, emission = Emission{0, 0, 0} }
Mesh{ filename = "two.obj", material = m, render_submeshes = "none",
      sub_meshes = { { name = "second", material = m2 } } }
Mesh{ filename = "two.obj", material = m, render_submeshes = "all" }
"""
    s = LS.load(_write(tmp_path, text), workdir=str(tmp_path))
    # triangle, then "second" only (1 + 2 fan triangles) with m2, then all (1 + 3) with m
    assert s.scene.n_tris == 1 + 3 + 4
    assert list(s.scene.mat_ids) == [0, 1, 1, 1, 0, 0, 0, 0]
    assert np.any(s.scene.normals[1] != 0) and np.all(s.scene.normals[2] == 0)


SEMANTICS = r"""
local t = {10, 20, 30, x = "a", [5] = 50}
print(#t, t[5], t.x)
local s = 0
for i = 1, 10, 3 do s = s + i end
print(s)
for i = 3, 1, -1 do io_ = (io_ or "") .. i end
print(io_)
local function fib(n) if n < 2 then return n end return fib(n - 1) + fib(n - 2) end
print(fib(15))
local function counter()
    local c = 0
    return function() c = c + 1; return c end
end
local c1, c2 = counter(), counter()
c1(); c1()
print(c1(), c2())
print(2 ^ 3 ^ 2, -2 ^ 2, 7 // 2, -7 // 2, 7 % -3, -7 % 3, 1 / 0 > 0)
print("a" .. "b" .. 1 .. 2, 10 == 10.0, "10" + 5, #"hello")
print(1 and 2, nil and 1, false or "z", not nil, nil == false)
local n = 0
while true do n = n + 1; if n > 4 then break end end
repeat n = n - 1 until n == 0
print(n)
local keys = {}
for k, v in pairs({3, 4, 5}) do keys[#keys + 1] = k * v end
print(table.concat(keys, ","))
local acc = {}
for i, v in ipairs({"p", "q", nil, "r"}) do acc[i] = v end
print(#acc, select("#", 1, 2, 3), select(2, "a", "b", "c"))
print(string.format("%5.2f|%d|%s|%x", 3.14159, 42, "s", 255), tostring(nil), type(print))
print(math.floor(-1.5), math.max(3, 9, 1), math.huge == 1 / 0, tonumber("0x10"), tonumber("z"))
local obj = {v = 2}
function obj:twice(k) return self.v * k end
print(obj:twice(21), 3 / 2, 4 / 2, 2^53 // 1, 10 // 3.0, math.type(1), math.type(1.0), 1e3)
local function va(...) local a, b = ... return select("#", ...), a, b end
print(va(7, 8, 9))
print(("abc"):upper(), string.sub("hello", 2, -2), string.rep("ab", 3))
--[[ a long
comment ]] print([[long
string]])
local tt = {a = 1, b = 2, c = 3, 10, 20}
local cnt = 0
for k, v in pairs(tt) do tt[k] = nil; cnt = cnt + 1 end
print(cnt, next(tt))
"""

EXPECTED = """3\t50\ta
22
321
610
3\t1
512.0\t-4.0\t3\t-4\t-2\t2\ttrue
ab12\ttrue\t15\t5
2\tnil\tz\ttrue\tfalse
0
3,8,15
2\t3\tb\tc
 3.14|42|s|ff\tnil\tfunction
-2\t9\ttrue\t16\tnil
42\t1.5\t2.0\t9.007199254741e+15\t3.0\tinteger\tfloat\t1000.0
3\t7\t8
ABC\tell\tababab
long
string
5\tnil
"""


def test_interpreter_semantics(tmp_path):
    # one Triangle so the scene extraction has what it needs; the output is the point
    s = LS.load(_write(tmp_path, SEMANTICS + BASE))
    assert s.printed == EXPECTED


def test_glibc_random_sequence_matches_libc(tmp_path):
    text = BASE + """
math.randomseed(1234)
local out = {}
for i = 1, 5 do out[i] = string.format("%.17g", math.random()) end
print(table.concat(out, " "))
print(math.random(6), math.random(10, 20))
"""
    s = LS.load(_write(tmp_path, text))
    lines = s.printed.strip().split("\n")
    libc = ctypes.CDLL("libc.so.6")
    libc.random.restype = ctypes.c_long
    libc.srandom(1234)
    libc.random()  # math.randomseed discards one draw
    want = [libc.random() / 2.0 ** 31 for _ in range(5)]
    assert [float(x) for x in lines[0].split()] == want
    r6 = int(np.floor(libc.random() / 2.0 ** 31 * 6)) + 1
    r10 = int(np.floor(libc.random() / 2.0 ** 31 * 11)) + 10
    assert lines[1].split("\t") == [str(r6), str(r10)]
