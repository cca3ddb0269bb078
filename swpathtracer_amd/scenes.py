"""Synthetic, seeded scenes for the benchmark configurations and tests.

Every scene is plain arrays in the reference's primitive order (prim id =
index, core/scene.cpp:26-30): triangles (n, 9) float32, per-vertex normals
(n, 9) float32 (all-zero rows = flat Triangle, primitives/triangle.cpp:5-6;
otherwise SmoothTriangle, primitives/smooth_triangle.h), material ids and a
material table. Nothing here is timed: scene generation is host set-up.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import api
from . import _lib as L


@dataclass
class SceneData:
    verts: np.ndarray                 # (n, 9) float32
    normals: Optional[np.ndarray]     # (n, 9) float32 or None (all flat)
    mat_ids: np.ndarray               # (n,) uint32
    materials: List[L.Material]
    camera: Optional[L.Camera] = None
    integrator: Optional[api.Integrator] = None
    name: str = ""
    kinds: Optional[np.ndarray] = None  # (n,) uint8 SPT_PRIM_* per primitive; None: all triangles

    @property
    def n_tris(self) -> int:
        return self.verts.shape[0]


def concat(parts: List[SceneData], materials: List[L.Material], name="") -> SceneData:
    verts = np.concatenate([p.verts for p in parts]).astype(np.float32)
    if any(p.normals is not None for p in parts):
        normals = np.concatenate([p.normals if p.normals is not None else np.zeros_like(p.verts) for p in parts])
    else:
        normals = None
    mats = np.concatenate([p.mat_ids for p in parts]).astype(np.uint32)
    kinds = None
    if any(p.kinds is not None for p in parts):
        kinds = np.concatenate([p.kinds if p.kinds is not None else np.zeros(p.n_tris, np.uint8) for p in parts])
    return SceneData(verts, normals, mats, materials, name=name, kinds=kinds)


def spheres(centres_radii, mats) -> SceneData:
    """Sphere primitives (primitives/sphere.cpp): rows (cx, cy, cz, r) as 9-float
    primitive rows with kind SPT_PRIM_SPHERE."""
    cr = np.asarray(centres_radii, np.float32).reshape(-1, 4)
    v = np.zeros((cr.shape[0], 9), np.float32)
    v[:, 0:4] = cr
    m = np.broadcast_to(np.asarray(mats, np.uint32).reshape(-1), (cr.shape[0],)).copy()
    return SceneData(v, None, m, [],
                     kinds=np.full(cr.shape[0], L.SPT_PRIM_SPHERE, np.uint8))


# ----------------------------------------------------------------------------
# geometry helpers
# ----------------------------------------------------------------------------

def quad(p0, p1, p2, p3, normal, mat: int) -> SceneData:
    """Two triangles (p0 p1 p2), (p0 p2 p3) wound so cross(e1, e2) points along `normal`."""
    p = [np.asarray(x, np.float64) for x in (p0, p1, p2, p3)]
    n = np.cross(p[1] - p[0], p[2] - p[0])
    if np.dot(n, normal) < 0:
        p = [p[0], p[3], p[2], p[1]]
    v = np.array([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])], np.float32)
    return SceneData(v, None, np.full(2, mat, np.uint32), [])


def box_walls(lo, hi, white: int, red: int, green: int, front: bool = False) -> List[SceneData]:
    """Closed-on-five-sides axis-aligned box, inward-facing (10 triangles; 12 with front)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    parts = [
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, 1, 0), white),   # floor
        quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1), (0, -1, 0), white),  # ceiling
        quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (0, 0, 1), white),   # back
        quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1), (1, 0, 0), red),     # left
        quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (-1, 0, 0), green),  # right
    ]
    if front:
        parts.append(quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, -1), white))
    return parts


def icosphere(subdiv: int, center=(0, 0, 0), radius=1.0, rotation: Optional[np.ndarray] = None,
              mat: int = 0, smooth: bool = False) -> SceneData:
    """Icosahedron subdivided `subdiv` times: 20 * 4**subdiv outward-wound triangles."""
    t = (1.0 + 5 ** 0.5) / 2.0
    vs = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
          (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in vs]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(verts)
    F = np.array(faces, np.int64)
    for _ in range(subdiv):
        cache = {}
        newv = [V]
        nv = V.shape[0]
        extra = []

        def mid(a, b):
            nonlocal nv
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                extra.append(m / np.linalg.norm(m))
                cache[key] = nv
                nv += 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        V = np.concatenate([V, np.array(extra)])
        F = np.array(nf, np.int64)
    if rotation is not None:
        V = V @ rotation.T
    # outward winding
    tri = V[F]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    flip = np.einsum("ij,ij->i", n, tri.sum(1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    P = V * radius + np.asarray(center, np.float64)
    verts = P[F].reshape(-1, 9).astype(np.float32)
    normals = V[F].reshape(-1, 9).astype(np.float32) if smooth else None
    return SceneData(verts, normals, np.full(F.shape[0], mat, np.uint32), [])


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# Improved Perlin noise (Perlin, "Improving Noise", SIGGRAPH 2002).
def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def _grad(h, x, y, z):
    h = h & 15
    u = np.where(h < 8, x, y)
    v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
    return np.where((h & 1) == 0, u, -u) + np.where((h & 2) == 0, v, -v)


def perlin3(p: np.ndarray, seed: int) -> np.ndarray:
    perm = np.random.default_rng(seed).permutation(256)
    P = np.concatenate([perm, perm]).astype(np.int64)
    fl = np.floor(p)
    X, Y, Z = [(fl[:, i].astype(np.int64) & 255) for i in range(3)]
    x, y, z = [p[:, i] - fl[:, i] for i in range(3)]
    u, v, w = _fade(x), _fade(y), _fade(z)
    A = P[X] + Y
    AA, AB = P[A] + Z, P[A + 1] + Z
    B = P[X + 1] + Y
    BA, BB = P[B] + Z, P[B + 1] + Z
    lerp = lambda t, a, b: a + t * (b - a)
    return lerp(w, lerp(v, lerp(u, _grad(P[AA], x, y, z), _grad(P[BA], x - 1, y, z)),
                        lerp(u, _grad(P[AB], x, y - 1, z), _grad(P[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, _grad(P[AA + 1], x, y, z - 1), _grad(P[BA + 1], x - 1, y, z - 1)),
                     lerp(u, _grad(P[AB + 1], x, y - 1, z - 1), _grad(P[BB + 1], x - 1, y - 1, z - 1))))


def displaced_sphere(rings: int = 512, segments: int = 1024, amp: float = 0.1, freq: float = 4.0,
                     seed: int = 2, center=(0, 0, 0), scale: float = 1.0, mat: int = 0) -> SceneData:
    """Lat-long sphere, rings x segments x 2 triangles, radius 1 + amp * perlin(freq * p),
    smooth per-vertex normals. Pole triangles are degenerate (rejected by the
    Moller-Trumbore determinant test, as in the reference)."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, segments, endpoint=False)
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    unit = np.stack([st * np.cos(ph)[None, :], np.broadcast_to(ct, (rings + 1, segments)),
                     st * np.sin(ph)[None, :]], -1).reshape(-1, 3)
    r = 1.0 + amp * perlin3(unit * freq, seed)
    P = unit * r[:, None] * scale + np.asarray(center, np.float64)
    idx = np.arange((rings + 1) * segments).reshape(rings + 1, segments)
    a = idx[:-1, :]
    b = np.roll(idx[:-1, :], -1, axis=1)
    c = idx[1:, :]
    d = np.roll(idx[1:, :], -1, axis=1)
    # outward: (a, c, b) and (b, c, d) with theta increasing downwards
    F = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], 2).reshape(-1, 3)
    tri = P[F]
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    out = np.einsum("ij,ij->i", fn, tri.sum(1) - 3 * np.asarray(center, np.float64)) < 0
    if out.mean() > 0.5:
        F = F[:, [0, 2, 1]]
        fn = -fn
    vn = np.zeros_like(P)
    for k in range(3):
        np.add.at(vn, F[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    verts = P[F].reshape(-1, 9).astype(np.float32)
    normals = vn[F].reshape(-1, 9).astype(np.float32)
    return SceneData(verts, normals, np.full(F.shape[0], mat, np.uint32), [])


def incoherent_rays(verts: np.ndarray, n: int, seed: int = 3, offset: float = 1e-3) -> np.ndarray:
    """n rays (n, 8) float32: origins uniform on the surface (triangle chosen
    proportionally to area, barycentrics (1 - sqrt r1, r2 sqrt r1)) moved by
    offset * d off the surface (integrators/pathtracer.cpp:38), directions
    uniform on the sphere."""
    rng = np.random.default_rng(seed)
    v = verts.reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    cdf = np.cumsum(area)
    tri = np.searchsorted(cdf, rng.random(n) * cdf[-1], side="right").clip(0, len(area) - 1)
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    b0, b1, b2 = 1 - s, r2 * s, (1 - r2) * s
    p = b0[:, None] * v[tri, 0] + b1[:, None] * v[tri, 1] + b2[:, None] * v[tri, 2]
    z = 1 - 2 * rng.random(n)
    phi = 2 * np.pi * rng.random(n)
    rr = np.sqrt(np.maximum(0, 1 - z * z))
    d = np.stack([rr * np.cos(phi), rr * np.sin(phi), z], 1)
    o = p + offset * d
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = o
    out[:, 4:7] = d
    return out


# ----------------------------------------------------------------------------
# benchmark configurations
# ----------------------------------------------------------------------------

def _box_materials():
    white = api.material([api.lambertian((0.75, 0.75, 0.75))])
    red = api.material([api.lambertian((0.75, 0.0, 0.0))])
    green = api.material([api.lambertian((0.0, 0.75, 0.0))])
    light = api.material([api.lambertian((0.0, 0.0, 0.0))], emission=(15.0, 15.0, 15.0))
    return [white, red, green, light]


def cornell_icosphere(seed: int = 1, width: int = 256, height: int = 256, spp: int = 16, subdiv: int = 5,
                      max_depth: int = 5) -> SceneData:
    """Box of 10 Lambertian triangles (0.75 grey, red/green side walls), a
    2-triangle ceiling light (radiance 15), a seeded icosphere of 20*4**subdiv
    triangles (albedo 0.8); pinhole camera, jittered spp, fixed max depth."""
    rng = np.random.default_rng(seed)
    mats = _box_materials() + [api.material([api.lambertian((0.8, 0.8, 0.8))])]
    parts = box_walls((-1, -1, -1), (1, 1, 1), 0, 1, 2)
    parts.append(quad((-0.25, 0.999, -0.25), (0.25, 0.999, -0.25), (0.25, 0.999, 0.25), (-0.25, 0.999, 0.25),
                      (0, -1, 0), 3))
    center = (rng.uniform(-0.2, 0.2), -0.5 + rng.uniform(-0.05, 0.05), rng.uniform(-0.2, 0.2))
    radius = 0.45 + rng.uniform(-0.03, 0.03)
    parts.append(icosphere(subdiv, center, radius, random_rotation(rng), mat=4))
    sc = concat(parts, mats, name=f"cornell_icosphere_{20 * 4 ** subdiv}")
    sc.camera = api.pinhole(width, height, position=(0, 0, 3.6), look_at=(0, 0, 0), fov=40.0)
    sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="max-length",
                                   path_length=max_depth, seed=seed)
    return sc


def displaced_mesh(rings=512, segments=1024, seed=2, mat=0) -> SceneData:
    """1,048,576-triangle sphere with radial Perlin displacement (amplitude 0.1)."""
    return displaced_sphere(rings, segments, 0.1, 4.0, seed, mat=mat)


def layered_gold():
    """2-layer BSDF: smooth dielectric (IOR 1.5) over a rough GGX conductor (alpha 0.2, gold)."""
    return api.material([api.smooth_dielectric(1.5), api.ggx_conductor(0.2)])


def mesh_in_box(rings=512, segments=1024, width=800, height=600, spp=1024, seed=2) -> SceneData:
    """The displaced mesh with the layered gold BSDF inside the Lambertian box
    with an emissive ceiling quad; RR from depth 3, max depth 16."""
    mats = _box_materials() + [layered_gold()]
    parts = box_walls((-2, -1.5, -2), (2, 2.5, 2), 0, 1, 2)
    parts.append(quad((-0.6, 2.499, -0.6), (0.6, 2.499, -0.6), (0.6, 2.499, 0.6), (-0.6, 2.499, 0.6),
                      (0, -1, 0), 3))
    box = concat(parts, [])
    mesh = displaced_mesh(rings, segments, seed, mat=4)
    sc = concat([box, mesh], mats, name=f"mesh_in_box_{mesh.n_tris}")
    sc.camera = api.pinhole(width, height, position=(0, 0.4, 6.8), look_at=(0, 0.2, 0), fov=40.0)
    sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=3, max_depth=16, seed=seed)
    return sc


def slit_box(rings=512, segments=1024, width=800, height=600, spp=1024, seed=2, slit_frac=0.02) -> SceneData:
    """Dark closed box lit only through a ceiling slit 2% of the wall wide; the
    emitter sits above the slit outside the box. Glossy GGX floor (alpha 0.1)
    and the layered-BSDF mesh; max depth 16."""
    W = 4.0
    s = slit_frac * W
    white = api.material([api.lambertian((0.75, 0.75, 0.75))])
    red = api.material([api.lambertian((0.75, 0.0, 0.0))])
    green = api.material([api.lambertian((0.0, 0.75, 0.0))])
    light = api.material([api.lambertian((0.0, 0.0, 0.0))], emission=(60.0, 60.0, 60.0))
    floor = api.material([api.ggx_conductor(0.1, eta=(1.2, 1.2, 1.2), k=(2.5, 2.5, 2.5))])
    mats = [white, red, green, light, floor, layered_gold()]
    y0, y1 = -1.5, 2.5
    parts = [quad((-2, y0, -2), (2, y0, -2), (2, y0, 2), (-2, y0, 2), (0, 1, 0), 4),
             quad((-2, y0, -2), (2, y0, -2), (2, y1, -2), (-2, y1, -2), (0, 0, 1), 0),
             quad((-2, y0, 2), (2, y0, 2), (2, y1, 2), (-2, y1, 2), (0, 0, -1), 0),
             quad((-2, y0, -2), (-2, y1, -2), (-2, y1, 2), (-2, y0, 2), (1, 0, 0), 1),
             quad((2, y0, -2), (2, y1, -2), (2, y1, 2), (2, y0, 2), (-1, 0, 0), 2),
             # ceiling in two pieces leaving a slit along z at x in [-s/2, s/2]
             quad((-2, y1, -2), (-s / 2, y1, -2), (-s / 2, y1, 2), (-2, y1, 2), (0, -1, 0), 0),
             quad((s / 2, y1, -2), (2, y1, -2), (2, y1, 2), (s / 2, y1, 2), (0, -1, 0), 0),
             # light above the slit, facing down, outside the box
             quad((-0.5, y1 + 0.3, -2), (0.5, y1 + 0.3, -2), (0.5, y1 + 0.3, 2), (-0.5, y1 + 0.3, 2),
                  (0, -1, 0), 3)]
    box = concat(parts, [])
    mesh = displaced_mesh(rings, segments, seed, mat=5)
    sc = concat([box, mesh], mats, name=f"slit_box_{mesh.n_tris}")
    sc.camera = api.pinhole(width, height, position=(0, 0.3, 1.95), look_at=(0, -0.2, 0), fov=70.0)
    sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=3, max_depth=16, seed=seed)
    return sc


def furnace(subdiv: int = 3, albedo=(0.999, 0.999, 0.999), layer=None, width=64, height=64, spp=4) -> SceneData:
    """White furnace (samples/furnace_test.lua:1-150): an orthographic view of a
    unit sphere (triangulated) inside a kd=0 emitter sphere of radiance 0.5."""
    mats = [api.material([api.lambertian((0, 0, 0))], emission=(0.5, 0.5, 0.5)),
            api.material([layer if layer is not None else api.lambertian(albedo)])]
    outer = icosphere(2, (0, 0, 0), 10.0, mat=0)
    inner = icosphere(subdiv, (0, 0, 0), 1.0, mat=1, smooth=False)
    sc = concat([outer, inner], mats, name="furnace")
    sc.camera = api.orthographic(width, height, -1.25, 1.25, -1.25, 1.25, position=(0, 0, 6))
    sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=3, seed=7)
    return sc


def reference_furnace_materials():
    """The materials samples/furnace_test.lua:48-133 defines (importance sampling)."""
    perf = (0.999, 0.999, 0.999)
    return {
        "lambertian": api.material([api.lambertian(perf)]),
        "smooth_conductor": api.material([api.smooth_conductor(perf)]),
        "smooth_dielectric": api.material([api.smooth_dielectric(1.49)]),
        "cook_torrance": api.material([api.cook_torrance(0.2, perf)]),
        "plastic": api.material([api.smooth_dielectric(1.49), api.lambertian(perf)]),
        "coating": api.material([api.smooth_dielectric(1.49), api.cook_torrance(0.2, perf)]),
    }


def reference_furnace(material: str = "cook_torrance", width: int = 512, height: int = 512, spp: int = 256,
                      seed: int = 0) -> SceneData:
    """samples/furnace_test.lua as written: orthographic camera at (0, 0, 6)
    over [-1.25, 1.25]^2, jittered spp, Russian roulette from path length 3,
    background 0; an emitter sphere of radius 10 (kd 0, emission 0.5) around a
    unit sphere of the chosen material (the script's default: Cook-Torrance)."""
    light = api.material([api.lambertian((0.0, 0.0, 0.0))], emission=(0.5, 0.5, 0.5))
    mats = [light, reference_furnace_materials()[material]]
    sc = spheres([[0, 0, 0, 10], [0, 0, 0, 1]], [0, 1])
    sc.materials = mats
    sc.name = f"reference_furnace_{material}"
    sc.camera = api.orthographic(width, height, -1.25, 1.25, -1.25, 1.25, position=(0, 0, 6))
    sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="russian-roulette",
                                   path_length=3, seed=seed)
    return sc


def batch_scenes(n: int = 64, seed: int = 4, min_tris: int = 1000, max_tris: int = 500_000, min_side: int = 128,
                 max_side: int = 1024, spp_choices=(16, 64, 256), max_depth: int = 16) -> List[SceneData]:
    """Independent scenes for batched rendering: each a Lambertian box with a
    ceiling light around a seeded icosphere or Perlin-displaced sphere whose
    triangle count is log-uniform in [min_tris, max_tris]; square resolution
    uniform in [min_side, max_side]; spp from spp_choices; object material
    Lambertian, GGX (alpha 0.05-0.5, gold) or smooth dielectric (IOR 1.5);
    Russian roulette from depth 3, max depth max_depth."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = float(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        kind = rng.integers(3)
        if kind == 0:
            obj = api.material([api.lambertian(tuple(rng.uniform(0.2, 0.9, 3)))])
        elif kind == 1:
            obj = api.material([api.ggx_conductor(float(rng.uniform(0.05, 0.5)))])
        else:
            obj = api.material([api.smooth_dielectric(1.5)])
        mats = _box_materials() + [obj]
        parts = box_walls((-1, -1, -1), (1, 1, 1), 0, 1, 2)
        parts.append(quad((-0.3, 0.999, -0.3), (0.3, 0.999, -0.3), (0.3, 0.999, 0.3), (-0.3, 0.999, 0.3),
                          (0, -1, 0), 3))
        if rng.integers(2) == 0:
            sub = int(np.clip(np.round(np.log(max(target, 20) / 20.0) / np.log(4.0)), 1, 7))
            center = (rng.uniform(-0.2, 0.2), -0.45, rng.uniform(-0.2, 0.2))
            mesh = icosphere(sub, center, 0.45, random_rotation(rng), mat=4)
        else:
            rings = max(4, int(round(np.sqrt(target / 4.0))))
            mesh = displaced_sphere(rings, 2 * rings, 0.1, 4.0, int(rng.integers(1 << 30)), center=(0.0, -0.45, 0.0),
                                    scale=0.45, mat=4)
        sc = concat(parts + [mesh], mats, name=f"batch{i}_{mesh.n_tris}")
        side = int(rng.integers(min_side, max_side + 1))
        spp = int(spp_choices[rng.integers(len(spp_choices))])
        sc.camera = api.pinhole(side, side, position=(0, 0, 3.6), look_at=(0, 0, 0), fov=40.0)
        sc.integrator = api.Integrator(spp=spp, pixel_sampler="jittered", termination="russian-roulette",
                                       path_length=3, max_depth=max_depth, seed=int(rng.integers(1 << 31)))
        out.append(sc)
    return out
