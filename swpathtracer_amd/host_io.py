"""Host-side scene input and image output around the core (no GPU work).

- `load_obj`: Wavefront OBJ meshes as the reference's Scene::loadMesh sees
  them after Assimp's triangulation (core/scene.cpp:32-68): polygons become
  fans, every face corner is its own vertex, faces are emitted mesh by mesh
  in file order, a mesh's material comes from the `submeshes` name map or the
  default id, and meshes not in the map are skipped unless
  `render_all_submeshes`. Vertex normals come from `vn`; a face without them
  is flat (all-zero normal row, primitives/triangle.cpp:5-6).
- `to_display` / `write_ppm`: ImageBuffer::convertTosRGB and saveToFile
  (core/image_buffer.cpp:22-47): clamp, x^(1/2.2), P3 text with
  int(v * 255 + 0.5).
- `normal_image`, `depth_image`, `count_heatmap`: the ray casters'
  saveImageToFile post-processing (integrators/normal_raycaster.cpp:6-17,
  integrators/depth_raycaster.cpp:9-34,
  integrators/int_test_count_raycaster.cpp:10-73).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from .scenes import SceneData


def _index(tok: str, n: int) -> int:
    """OBJ index (1-based, negative = relative to the end) -> 0-based."""
    i = int(tok)
    return i - 1 if i > 0 else n + i


def read_obj(path: str) -> List[Tuple[str, np.ndarray, np.ndarray]]:
    """Meshes of an OBJ file in file order: (name, triangles (m, 9) float32,
    corner normals (m, 9) float32, zero rows where a face has no normals).
    A new mesh starts at each `o` or `g` statement that is followed by faces;
    faces before any name belong to the mesh "default"."""
    pos: List[List[float]] = []
    nrm: List[List[float]] = []
    meshes: List[Tuple[str, List[np.ndarray], List[np.ndarray]]] = []
    name = "default"
    cur: Optional[Tuple[str, List[np.ndarray], List[np.ndarray]]] = None
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "v":
                pos.append([float(x) for x in t[1:4]])
            elif t[0] == "vn":
                nrm.append([float(x) for x in t[1:4]])
            elif t[0] in ("o", "g"):
                name = " ".join(t[1:]) if len(t) > 1 else "default"
                cur = None
            elif t[0] == "f":
                if cur is None:
                    cur = (name, [], [])
                    meshes.append(cur)
                corners = []
                for c in t[1:]:
                    parts = c.split("/")
                    vi = _index(parts[0], len(pos))
                    ni = _index(parts[2], len(nrm)) if len(parts) > 2 and parts[2] else None
                    corners.append((vi, ni))
                for k in range(1, len(corners) - 1):  # fan triangulation
                    tri = [corners[0], corners[k], corners[k + 1]]
                    v = np.array([pos[i] for i, _ in tri], np.float64)
                    # degenerate faces (repeated corners) are dropped, as
                    # Assimp's degenerate finder turns them into points/lines
                    if len({i for i, _ in tri}) < 3:
                        continue
                    if all(n is not None for _, n in tri):
                        nn = np.array([nrm[n] for _, n in tri], np.float64)
                    else:
                        nn = np.zeros((3, 3))
                    cur[1].append(v.reshape(9))
                    cur[2].append(nn.reshape(9))
    out = []
    for nm, vs, ns in meshes:
        if vs:
            out.append((nm, np.array(vs, np.float32), np.array(ns, np.float32)))
    return out


def load_obj(path: str, default_material_id: int = 0, render_all_submeshes: bool = True,
             submeshes: Optional[Dict[str, int]] = None) -> SceneData:
    """Triangles, normals and material ids of an OBJ file in the reference's
    primitive order (core/scene.cpp:32-68)."""
    submeshes = submeshes or {}
    vs, ns, ms = [], [], []
    for name, v, n in read_obj(path):
        if name not in submeshes and not render_all_submeshes:
            continue
        mat = submeshes.get(name, default_material_id)
        vs.append(v)
        ns.append(n)
        ms.append(np.full(len(v), mat, np.uint32))
    if not vs:
        return SceneData(np.zeros((0, 9), np.float32), None, np.zeros(0, np.uint32), [])
    normals = np.concatenate(ns)
    return SceneData(np.concatenate(vs), normals if np.any(normals) else None, np.concatenate(ms), [])


def to_display(rgb: np.ndarray) -> np.ndarray:
    """clamp to [0, 1] then x^(1/2.2) per channel, in float32."""
    c = np.clip(np.asarray(rgb, np.float32), 0.0, 1.0)
    return np.power(c.astype(np.float64), 1.0 / np.float32(2.2)).astype(np.float32)


def write_ppm(path: str, rgb: np.ndarray):
    """PPM P3 of an (h, w, 3) image of values in [0, 1] (clamped), rows top
    to bottom, each value as int(v * 255 + 0.5) followed by a space."""
    img = np.clip(np.asarray(rgb, np.float32), 0.0, 1.0)
    h, w = img.shape[:2]
    q = (img * np.float32(255.0) + np.float32(0.5)).astype(np.int32)
    with open(path, "w") as fh:
        fh.write(f"P3\n{w} {h}\n255\n")
        fh.write("".join(f"{v} " for v in q.reshape(-1)))


def normal_image(normals: np.ndarray) -> np.ndarray:
    """Normal ray caster output: n * 0.5 + 0.5 where any component is
    non-zero, black (the miss value) elsewhere."""
    n = np.asarray(normals, np.float32)
    hit = np.any(n != 0.0, axis=-1, keepdims=True)
    return np.where(hit, n * np.float32(0.5) + np.float32(0.5), np.float32(0.0)).astype(np.float32)


def depth_image(depth: np.ndarray, prescribed_min: float = -1.0, prescribed_max: float = -1.0) -> np.ndarray:
    """Depth ray caster output: (t - min) / (max - min) as grey over hits
    (-1 marks a miss: black); min over positive depths and max over all
    values unless prescribed (-1 = not prescribed); a zero range paints hits
    white."""
    d = np.asarray(depth, np.float32)
    mx = np.float32(d.max()) if d.size else np.float32(-np.inf)
    pos = d[d > 0.0]
    mn = np.float32(pos.min()) if pos.size else np.float32(np.inf)
    if prescribed_min != -1.0:
        mn = np.float32(prescribed_min)
    if prescribed_max != -1.0:
        mx = np.float32(prescribed_max)
    rng = np.float32(mx - mn)
    if rng:
        v = (d - mn) / rng
    else:
        v = np.ones_like(d)
    v = np.where(d != -1.0, v, np.float32(0.0)).astype(np.float32)
    return np.repeat(v[..., None], 3, axis=-1)


def _hue2rgb(v1, v2, h):
    h = np.where(h < 0.0, h + np.float32(1.0), h)
    h = np.where(h > 1.0, h - np.float32(1.0), h)
    return np.where(6.0 * h < 1.0, v1 + (v2 - v1) * np.float32(6.0) * h,
                    np.where(2.0 * h < 1.0, v2,
                             np.where(3.0 * h < 2.0, v1 + (v2 - v1) * (np.float32(2.0 / 3.0) - h) * np.float32(6.0),
                                      v1))).astype(np.float32)


def count_heatmap(counts: np.ndarray, prescribed_min: int = -1, prescribed_max: int = -1) -> np.ndarray:
    """Intersection-test-count ray caster output: counts normalised to [0, 1]
    between the per-pixel minimum and maximum (or the prescribed ones), then a
    blue-to-red heat map through HSL (hue 0.65 (1 - t), saturation 1,
    lightness t / 2)."""
    c = np.asarray(counts, np.uint64)
    mn = np.uint64(prescribed_min) if prescribed_min != -1 else (c.min() if c.size else np.uint64(0))
    mx = np.uint64(prescribed_max) if prescribed_max != -1 else (c.max() if c.size else np.uint64(0))
    rng = np.uint32((int(mx) - int(mn)) & 0xFFFFFFFF)
    if rng > 0:
        t = ((c - mn).astype(np.float32) / np.float32(rng)).astype(np.float32)
    elif mx > 0:
        t = np.ones(c.shape, np.float32)
    else:
        t = np.zeros(c.shape, np.float32)
    h = np.float32(0.65) - t * np.float32(0.65)
    s = np.float32(1.0)
    l = t * np.float32(0.5)
    v2 = np.where(l < 0.5, l * (np.float32(1.0) + s), (l + s) - (s * l)).astype(np.float32)
    v1 = (np.float32(2.0) * l - v2).astype(np.float32)
    third = np.float32(1.0 / 3.0)
    return np.stack([_hue2rgb(v1, v2, h + third), _hue2rgb(v1, v2, h), _hue2rgb(v1, v2, h - third)], -1)
