"""Writes the golden fixtures of this directory from the oracle (the CPU
restatement of the reference). Each fixture holds its own inputs (triangles,
rays, camera and integrator settings) and the expected outputs, so a test can
check both the oracle (regression) and the HIP core (parity) against it.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as O  # noqa: E402
from swpathtracer_amd import scenes  # noqa: E402


def main():
    sc = scenes.cornell_icosphere(subdiv=2, width=16, height=16, spp=4)
    rays = scenes.incoherent_rays(sc.verts, 2000, seed=21)
    o = O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)
    t, prim, tests, isects = o.trace(rays)
    np.savez_compressed(os.path.join(HERE, "trace_cornell_332.npz"), verts=sc.verts, rays=rays, t=t, prim=prim,
                        tests=tests, isects=isects)
    p = sc.integrator.params()
    _, _, paths, _ = o.render(sc.camera, p, per_path=True)
    cam = np.array([sc.camera.width, sc.camera.height] + list(sc.camera.position) + list(sc.camera.look_at) +
                   list(sc.camera.up) + [sc.camera.fov_deg], np.float64)
    integ = np.array([p.spp, p.max_depth, p.rr_depth, p.seed], np.int64)
    np.savez_compressed(os.path.join(HERE, "paths_cornell_16x16x4.npz"), verts=sc.verts, mat_ids=sc.mat_ids,
                        camera=cam, integrator=integ, paths=paths)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
