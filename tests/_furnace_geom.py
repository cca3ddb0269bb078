"""Geometry of each path's first vertex in a furnace render, from the
oracle's camera, pixel sampler and closest hit: the cosine between the hit's
shading normal and the direction back to the camera, and whether the camera
ray met the object (material 1) or the emitter. Shared by the CPU pins and
the GPU tests of the furnace closed forms."""
import numpy as np

import pyoracle as O


def first_vertex(sc, orc):
    """Returns (cos_i, hits_object) per path, in per-path output order
    (pixel * spp + sample, path index = that too for a whole-image render)."""
    cam, integ = sc.camera, sc.integrator
    p = integ.params()
    spp = integ.effective_spp()
    rays = np.zeros((cam.width * cam.height * spp, 8), np.float32)
    i = 0
    for y in range(cam.height):
        for x in range(cam.width):
            for s in range(spp):
                path = (y * cam.width + x) * spp + s
                sx, sy, _ = O.pixel_sample(p.pixel_sampler, p.spp, s, p.seed, path)
                xs = np.float32(x) + np.float32(sx)  # pathtracer.cpp:110-111, float adds
                ys = np.float32(y) + np.float32(sy)
                rays[i] = O.camera_ray(cam, float(xs), float(ys))
                i += 1
    t, prim, _, _, nrm = orc.trace(rays, normals=True)
    d = rays[:, 4:7].astype(np.float64)
    cos_i = -np.sum(nrm.astype(np.float64) * d, axis=1) / np.linalg.norm(d, axis=1)
    obj = (prim != 0xFFFFFFFF) & (sc.mat_ids[np.minimum(prim, len(sc.mat_ids) - 1)] == 1)
    return cos_i, obj


def schlick_r0(r0, c):
    """FresnelSchlick(R0) (materials/fresnel_schlick.cpp:11-18) in float64."""
    r0 = np.asarray(r0, np.float64)
    return r0[None, :] + (1.0 - r0[None, :]) * ((1.0 - c) ** 5)[:, None]


def schlick_ior(ior_i, ior_t, c):
    """FresnelSchlick(ior_i, ior_t) for cos >= 0 (materials/fresnel_schlick.cpp:20-38)."""
    r0 = ((ior_i - ior_t) / (ior_i + ior_t)) ** 2
    return r0 + (1.0 - r0) * (1.0 - c) ** 5
