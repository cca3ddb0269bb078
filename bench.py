"""Benchmark: plain GPU path tracing on one MI355X.

Workload (the value): the 1,048,576-triangle Perlin-displaced sphere with a
2-layer BSDF (smooth dielectric, IOR 1.5, over GGX alpha 0.2 gold) inside the
Lambertian box with an emissive ceiling quad, 800x600, jittered, Russian
roulette from depth 3, max depth 16. One step = one render call of the full
1024-spp image (491.52M camera paths); the default is 1 timed step after 1
warm-up step. Inputs (tree, triangles, materials) are resident in HBM before
timing; the value is camera paths per second over the timed steps.

Also reported in the same JSON line:
- roofline of the dominant kernel (k_extend, the closest-hit traversal of
  extension rays) from its HIP-event launch times and its algorithmic bytes;
- converged-image RMSE of the 1024-spp image against an FP64-film oracle
  image of 65,536 spp (independent seed), with the oracle image's own noise
  floor;
- closest hit on the same 1M-triangle mesh: 1920x1080 coherent primary rays
  and 16,777,216 incoherent rays (rays/s, algorithmic fetch GB/s);
- the 256x256, 16-spp, depth-5 box + 20,480-triangle icosphere workload;
- the guided vs plain time to equal RMSE on the slit-lit scene;
- the 64-scene batch;
- cpu_baseline: the oracle (the C++ restatement of the reference's OpenMP CPU
  path) on the host cores: a viewport slice of the 1024-spp workload (the
  value's unit), the full 256x256 box image, the coherent frame and 1M of the
  incoherent rays, with omp_get_max_threads(), nproc and the CPU model.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
REFERENCE_MPATHS = 0.39  # reference README.md:11-19 (Buddha, 800x600x4096 spp in ~1h24m, i5, OpenMP)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(sc3, c1, mesh, coherent_rays, incoherent_rays):
    """The oracle (reference semantics: recursive traversal, heap nodes,
    virtual primitives, OpenMP dynamic rows / dynamic ray chunks) on the host
    cores, each workload sized to a few seconds:
    - the 1024-spp workload: an 80x60 centre viewport (the reference's own
      viewport mechanism), samples sized to ~8 s;
    - the box + icosphere workload in full (256x256, 16 spp, depth 5);
    - closest hit on the 1M-triangle mesh: the 1920x1080 coherent frame and a
      seeded 1,048,576-ray subset of the incoherent rays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    import ctypes as C

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = O.omp_max_threads()
    out = {"cores": threads, "nproc": os.cpu_count(), "omp_max_threads": omp, "cpu_model": _cpu_model(),
           "kind": "port"}

    # 1024-spp workload, viewport slice
    t0 = time.perf_counter()
    orc = O.Scene(sc3.verts, sc3.normals, sc3.mat_ids, materials=sc3.materials)
    build_s = time.perf_counter() - t0
    cam = sc3.camera
    vw, vh = 80, 60
    vp = (cam.height // 2 - vh // 2, cam.width // 2 - vw // 2, vw, vh)
    p = sc3.integrator.params()
    p.viewport = (C.c_uint32 * 4)(*vp)
    p.spp_begin, p.spp_count = 0, 1
    t0 = time.perf_counter()
    orc.render(cam, p, threads=threads)
    probe = time.perf_counter() - t0
    spp = int(max(1, min(1024, 8.0 / max(probe, 1e-3))))
    p.spp_count = spp
    t0 = time.perf_counter()
    orc.render(cam, p, threads=threads)
    dt = time.perf_counter() - t0
    paths = vw * vh * spp
    orc.close()
    out.update({"value": paths / dt / 1e6, "unit": "Mpaths/s",
                "sample": f"{vw}x{vh} centre viewport of the 800x600 1024-spp workload x {spp} spp = {paths} paths "
                          f"in {dt:.2f} s (tree build {build_s:.1f} s not timed)"})

    # box + icosphere, full image
    orc = O.Scene(c1.verts, c1.normals, c1.mat_ids, materials=c1.materials)
    t0 = time.perf_counter()
    orc.render(c1.camera, c1.integrator.params(), threads=threads)
    dt = time.perf_counter() - t0
    n1 = c1.camera.width * c1.camera.height * c1.integrator.effective_spp()
    orc.close()
    out["box_icosphere_full_image"] = {"mpaths_s": round(n1 / dt / 1e6, 4), "paths": n1, "s": round(dt, 3)}

    # closest hit, 1M-triangle mesh
    t0 = time.perf_counter()
    orc = O.Scene(mesh.verts, mesh.normals, mesh.mat_ids)
    build_s = time.perf_counter() - t0
    for name, rays in (("coherent_1920x1080", coherent_rays), ("incoherent_1M_subset", incoherent_rays)):
        if len(rays) == 0:
            continue
        t0 = time.perf_counter()
        orc.trace(rays, threads=threads)
        dt = time.perf_counter() - t0
        out[name] = {"mrays_s": round(len(rays) / dt / 1e6, 4), "rays": int(len(rays)), "s": round(dt, 3)}
    out["mesh_tree_build_s_not_timed"] = round(build_s, 2)
    orc.close()
    return out


def closest_hit_bench(dev, rays, reps=5):
    """Closest hit of a resident ray batch on the device's current scene:
    rays/s over `reps` launches (HIP events on the context stream) and the
    algorithmic fetch rate (rays and hits plus the tree data and triangle
    records each ray fetched, counted by one FETCH launch on the same rays)."""
    import swpathtracer_amd._lib as L

    n_rays = len(rays)
    d_rays = dev.to_device(rays)
    d_hits = dev.alloc(8 * n_rays)
    d_nodes = dev.alloc(4 * n_rays)
    d_tris = dev.alloc(4 * n_rays)
    dev.trace(d_rays.ptr, n_rays, d_hits.ptr, L.SPT_TRACE_CLOSEST)
    dev.sync()
    dev.event_record(0)
    for _ in range(reps):
        dev.trace(d_rays.ptr, n_rays, d_hits.ptr, L.SPT_TRACE_CLOSEST)
    dev.event_record(1)
    ms = dev.event_elapsed(0, 1) / reps
    dev.sync()
    dev.trace(d_rays.ptr, n_rays, d_hits.ptr, L.SPT_TRACE_FETCH, d_nodes.ptr, d_tris.ptr)
    dev.sync()
    nodes = d_nodes.download(np.uint32, n_rays).astype(np.float64)
    tris = d_tris.download(np.uint32, n_rays).astype(np.float64)
    hits = d_hits.download(np.uint32, (n_rays, 2))
    byts = 32 * n_rays + 8 * n_rays + 32 * nodes.sum() + 48 * tris.sum()  # tree data in 32-B units
    for b in (d_rays, d_hits, d_nodes, d_tris):
        b.free()
    return {"rays": n_rays, "ms": ms, "grays_per_s": n_rays / (ms * 1e-3) / 1e9,
            "mean_tree_units_32B": nodes.mean(), "mean_tris": tris.mean(), "hit_frac": float((hits[:, 1] != L.SPT_MISS).mean()),
            "fetch_gbs": byts / (ms * 1e-3) / 1e9}


def c3_converged_rmse(dev, sc, img1024, oracle_spp):
    """Per-pixel RMSE of the 1024-spp image against an FP64-film image of
    oracle_spp samples per pixel with an independent seed, rendered in
    256-spp calls. The oracle image carries its own noise: for independent
    estimates, E[RMSE^2] = noise(1024)^2 + noise(oracle_spp)^2, so the floor
    the reference image adds is rmse * sqrt(1024 / oracle_spp) to first order."""
    import swpathtracer_amd._lib as L

    n = int(round(np.sqrt(oracle_spp)))
    assert n * n == oracle_spp, "the FP64 oracle image needs a square jittered spp"
    cam, integ = sc.camera, sc.integrator
    npix = cam.width * cam.height
    film64 = dev.alloc(npix * 3 * 8)
    film64.zero()
    t0 = time.perf_counter()
    for b in range(0, oracle_spp, 256):
        p = integ.params()
        p.spp, p.seed = oracle_spp, 0x0FAC1E5
        p.spp_begin, p.spp_count = b, min(256, oracle_spp - b)
        p.flags = L.SPT_RENDER_FILM64
        dev.render(cam, p, None, film64, None)
    dev.sync()
    ref = film64.download(np.float64, (cam.height, cam.width, 3)) / oracle_spp
    film64.free()
    dt = time.perf_counter() - t0
    rmse = _rmse(img1024, ref)
    floor = rmse * np.sqrt(1024.0 / oracle_spp)
    log(f"converged image: RMSE {rmse:.6f} vs {oracle_spp}-spp FP64 image ({dt:.1f} s), floor {floor:.6f}")
    return {"oracle_spp": oracle_spp, "oracle_film": "FP64", "oracle_seed": "independent", "rmse_1024spp": round(rmse, 6),
            "oracle_noise_floor": round(float(floor), 6), "oracle_mean": [round(float(x), 5) for x in ref.reshape(-1, 3).mean(0)],
            "image_mean": [round(float(x), 5) for x in img1024.reshape(-1, 3).mean(0)],
            "oracle_render_s": round(dt, 1)}


def _rmse(img, ref):
    return float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))


def guide_experiment(args):
    """Time to equal RMSE on the slit-lit scene: the plain tracer renders
    1024 spp in one call (timed); the guided tracer renders warm-up passes
    that only train the guide, then counted passes (training continues after each)
    until its running image reaches the plain image's RMSE against an
    FP64-film reference of oracle_spp samples (independent seed). All render
    and training time is counted, warm-up included; RMSE evaluation between
    passes is not."""
    from swpathtracer_amd import api, scenes
    import swpathtracer_amd._lib as L

    W, H = args.guide_res
    t0 = time.perf_counter()
    sc = scenes.slit_box(rings=args.rings, segments=args.segments, width=W, height=H, spp=1024)
    dev = api.Device(0)
    bvh = api.Bvh(sc.verts, n_threads=os.cpu_count() or 8)
    dev.set_scene(bvh, sc.normals, sc.mat_ids)
    dev.set_materials(sc.materials)
    log(f"slit scene {sc.n_tris} triangles, {W}x{H} ({time.perf_counter() - t0:.1f} s)")
    cam, integ = sc.camera, sc.integrator
    npix = W * H

    # reference image: FP64 film, independent seed, a square jittered spp so
    # that every stratum of every pixel is sampled equally
    osp = args.oracle_spp
    assert int(round(np.sqrt(osp))) ** 2 == osp, "--oracle-spp must be a square"
    film64 = dev.alloc(npix * 3 * 8)
    film64.zero()
    t0 = time.perf_counter()
    chunk = 256
    for b in range(0, osp, chunk):
        p = integ.params()
        p.spp, p.seed = osp, 0x5EED0F0A
        p.spp_begin, p.spp_count = b, min(chunk, osp - b)
        p.flags = L.SPT_RENDER_FILM64
        dev.render(cam, p, None, film64, None)
    dev.sync()
    ref = film64.download(np.float64, (H, W, 3)) / osp
    log(f"reference {osp} spp: {time.perf_counter() - t0:.1f} s, mean {ref.reshape(-1, 3).mean(0)}")

    film = dev.alloc(npix * 3 * 4)

    def plain(spp):
        film.zero()
        p = integ.params()
        p.spp = spp
        p.flags = L.SPT_RENDER_FILM32
        dev.sync()
        t = time.perf_counter()
        dev.render(cam, p, film, None, None)
        dev.sync()
        return time.perf_counter() - t, film.download(np.float32, (H, W, 3)) / spp

    plain(16)  # warm-up
    t_plain, img_plain = plain(1024)
    r_plain = _rmse(img_plain, ref)
    log(f"plain 1024 spp: {t_plain * 1e3:.1f} ms, RMSE {r_plain:.5f}")

    # guided: warm-up passes (guide trained after every pass, images
    # discarded but their time counted), then counted passes that keep
    # training; the running image of the counted passes is compared with the
    # reference after each pass (comparison time not counted)
    dev.guide_init(seed=1, lr=args.guide_lr, batch=262144, steps_per_pass=args.guide_steps, flags=args.guide_flags,
                   train_every=args.guide_train_every, bsdf_fraction=args.guide_bsdf_fraction)
    if args.guide_capacity:  # queue of the guided passes (the plain render above used the default)
        dev.set_tuning(path_capacity=args.guide_capacity)
    t_guided, t_train, hit = 0.0, 0.0, None
    curve = []
    k = 0

    def guided_pass(c, tgt):
        nonlocal k, t_guided, t_train
        k += 1
        p = integ.params()
        p.spp, p.seed = c, 0x600D0000 + k  # a complete jittered estimate of c spp
        p.flags = L.SPT_RENDER_FILM32 | L.SPT_RENDER_GUIDED | L.SPT_RENDER_GUIDE_TRAIN
        dev.sync()
        t = time.perf_counter()
        dev.render(cam, p, tgt, None, None)
        t1 = time.perf_counter()
        dev.guide_train_collected(wait=False)
        dev.sync()
        t2 = time.perf_counter()
        t_guided += t2 - t
        t_train += t2 - t1

    warm = dev.alloc(npix * 3 * 4)
    warm.zero()
    for _ in range(args.guide_warmup):
        guided_pass(args.guide_warmup_spp, warm)
    t_warm = t_guided
    # counted passes run back to back; after each, the running film is copied
    # on the device and its RMSE is evaluated only after the run, so the
    # measurement leaves the GPU no idle gaps between passes
    film.zero()
    c = args.guide_pass_spp
    snaps, marks = [], []
    done = 0
    while done < 4096 and t_guided < 2.0 * t_plain:
        guided_pass(c, film)
        done += c
        snap = dev.alloc(npix * 3 * 4)
        dev._check(dev.lib.spt_memcpy_d2d(dev.h, snap.ptr, film.ptr, npix * 12))
        snaps.append(snap)
        marks.append((done, t_guided))
    dev.sync()
    host = np.empty((H, W, 3), np.float32)
    for snap, (spp_k, t_k) in zip(snaps, marks):
        snap.download(np.float32, (H, W, 3), out=host)
        r = _rmse(host / spp_k, ref)
        curve.append((spp_k, round(t_k * 1e3, 2), round(r, 6)))
        if hit is None and r <= r_plain:
            hit = (spp_k, t_k, r)
        snap.free()
    st = dev.guide_stats()
    flops = st["trained_records"] * 76.8e3
    mfma = flops / max(st["train_kernel_ms"] * 1e-3, 1e-12) / 2.5e15
    # inference: one more 144-spp guided pass (trained guide, no capture) with
    # the inference launches timed by HIP events; 25,600 useful FLOP per query
    # (32x64 + 64x64 + 64x64 + 64x40 multiply-adds)
    # (one wavefront, so that no other wavefront's kernels share the timed
    # launches)
    dev.set_tuning(wavefronts=1, path_capacity=args.guide_capacity)
    p = integ.params()
    p.spp, p.seed = 144, 0x1FE12
    p.flags = L.SPT_RENDER_FILM32 | L.SPT_RENDER_GUIDED | L.SPT_RENDER_TIME_KERNELS
    film.zero()
    dev.render(cam, p, film, None, None)
    dev.sync()
    rs = dev.render_stats()
    dev.set_tuning(path_capacity=args.guide_capacity)
    inf_tflops = rs["guide_queries"] * 25600 / max(rs["guide_ms"] * 1e-3, 1e-12) / 1e12
    # its memory side: per query 104 B read (queue index 4, hit 16, sort key 4,
    # origin 16, direction 16, the triangle's shading record 48) and 80 B of
    # half-float outputs written
    inf_gbs = rs["guide_queries"] * 184 / max(rs["guide_ms"] * 1e-3, 1e-12) / 1e9
    infer = {"queries": int(rs["guide_queries"]), "kernel_ms": round(rs["guide_ms"], 3),
             "tflops": round(inf_tflops, 1), "mfma_util": round(inf_tflops / 2500.0, 4),
             "algorithmic_gbs": round(inf_gbs, 1), "hbm_frac": round(inf_gbs / HBM_PEAK_GBS, 4)}
    res = {"scene": f"slit box, {sc.n_tris} triangles, {W}x{H}", "reference_spp": osp,
           "plain_ms": round(t_plain * 1e3, 2), "plain_rmse": round(r_plain, 6),
           "plain_mpaths_s": round(npix * 1024 / t_plain / 1e6, 3),
           "guided_spp_to_equal_rmse": hit[0] if hit else None,
           "guided_ms_to_equal_rmse": round(hit[1] * 1e3, 2) if hit else None,
           "time_ratio": round(hit[1] / t_plain, 4) if hit else None,
           "guided_train_ms": round(t_train * 1e3, 2), "train_steps": st["steps"],
           "warmup_passes": args.guide_warmup, "warmup_ms": round(t_warm * 1e3, 2),
           "train_kernel_ms": round(st["train_kernel_ms"], 3), "trained_records": st["trained_records"],
           "train_mfma_util": round(mfma, 6), "final_loss": st["last_loss"], "inference": infer,
           "curve": curve[:: max(1, len(curve) // 24)]}
    log(f"guide experiment: {res}")
    dev.close()
    return res


def batch_bench(args):
    """64 independent scenes (box + seeded icosphere or displaced sphere,
    1k-500k triangles log-uniform, 128^2-1024^2, 16/64/256 spp, Lambertian /
    GGX / dielectric objects) rendered in one call after autotuning tile and
    queue sizes, launch shapes and the number of concurrent wavefronts on a
    1/64 sample of every scene's samples."""
    from swpathtracer_amd import api, scenes

    t0 = time.perf_counter()
    scs = scenes.batch_scenes(n=64, seed=4)
    log(f"batch: {len(scs)} scenes, {sum(s.n_tris for s in scs)} triangles ({time.perf_counter() - t0:.1f} s)")
    dev = api.Device(0)
    t0 = time.perf_counter()
    b = api.Batch(dev, scs, n_threads=os.cpu_count() or 8)
    log(f"batch trees built and uploaded ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    tuning = b.autotune(sample_div=64)
    log(f"autotuned in {time.perf_counter() - t0:.1f} s: {tuning}")
    b.clear()
    dev.sync()
    t0 = time.perf_counter()
    b.render()
    dev.sync()
    wall = time.perf_counter() - t0
    st, ms = b.stats()
    paths = st["paths"]
    res = {"scenes": len(scs), "triangles": int(st["triangles"]), "paths": int(paths),
           "wall_ms": round(wall * 1e3, 2), "mpaths_s": round(paths / wall / 1e6, 3),
           "grays_s": round(st["rays"] / wall / 1e9, 4),
           "scene_ms_min": round(float(ms.min()), 2), "scene_ms_median": round(float(np.median(ms)), 2),
           "scene_ms_max": round(float(ms.max()), 2), "tuning": tuning,
           "device_mb": round(st["device_bytes"] / 2 ** 20, 1)}
    # the same render with the bounce loop captured once in a HIP graph and
    # replayed (the tuner's other settings kept), whatever the tuner chose
    img0 = b.image(0)
    dev.set_tuning(**{**dev.get_tuning(), "use_graph": 1})
    b.clear()
    dev.sync()
    t0 = time.perf_counter()
    b.render()
    dev.sync()
    wall_g = time.perf_counter() - t0
    res["graph_replay"] = {"wall_ms": round(wall_g * 1e3, 2), "mpaths_s": round(paths / wall_g / 1e6, 3),
                           "scene0_image_max_rel_diff": float(np.max(np.abs(b.image(0) - img0)) /
                                                              max(float(np.max(np.abs(img0))), 1e-30))}
    log(f"batch: {res}")
    b.close()
    dev.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=1024,
                    help="samples per pixel of one step (one render call); 1 step = the 1024-spp image")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-closest-hit", action="store_true")
    ap.add_argument("--no-sort", action="store_true", help="skip the material sort between stages")
    ap.add_argument("--rings", type=int, default=512)
    ap.add_argument("--segments", type=int, default=1024)
    ap.add_argument("--no-guide", action="store_true", help="skip the guided vs plain equal-RMSE experiment")
    ap.add_argument("--guide-only", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched render")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--guide-full", action="store_true", help="keep rendering after equal RMSE (full curve)")
    ap.add_argument("--guide-res", type=int, nargs=2, default=(800, 600))
    ap.add_argument("--oracle-spp", type=int, default=65536,
                    help="spp of the slit scene's FP64-film oracle image (a square)")
    ap.add_argument("--c3-oracle-spp", type=int, default=65536,
                    help="spp of the 1024-spp workload's FP64-film oracle image (a square; 0: skip)")
    ap.add_argument("--guide-lr", type=float, default=1e-2)
    ap.add_argument("--guide-warmup", type=int, default=4, help="training passes before the counted image")
    ap.add_argument("--guide-warmup-spp", type=int, default=4, help="spp of each warm-up training pass")
    ap.add_argument("--guide-pass-spp", type=int, default=144, help="spp of each counted guided pass (a square)")
    ap.add_argument("--guide-steps", type=int, default=24, help="Adam steps per training pass")
    ap.add_argument("--guide-capacity", type=int, default=0, help="path queue of the guided passes (0: default)")
    ap.add_argument("--guide-bsdf-fraction", type=float, default=0.5, help="one-sample mixture weight of BSDF sampling")
    ap.add_argument("--guide-train-every", type=int, default=0,
                    help="> 0: train inside each guided call every this many bounces")
    ap.add_argument("--guide-flags", type=int, default=0,
                    help="SPT_GUIDE_* options: 1 guide layers under a smooth coat, 2 radiance x BSDF-lobe product")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for {args.gpus}); "
              "the path shards by pixel tiles / sample ranges across GPUs but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    from swpathtracer_amd import api, scenes
    import swpathtracer_amd._lib as L
    import ctypes as C

    if args.copy_probe:
        dev = api.Device(0)
        nbytes = 256 << 20
        a, b = dev.alloc(nbytes), dev.alloc(nbytes)
        a.zero()
        for _ in range(3):
            dev._check(dev.lib.spt_memcpy_d2d(dev.h, b.ptr, a.ptr, nbytes))
        dev.sync()
        print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))
        return

    if args.guide_only:
        print(json.dumps({"guide_experiment": guide_experiment(args)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_bench(args)}), flush=True)
        return

    t0 = time.perf_counter()
    sc = scenes.mesh_in_box(rings=args.rings, segments=args.segments)
    log(f"scene {sc.name}: {sc.n_tris} triangles ({time.perf_counter() - t0:.1f} s)")
    dev = api.Device(0)
    t0 = time.perf_counter()
    bvh = api.Bvh(sc.verts, n_threads=os.cpu_count() or 8)
    info = bvh.info
    log(f"tree: {info['n_pair_nodes']} pair nodes, depth {info['max_depth']}, build {time.perf_counter() - t0:.1f} s")
    dev.set_scene(bvh, sc.normals, sc.mat_ids)
    dev.set_materials(sc.materials)
    cam = sc.camera
    integ = sc.integrator
    spp_total = integ.effective_spp()
    sps = args.spp_per_step
    npix = cam.width * cam.height
    film = dev.alloc(npix * 3 * 4)
    film.zero()

    def step(i, flags=L.SPT_RENDER_FILM32):
        p = integ.params()
        p.spp_begin = (i * sps) % spp_total
        p.spp_count = sps
        p.flags = flags | (L.SPT_RENDER_NO_SORT if args.no_sort else 0)
        dev.render(cam, p, film, None, None)

    for i in range(args.warmup):
        step(i)
    dev.sync()
    log(f"warmup {args.warmup} steps done")
    film.zero()
    # timed region: K steps, device synchronised on both sides
    dev.sync()
    t0 = time.perf_counter()
    ext_ms = 0.0
    ext_n = 0
    rays = 0
    for i in range(args.steps):
        step(i, L.SPT_RENDER_FILM32 | L.SPT_RENDER_TIME_KERNELS)
        st = dev.render_stats()
        ext_ms += st["extend_ms"]
        ext_n += st["extend_launches"]
        rays += st["rays"]
    dev.sync()
    wall = time.perf_counter() - t0
    paths = args.steps * sps * npix
    log(f"timed {args.steps} steps: {wall:.3f} s, {paths / wall / 1e6:.1f} Mpaths/s, "
        f"{rays / wall / 1e9:.3f} Grays/s extension rays, extend kernel {ext_ms / 1e3:.3f} s")
    img = film.download(np.float32, (cam.height, cam.width, 3)) / (args.steps * sps)
    # algorithmic bytes of the extension kernel: one instrumented step (into
    # the same film, after the image of the timed steps was read)
    step(0, L.SPT_RENDER_FILM32 | L.SPT_RENDER_FETCH_STATS)
    dev.sync()
    fs = dev.render_stats()
    per_ray_nodes = fs["node_fetches"] / fs["rays"]
    per_ray_tris = fs["tri_fetches"] / fs["rays"]
    # per extension ray: queue index 4 + origin 16 + direction 16 in, hit 16 + sort key 4 out
    bytes_per_ray = 56 + 32 * per_ray_nodes + 48 * per_ray_tris  # tree data counted in 32-B units
    achieved = bytes_per_ray * rays / (ext_ms * 1e-3) / 1e9
    # The renderer runs concurrent wavefronts, so an extension launch shares
    # the GPU with another wavefront's shading and lasts longer than it would
    # alone. Beside the per-launch figure: the extension rays' algorithmic
    # bytes over the whole timed wall time, and one untimed step on a single
    # wavefront (the kernel with the GPU to itself).
    aggregate = bytes_per_ray * rays / wall / 1e9
    tn_default = dev.get_tuning()
    dev.set_tuning(wavefronts=1)  # the other fields at their one-wavefront defaults
    step(0, L.SPT_RENDER_FILM32 | L.SPT_RENDER_TIME_KERNELS)
    dev.sync()
    si = dev.render_stats()
    dev.set_tuning()  # back to the library defaults
    single = {"avg_launch_ms": round(si["extend_ms"] / max(si["extend_launches"], 1), 4),
              "launches": int(si["extend_launches"]),
              "achieved": round(bytes_per_ray * si["rays"] / (si["extend_ms"] * 1e-3) / 1e9, 1)}
    single["frac"] = round(single["achieved"] / HBM_PEAK_GBS, 4)

    # HBM bytes per extension launch of this workload, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-calibrated)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_pt_mesh_in_box.json")
    if os.path.exists(tf) and args.rings == 512 and args.segments == 1024:
        with open(tf) as fh:
            tj = json.load(fh)
        # measured on this workload's launches, with as many wavefronts
        if tj.get("spp_per_step") == sps and tj.get("wavefronts", 1) == tn_default["wavefronts"]:
            traffic = tj["hbm_bytes_per_launch"]

    extra = {}
    if args.c3_oracle_spp and args.steps * sps == spp_total:  # the timed image is the full 1024-spp estimate
        extra["converged_image"] = c3_converged_rmse(dev, sc, img, args.c3_oracle_spp)

    # the box + 20,480-triangle icosphere workload: 256x256, 16 spp, depth 5
    c1 = scenes.cornell_icosphere()
    dev.set_scene(api.Bvh(c1.verts, n_threads=os.cpu_count() or 8), c1.normals, c1.mat_ids)
    dev.set_materials(c1.materials)
    c1_film = dev.alloc(c1.camera.width * c1.camera.height * 12)
    p1 = c1.integrator.params()
    p1.flags = L.SPT_RENDER_FILM32
    c1_paths = c1.camera.width * c1.camera.height * c1.integrator.effective_spp()

    def c1_time(reps=10):
        dev.render(c1.camera, p1, c1_film, None, None)
        dev.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            dev.render(c1.camera, p1, c1_film, None, None)
            dev.sync()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    c1_default = c1_time()
    # a 1M-path image is latency-bound: the autotuner picks a smaller
    # traversal grid and queue for it (the 1024-spp workload keeps its own)
    c1_tuning = dev.autotune(c1.camera, c1.integrator, sample_paths=c1_paths)
    c1_s = c1_time()
    dev.set_tuning()  # back to the library defaults
    extra["box_icosphere_256x256"] = {"triangles": int(c1.n_tris), "spp": c1.integrator.effective_spp(),
                                      "max_depth": 5, "paths": c1_paths, "ms": round(c1_s * 1e3, 3),
                                      "mpaths_s": round(c1_paths / c1_s / 1e6, 2), "tuning": c1_tuning,
                                      "default_tuning_mpaths_s": round(c1_paths / c1_default / 1e6, 2)}
    log(f"box + icosphere: {extra['box_icosphere_256x256']}")
    c1_film.free()

    mesh = coherent = incoherent = None
    if not args.no_closest_hit:
        mesh = scenes.displaced_mesh(args.rings, args.segments)
        dev.set_scene(api.Bvh(mesh.verts, n_threads=os.cpu_count() or 8), mesh.normals, mesh.mat_ids)
        cam2 = api.pinhole(1920, 1080, position=(0, 0, 3), look_at=(0, 0, 0), fov=40.0)
        coherent = dev.camera_rays(cam2)
        ch = closest_hit_bench(dev, coherent)
        log(f"closest hit, coherent: {ch}")
        extra["closest_hit_coherent"] = {k: round(v, 4) if isinstance(v, float) else v for k, v in ch.items()}
        incoherent = scenes.incoherent_rays(mesh.verts, 16_777_216, seed=3)
        ch = closest_hit_bench(dev, incoherent)
        log(f"closest hit, incoherent: {ch}")
        extra["closest_hit_incoherent"] = {k: round(v, 4) if isinstance(v, float) else v for k, v in ch.items()}
        incoherent = incoherent[np.sort(np.random.default_rng(21).choice(len(incoherent), 1 << 20, replace=False))]
    film.free()
    dev.close()  # each experiment below opens its own context
    if not args.no_guide:
        extra["guide_experiment"] = guide_experiment(args)
    if not args.no_batch:
        extra["batch"] = batch_bench(args)
    cpu = None
    if not args.no_cpu:
        if mesh is None:
            mesh = scenes.displaced_mesh(args.rings, args.segments)
            coherent = np.zeros((0, 8), np.float32)
            incoherent = scenes.incoherent_rays(mesh.verts, 1 << 20, seed=3)
        cpu = cpu_baselines(sc, c1, mesh, coherent, incoherent)
        log(f"cpu baseline: {cpu}")

    value = paths / wall / 1e6
    out = {
        "metric": "Mpaths/s & time-to-equal-RMSE, guided vs plain GPU PT, 1 MI355X",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_MPATHS, 1),
        "dtype": "f32",
        "data": "synthetic (seeded Perlin-displaced sphere, procedural box)",
        "config": {"workload": "plain path tracing, displaced 1M-triangle mesh with layered dielectric/GGX-gold "
                               "BSDF in a Lambertian box, 800x600, jittered, RR from depth 3, max depth 16",
                   "triangles": int(sc.n_tris), "resolution": f"{cam.width}x{cam.height}",
                   "spp_per_step": sps, "spp_timed": args.steps * sps, "paths_timed": paths,
                   "extension_rays_timed": rays,
                   "baseline_note": "vs_baseline divides by 0.39 Mpaths/s, the reference README's Buddha run "
                                    "(1M tris, 2-layer BSDF, 800x600, i5 CPU)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_extend (closest-hit traversal of extension rays)",
                     "bytes_per_ray": round(bytes_per_ray, 1), "tree_units_32B_per_ray": round(per_ray_nodes, 2),
                     "tri_fetches_per_ray": round(per_ray_tris, 2),
                     "avg_launch_ms": round(ext_ms / max(ext_n, 1), 4), "launches": ext_n,
                     "wavefronts": tn_default["wavefronts"],
                     "aggregate_achieved": round(aggregate, 1), "aggregate_frac": round(aggregate / HBM_PEAK_GBS, 4),
                     "single_wavefront": single},
        "cpu_baseline": cpu,
        "image_mean": [round(float(x), 5) for x in img.reshape(-1, 3).mean(0)],
    }
    if cpu:
        out["cpu_baseline"]["value"] = round(cpu["value"], 4)
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
