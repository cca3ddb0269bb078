"""Benchmark: plain GPU path tracing on one MI355X.

Workload: the 1,048,576-triangle Perlin-displaced sphere with a 2-layer BSDF
(smooth dielectric, IOR 1.5, over GGX alpha 0.2 gold) inside the Lambertian
box with an emissive ceiling quad, 800x600, jittered, Russian roulette from
depth 3, max depth 16. One step = 16 samples of every pixel (7.68M camera
paths); the default 64 steps render the full 1024 spp. Inputs (tree,
triangles, materials) are resident in HBM before timing.

Also reported (same JSON line): closest-hit throughput on 16,777,216
incoherent rays over the same mesh, the roofline of the dominant kernel (the
closest-hit extension kernel inside the path tracer) and the oracle (the CPU
restatement of the reference's OpenMP path tracer) timed on the host cores.

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


def cpu_baseline(sc, target_s=10.0):
    """Oracle (reference semantics: recursive traversal, heap nodes, virtual
    primitives, OpenMP dynamic rows) on a viewport slice of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    import ctypes as C

    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    orc = O.Scene(sc.verts, sc.normals, sc.mat_ids, materials=sc.materials)
    build_s = time.perf_counter() - t0
    cam = sc.camera
    vw, vh = 80, 60
    vp = (cam.height // 2 - vh // 2, cam.width // 2 - vw // 2, vw, vh)
    p = sc.integrator.params()
    p.viewport = (C.c_uint32 * 4)(*vp)
    # probe with 1 sample per pixel, then size the sample to ~target_s
    p.spp_begin, p.spp_count = 0, 1
    t0 = time.perf_counter()
    orc.render(cam, p, threads=threads)
    probe = time.perf_counter() - t0
    spp = int(max(1, min(1024, target_s / max(probe, 1e-3))))
    p.spp_count = spp
    t0 = time.perf_counter()
    orc.render(cam, p, threads=threads)
    dt = time.perf_counter() - t0
    paths = vw * vh * spp
    orc.close()
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{vw}x{vh} centre viewport x {spp} spp = {paths} paths in {dt:.2f} s "
                      f"(tree build {build_s:.1f} s not timed)"}


def closest_hit_bench(dev, mesh, n_rays, reps=5):
    from swpathtracer_amd import api, scenes
    import swpathtracer_amd._lib as L

    bvh = api.Bvh(mesh.verts, n_threads=os.cpu_count() or 8)
    dev.set_scene(bvh, mesh.normals, mesh.mat_ids)
    rays = scenes.incoherent_rays(mesh.verts, n_rays, seed=3)
    d_rays = dev.to_device(rays)
    del rays
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
    return {"rays": n_rays, "ms": ms, "grays_per_s": n_rays / (ms * 1e-3) / 1e9,
            "mean_tree_units_32B": nodes.mean(), "mean_tris": tris.mean(), "hit_frac": float((hits[:, 1] != L.SPT_MISS).mean()),
            "fetch_gbs": byts / (ms * 1e-3) / 1e9}


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

    # reference image: FP64 film, independent seed
    osp = args.oracle_spp
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
    res = {"scene": f"slit box, {sc.n_tris} triangles, {W}x{H}", "reference_spp": osp,
           "plain_ms": round(t_plain * 1e3, 2), "plain_rmse": round(r_plain, 6),
           "plain_mpaths_s": round(npix * 1024 / t_plain / 1e6, 3),
           "guided_spp_to_equal_rmse": hit[0] if hit else None,
           "guided_ms_to_equal_rmse": round(hit[1] * 1e3, 2) if hit else None,
           "time_ratio": round(hit[1] / t_plain, 4) if hit else None,
           "guided_train_ms": round(t_train * 1e3, 2), "train_steps": st["steps"],
           "warmup_passes": args.guide_warmup, "warmup_ms": round(t_warm * 1e3, 2),
           "train_kernel_ms": round(st["train_kernel_ms"], 3), "trained_records": st["trained_records"],
           "train_mfma_util": round(mfma, 6), "final_loss": st["last_loss"],
           "curve": curve[:: max(1, len(curve) // 24)]}
    log(f"guide experiment: {res}")
    dev.close()
    return res


def batch_bench(args):
    """64 independent scenes (box + seeded icosphere or displaced sphere,
    1k-500k triangles log-uniform, 128^2-1024^2, 16/64/256 spp, Lambertian /
    GGX / dielectric objects) rendered by one wavefront after autotuning tile
    and queue sizes on a 1/64 sample of every scene's samples."""
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
    ap.add_argument("--oracle-spp", type=int, default=8192, help="spp of the FP64-film reference image")
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
    # algorithmic bytes of the extension kernel: one instrumented step
    step(0, L.SPT_RENDER_FILM32 | L.SPT_RENDER_FETCH_STATS)
    dev.sync()
    fs = dev.render_stats()
    per_ray_nodes = fs["node_fetches"] / fs["rays"]
    per_ray_tris = fs["tri_fetches"] / fs["rays"]
    # per extension ray: queue index 4 + origin 16 + direction 16 in, hit 16 + sort key 4 out
    bytes_per_ray = 56 + 32 * per_ray_nodes + 48 * per_ray_tris  # tree data counted in 32-B units
    achieved = bytes_per_ray * rays / (ext_ms * 1e-3) / 1e9
    img = film.download(np.float32, (cam.height, cam.width, 3)) / (args.steps * sps)

    # HBM bytes per extension launch of this workload, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-calibrated)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_pt_mesh_in_box.json")
    if os.path.exists(tf) and args.rings == 512 and args.segments == 1024:
        with open(tf) as fh:
            tj = json.load(fh)
        if tj.get("spp_per_step") == sps:  # measured on this workload's launches
            traffic = tj["hbm_bytes_per_launch"]

    extra = {}
    if not args.no_closest_hit:
        mesh = scenes.displaced_mesh(args.rings, args.segments)
        ch = closest_hit_bench(dev, mesh, 16_777_216)
        log(f"closest hit: {ch}")
        extra["closest_hit_incoherent"] = {k: round(v, 4) if isinstance(v, float) else v for k, v in ch.items()}
    film.free()
    dev.close()  # each experiment below opens its own context
    if not args.no_guide:
        extra["guide_experiment"] = guide_experiment(args)
    if not args.no_batch:
        extra["batch"] = batch_bench(args)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc)
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
                     "avg_launch_ms": round(ext_ms / max(ext_n, 1), 4), "launches": ext_n},
        "cpu_baseline": cpu,
        "image_mean": [round(float(x), 5) for x in img.reshape(-1, 3).mean(0)],
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
