"""Throughput of the gfx950 PatchStereo engine on one MI355X.

Workload (the reference hot path's single-GPU configuration): seeded fBm terrain (200 x 200 m,
+-15 m relief, random albedo) seen by 7 pinhole cameras 30 m up at 45 deg oblique, f = 500 px,
1920 x 1080 gray with Gaussian noise sigma = 0.01, quantised to u8. Depth maps for all 7 views,
each view using the other 6 as sources (1 reference + 6 sources, as the geometric pass needs the
sources' depth), window radius 5 step 2, depth range 15-120 m, FP32:
7 photometric iterations from random initialisation, one geometric-consistency round of 2
iterations, read-out filter and count threshold (crates/dense/src/densify.rs:350-454 single-level
schedule), all on the device (c3d_pm_densify).

A step = that whole job. Images are uploaded before the timed region; states stay in HBM.
value = sum of final W*H over the 7 depth maps / step time (Mpx/s).

Other workloads (--workload):
  batched48  48 views of a formation flight, resolutions drawn from 960x540 / 1280x720 /
             1920x1080, 2-8 sources each, one batch per half-step, the whole schedule replayed as a
             hipGraph.
  zone108    108 views (36 stations x 3 cameras) at 1920x1080, 4-8 sources each: plain PatchStereo
             (the default schedule above) against the learned initialiser (network on the quarter-
             resolution image, 2 refinement iterations, one geometric round of 2), the accuracy of
             the learned result against the plain one (<= 1 % relative depth error on >= 99 % of
             the plain method's valid pixels) and the fused clouds of both (median point distance,
             completeness within 5 cm). value = the learned method's Mpx/s.
"""
from __future__ import annotations

import argparse
import dataclasses
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cumulus3d_amd import abi, pm, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per half-step launch on the default workload (7 x 1920x1080, 6 sources, 7 + 1x2
# iterations), averaged over its 18 launches: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors
# from a 512 MiB copy in the same run), profiles/k_half-zone7-7x1080p-6src.md. FETCH_SIZE counts L2
# misses, which the Infinity Cache mostly serves: an upper bound on HBM reads.
K_HALF_TRAFFIC_ZONE7 = 25.995e9
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_SAMPLE = 28  # projective transform 9, reciprocal 1, bilinear 7, scale 1, accumulations 6 (rounded up)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync():
    try:
        hip = C.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
    except OSError:
        pass


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (hipMemcpy, run by the runtime's copy kernel), so
    that the HBM counters of a profile can be scaled by the same counters on a copy of known size."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {mib} MiB device-to-device")


def make_workload(width, height, nviews, noise, seed, window=(5, 2)):
    rig = scenes.oblique_rig(nviews, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0, seed=seed, noise_sigma=noise, noise_seed=seed + 1)
    src = [[j for j in range(nviews) if j != v] for v in range(nviews)]
    pmp, fl = abi.fast_profile()
    pmp = abi.PmParams(**{**pmp.__dict__, "window_radius": window[0], "window_step": window[1]})
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * nviews, pmp, fl)
    return sc, inp, src


BATCH_SIZES = [(960, 540), (1280, 720), (1920, 1080)]


def make_batched_workload(nviews, noise, seed):
    """48 views of a formation flight (16 stations x 3 cameras), resolutions drawn uniformly from
    960x540 / 1280x720 / 1920x1080, 2-8 sources each (uniform), seeded."""
    rng = np.random.default_rng(seed)
    cams = scenes.formation_rig(nviews // 3, 3)
    sizes = [BATCH_SIZES[k] for k in rng.integers(0, 3, len(cams))]
    sc = scenes.make_mixed_scene(cams, sizes, seed=seed, noise_sigma=noise, noise_seed=seed + 1)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(2, 9, len(cams))])
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl)
    return sc, inp, src


def make_fast960_workload(nviews, noise, seed, levels_fn):
    """The reference's published densification setting (fast profile, 3 scales, views of 960 px:
    the V100 timing of 240 views) on synthetic terrain: a formation flight of `nviews` views at
    960x540, each with its 10 nearest overlapping views as sources (NeighborParams::num_views),
    gray pyramids built on the device by `levels_fn`."""
    cams = scenes.formation_rig(nviews // 3, 3)
    sc = scenes.make_mixed_scene(cams, [(960, 540)] * len(cams), seed=seed, noise_sigma=noise, noise_seed=seed + 1)
    src = scenes.nearest_sources(sc, [10] * len(cams))
    pmp, fl = abi.fast_profile()
    nl = pm.num_levels(540, 3, 64)
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl, levels=[levels_fn(v, nl) for v in sc.views])
    return sc, inp, src


def cpu_baseline(inp, rows, threads, photometric_iters, view=0):
    """The CPU oracle (C, OpenMP over rows) on a bounded sample: a band of `rows` rows of `view`,
    random-init photometric run (evaluation of the band and its 24-row reach, then the iterations
    on the band only). Returns the rate in pixel-iterations per second (eval counted as one)."""
    import oracle

    o = oracle.Oracle(inp, threads=threads)
    h, w = inp.level_shape(view, 0)
    st = abi.ViewState(w, h)
    p = abi.run_params(0, False, True, photometric_iters, 1, False)
    r0 = h // 2 - rows // 2
    t = time.perf_counter()
    o.run_rows(view, st, p, r0, r0 + rows)
    dt = time.perf_counter() - t
    px_iters = w * rows * photometric_iters + w * min(h, rows + 48)
    return px_iters / dt, dt, f"view {view}, rows {r0}..{r0 + rows} of {w}x{h}, random init + {photometric_iters} photometric iterations"


def make_zone_workload(noise, seed, stations=36):
    """The 108-view zone: 36 stations x 3 cameras at 1920 x 1080, 4-8 sources each (seeded)."""
    rng = np.random.default_rng(seed)
    cams = scenes.formation_rig(stations, 3, spacing=8.0)
    sc = scenes.make_mixed_scene(cams, [(1920, 1080)] * len(cams), seed=seed, noise_sigma=noise, noise_seed=seed + 1)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(4, 9, len(cams))])
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl)
    return sc, inp, src


def zone_main(args):
    """Plain vs learned-init PatchStereo on the 108-view zone, with fusion and cloud comparison."""
    from cumulus3d_amd import fusion, init_net

    wpath = os.path.join(ROOT, "cumulus3d_amd", "weights", "init_unet.npy")
    if not os.path.exists(wpath):
        print(f"{wpath} is missing: train it with python -m cumulus3d_amd.train_init", file=sys.stderr)
        sys.exit(1)
    t0 = time.perf_counter()
    sc, inp, src = make_zone_workload(0.01, seed=5, stations=(args.views or 108) // 3)
    nviews = len(sc.views)
    views = list(range(nviews))
    job_mpx = sum(v.width * v.height for v in sc.views) / 1e6
    log(f"zone: {nviews} views, {job_mpx:.1f} Mpx rendered in {time.perf_counter() - t0:.1f} s")
    be = pm.HipPatchMatch()
    net = init_net.InitNet(be, 8, 272, 480)
    net.set_params(np.load(wpath, allow_pickle=False).astype(np.float32))
    # read-out post-processing of the fast profile (speckle removal + hole filling, densify.rs:455-462)
    post = abi.PostParams.fast() if args.post else None
    plain_s = abi.Schedule(photometric_iters=args.photometric_iters, geometric_rounds=args.geometric_rounds,
                           geometric_iters=args.geometric_iters, post=post)
    learned_s = init_net.LearnedSchedule(refine_iters=args.refine_iters, geometric_rounds=args.geometric_rounds,
                                         geometric_iters=args.geometric_iters, post=post)
    res = {}
    tune_rows = None
    with be.begin(inp) as s:
        if args.autotune:
            from cumulus3d_amd import autotune

            calib = views[:12]
            t = time.perf_counter()
            learned_s, tune_rows = autotune.tune_learned(s, net, calib, plain_s)
            learned_s = dataclasses.replace(learned_s, post=post)
            log(f"schedule autotuner on {len(calib)} views ({time.perf_counter() - t:.1f} s): "
                + ", ".join(f"{c}: {a:.4f} in {ms:.0f} ms" for c, a, ms in tune_rows)
                + f" -> refine {learned_s.refine_iters}, geometric {learned_s.geometric_rounds}x{learned_s.geometric_iters}, "
                  f"confidence {learned_s.conf_min}")
        # the timed jobs end at the read-out threshold, as the metric does; the downloaded job that
        # feeds the fusion adds the post-processing (densify.rs:455-462)
        timed_p = dataclasses.replace(plain_s, post=None)
        timed_l = dataclasses.replace(learned_s, post=None)
        for name, run in [("plain", lambda dl: s.densify(views, plain_s if dl else timed_p, download=dl)),
                          ("learned", lambda dl: init_net.densify_learned(s, net, views, learned_s if dl else timed_l, download=dl))]:
            for i in range(args.warmup):
                run(False)
            device_sync()
            t = time.perf_counter()
            tms = []
            for i in range(args.steps):
                r = run(False)
                tms.append(r[3])
            device_sync()
            ms = (time.perf_counter() - t) / args.steps * 1e3
            log(f"{name}: {ms:.1f} ms per job, {job_mpx / (ms / 1e3):.2f} Mpx/s")
            outs = run(True)
            res[name] = (ms, outs, tms)
        # two plain runs that differ only in their random streams: the agreement a method can have
        reseeded = s.densify(views, abi.Schedule(photometric_iters=args.photometric_iters, geometric_rounds=args.geometric_rounds,
                                                 geometric_iters=args.geometric_iters, first_run_id=101, post=post))
    # the plain method's read-out mask: raw depth > 0 and count >= max(1, min(min_num_consistent, nsrc))
    # (densify.rs:446-454), before any post-processing
    _, fl = abi.fast_profile()
    valid = [(res["plain"][1][2][v] > 0) & (res["plain"][1][1][v] >= max(1, min(fl.min_num_consistent, len(src[v])))) for v in views]
    # accuracy of the learned maps against the plain ones (raw depth and final normal on that mask)
    ok_d = ok_n = tot = 0
    for v in views:
        dp, dl = res["plain"][1][2][v], res["learned"][1][2][v]
        npl, nle = res["plain"][1][0][v].normal, res["learned"][1][0][v].normal
        m = valid[v]
        tot += int(m.sum())
        rel = np.abs(dl[m] - dp[m]) / dp[m]
        ok_d += int((rel <= 0.01).sum())
        cosang = np.clip((npl[m] * nle[m]).sum(-1), -1, 1)
        ok_n += int((np.degrees(np.arccos(cosang)) <= 2.0).sum())
    frac_d, frac_n = ok_d / max(tot, 1), ok_n / max(tot, 1)
    log(f"learned vs plain: depth within 1 % on {frac_d:.4f}, normal within 2 deg on {frac_n:.4f} of {tot} px")
    rd = rn = 0
    for v in views:
        m = valid[v]
        dp, dr = res["plain"][1][2][v], reseeded[2][v]
        rd += int((np.abs(dr[m] - dp[m]) <= 0.01 * dp[m]).sum())
        cosang = np.clip((res["plain"][1][0][v].normal[m] * reseeded[0][v].normal[m]).sum(-1), -1, 1)
        rn += int((np.degrees(np.arccos(cosang)) <= 2.0).sum())
    base_d, base_n = rd / max(tot, 1), rn / max(tot, 1)
    log(f"plain vs plain with other random streams: depth within 1 % on {base_d:.4f}, normal within 2 deg on {base_n:.4f}")
    # fused clouds (min views 3) and their comparison
    clouds = {}
    for name in ("plain", "learned"):
        outs = res[name][1][0]
        keep = []
        fv = fusion.fusion_views(sc, [o.depth for o in outs], [o.normal for o in outs], [o.cost for o in outs], src, src, keep)
        t = time.perf_counter()
        clouds[name] = fusion.fuse(be, fv, fusion.FusionParams(), int(sum((o.depth > 0).sum() for o in outs)) + 1)
        log(f"{name} fusion: {len(clouds[name].positions)} points in {time.perf_counter() - t:.1f} s")
    med, comp = fusion.compare(clouds["learned"], clouds["plain"], 0.05)
    log(f"clouds: median learned->plain distance {med * 100:.2f} cm, completeness at 5 cm {comp:.4f}")
    ms_l, ms_p = res["learned"][0], res["plain"][0]
    line = {
        "metric": "depth-map Mpx/s at <=1% rel. depth error vs plain PatchStereo; HBM GB/s",
        "value": round(job_mpx / (ms_l / 1e3), 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_l, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm terrain, rendered on the host)",
        "config": {
            "workload": f"zone: {nviews} views 1920x1080, 4-8 sources, learned initialiser (bf16 U-Net at 480x270) + "
                        f"{learned_s.refine_iters} refinement iterations + {learned_s.geometric_rounds}x{learned_s.geometric_iters} geometric, "
                        f"against plain PatchStereo ({args.photometric_iters} + {args.geometric_rounds}x{args.geometric_iters})",
            "views": nviews,
            "megapixels": round(job_mpx, 3),
        },
        "plain_mpx_s": round(job_mpx / (ms_p / 1e3), 3),
        "speedup_vs_plain": round(ms_p / ms_l, 3),
        "accuracy_vs_plain": {"depth_within_1pct": round(frac_d, 5), "normal_within_2deg": round(frac_n, 5),
                              "gate_depth_99pct": frac_d >= 0.99,
                              "plain_reseeded_depth_within_1pct": round(base_d, 5),
                              "plain_reseeded_normal_within_2deg": round(base_n, 5)},
        "learned_schedule": {"refine_iters": learned_s.refine_iters, "geometric_rounds": learned_s.geometric_rounds,
                             "geometric_iters": learned_s.geometric_iters, "conf_min": learned_s.conf_min,
                             "autotuned_on_views": 12 if tune_rows else 0,
                             "candidates": [[list(c), round(a, 5), round(ms, 1)] for c, a, ms in tune_rows] if tune_rows else None},
        "stage_ms_learned": {"init": round(res["learned"][2][-1].init_ms, 2), "photometric": round(res["learned"][2][-1].photometric_ms, 2),
                             "geometric": round(res["learned"][2][-1].geometric_ms, 2), "filter": round(res["learned"][2][-1].filter_ms, 2)},
        "post_processing": {"speckles": bool(post), "fill_holes": bool(post), "timed": False,
                            "plain_ms": round(res["plain"][1][3].post_ms, 2), "learned_ms": round(res["learned"][1][3].post_ms, 2),
                            "plain_removed": int(res["plain"][1][3].post_removed), "plain_filled": int(res["plain"][1][3].post_filled),
                            "learned_removed": int(res["learned"][1][3].post_removed),
                            "learned_filled": int(res["learned"][1][3].post_filled)},
        "clouds": {"plain_points": len(clouds["plain"].positions), "learned_points": len(clouds["learned"].positions),
                   "median_distance_m": round(med, 5), "completeness_5cm": round(comp, 5)},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["zone7", "batched48", "zone108", "fast960"], default="zone7",
                    help="zone7: 7 views 1920x1080, 6 sources each (default); batched48: 48 views of uneven "
                         "resolution and source count in one batch, schedule captured in a hipGraph")
    ap.add_argument("--window", type=int, nargs=2, default=[5, 2], metavar=("RADIUS", "STEP"),
                    help="default workload: NCC window radius and step (5 2 = fast profile, 5 1 = quality profile's window)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--views", type=int, default=None)
    ap.add_argument("--photometric-iters", type=int, default=7)
    ap.add_argument("--geometric-rounds", type=int, default=1)
    ap.add_argument("--geometric-iters", type=int, default=2)
    ap.add_argument("--graph", type=int, default=None, help="1 = replay the schedule as one hipGraph (default: batched48 only)")
    ap.add_argument("--refine-iters", type=int, default=2, help="zone108: refinement iterations after the learned init")
    ap.add_argument("--autotune", type=int, default=1, help="zone108: pick the learned schedule with the autotuner (1) or not (0)")
    ap.add_argument("--tile-autotune", type=int, default=1,
                    help="zone7/batched48: pick the half step's workgroup shape with c3d_pm_autotune before timing (1) or keep (4, 1) (0)")
    ap.add_argument("--post", type=int, default=1, help="zone108: the fast profile's speckle removal and hole filling (1) or none (0)")
    ap.add_argument("--cpu-rows", type=int, default=320)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of this many MiB (profiling calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box here has one); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if not pm.is_available():
        print("no gfx950 device: the engine has no CPU path", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    if args.workload == "zone108":
        zone_main(args)
        return
    batched = args.workload in ("batched48", "fast960")
    fast960 = args.workload == "fast960"
    use_graph = args.graph if args.graph is not None else int(batched)
    t0 = time.perf_counter()
    if fast960:
        nviews = args.views or 240
        pbe = pm.HipPatchMatch()
        sc, inp, src = make_fast960_workload(nviews, 0.01, 6, lambda v, nl: pbe.build_pyramid(v.gray, v.k, nl))
        pbe.close()
    elif batched:
        nviews = args.views or 48
        sc, inp, src = make_batched_workload(nviews, 0.01, seed=3)
    else:
        nviews = args.views or 7
        sc, inp, src = make_workload(args.width, args.height, nviews, 0.01, seed=1, window=tuple(args.window))
    top = inp.num_levels - 1
    sizes = [inp.level_shape(v, top) for v in range(nviews)]
    npix_v = [h * w for (h, w) in sizes]
    job_mpx = sum(npix_v) / 1e6
    log(f"scene {nviews} views, {job_mpx:.1f} Mpx rendered in {time.perf_counter() - t0:.1f} s")
    views = list(range(nviews))
    if fast960:
        # LevelSchedule of the fast profile (params.rs:496-503): 6 / 3 photometric iterations,
        # 2 geometric rounds of 2 per scale
        sched = abi.Schedule.fast()
    else:
        sched = abi.Schedule(photometric_iters=args.photometric_iters, geometric_rounds=args.geometric_rounds,
                             geometric_iters=args.geometric_iters)
    opts = abi.c3d_pm_options()
    opts.use_graph = use_graph
    be = pm.HipPatchMatch(0, opts)
    sess = be.begin(inp)
    log(f"session open on {be.name()}, graph {use_graph}")
    shape = (4, 1)
    if args.tile_autotune:
        # the half step's workgroup shape (rows, lanes per pixel), outside the timed region
        shape, shape_ms = sess.autotune(views, level=inp.num_levels - 1)  # the finest scale dominates
        log(f"tile autotuner: (rows, lanes per pixel) {shape}; two iterations (2,1) {shape_ms[0]:.1f} ms, "
            f"(4,1) {shape_ms[1]:.1f} ms, (4,2) {shape_ms[2]:.1f} ms")
    for i in range(args.warmup):
        sess.densify(views, sched, download=False)
        log(f"warmup {i + 1}/{args.warmup} done")
    device_sync()
    times, tms = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, _, _, tm = sess.densify(views, sched, download=False)
        times.append(time.perf_counter() - ts)
        tms.append(tm)
        log(f"step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms (device {tm.total_ms:.1f} ms)")
    device_sync()
    wall = time.perf_counter() - t_all
    ms_per_step = wall / args.steps * 1e3
    value = job_mpx / (ms_per_step / 1e3)
    if use_graph:
        # a graph replay has no per-launch events: one eager run outside the timed region gives the
        # half-step launch times and the stage split
        sess.close()
        be.close()
        eopts = abi.c3d_pm_options()
        eopts.tile_rows, eopts.threads_per_pixel = shape
        be = pm.HipPatchMatch(0, eopts)
        sess = be.begin(inp)
        _, _, _, tm = sess.densify(views, sched, download=False)
        tms = [tm]
        log(f"eager run for kernel times: {tm.total_ms:.1f} ms")

    # dominant kernel: the red-black half step. Algorithmic work per pixel-iteration (one pixel
    # updated once by a red + black pair): S = 15 nsrc WN^2 bilinear samples (upper bound: 9
    # propagated + 6 refinement hypotheses over every source), 28 FLOP each; compulsory HBM bytes
    # 58 + 2 (1 + nsrc), plus 8 nsrc for geometric iterations. One launch updates npix/2 pixels
    # of every view.
    wn = (2 * inp.pm.window_radius) // inp.pm.window_step + 1
    # every half-step launch of the schedule: (pixels of each view at its scale, geometric?)
    launches = []
    for lv in range(inp.num_levels):
        pix_l = [h * w for (h, w) in (inp.level_shape(v, lv) for v in range(nviews))]
        n_ph = 2 * (sched.photometric_iters if lv == 0 else sched.photometric_iters_upper)
        launches += [(pix_l, False)] * n_ph + [(pix_l, True)] * (2 * sched.geometric_rounds * sched.geometric_iters)
    avg_bytes = sum(sum(p / 2 * (58 + 2 * (1 + len(s)) + (8 * len(s) if geo else 0)) for p, s in zip(pix_l, src))
                    for pix_l, geo in launches) / len(launches)
    flop = sum(sum(FLOP_PER_SAMPLE * 15 * len(s) * wn * wn * p / 2 for p, s in zip(pix_l, src)) for pix_l, _ in launches) / len(launches)
    half_ms = sum(t.half_kernel_ms for t in tms) / sum(t.half_launches for t in tms)
    achieved_gbs = avg_bytes / (half_ms / 1e3) / 1e9
    default_zone7 = (not batched and nviews == 7 and (args.width, args.height) == (1920, 1080) and args.photometric_iters == 7
                     and args.geometric_rounds == 1 and args.geometric_iters == 2 and tuple(args.window) == (5, 2))
    achieved_tflops = flop / (half_ms / 1e3) / 1e12
    # the windows built with register-held weights; every other one runs the any-radius build
    reg_windows = {(5, 2), (4, 2), (3, 2), (2, 2), (3, 1), (2, 1), (1, 1)}
    kernel_name = (f"k_half<{inp.pm.window_radius},{inp.pm.window_step}>"
                   if (inp.pm.window_radius, inp.pm.window_step) in reg_windows else f"k_half<any radius,{inp.pm.window_step}>")
    nsrc_desc = f"{len(src[0])} src each" if len({len(s) for s in src}) == 1 else \
        f"{min(len(s) for s in src)}-{max(len(s) for s in src)} src"
    size_desc = f"{sizes[0][1]}x{sizes[0][0]}" if len(set(sizes)) == 1 else \
        "mixed " + "/".join(f"{w}x{h}" for (w, h) in BATCH_SIZES)
    if fast960:
        iters_desc = (f"fast profile over {inp.num_levels} scales ({sched.photometric_iters} / {sched.photometric_iters_upper} "
                      f"photometric, {sched.geometric_rounds}x{sched.geometric_iters} geometric per scale)")
    else:
        iters_desc = (f"{args.photometric_iters} photometric + {args.geometric_rounds}x{args.geometric_iters} "
                      f"geometric iterations")
    line = {
        "metric": "depth-map Mpx/s at <=1% rel. depth error vs plain PatchStereo; HBM GB/s",
        "value": round(value, 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm terrain, rendered on the host)",
        "config": {
            "workload": f"plain PatchStereo, {nviews} views {size_desc} u8, 1 ref + {nsrc_desc}, "
                        f"window {inp.pm.window_radius} step {inp.pm.window_step}, {iters_desc}, read-out filter, depth 15-120 m"
                        + (", hipGraph replay" if use_graph else ""),
            "views": nviews,
            "megapixels": round(job_mpx, 3),
            "sources": [len(s) for s in src] if len({len(s) for s in src}) > 1 else len(src[0]),
            "photometric_iters": sched.photometric_iters,
            "geometric_rounds": sched.geometric_rounds,
            "geometric_iters": sched.geometric_iters,
            "scales": inp.num_levels,
            "graph": use_graph,
            "tile": {"rows": shape[0], "lanes_per_pixel": shape[1], "autotuned": bool(args.tile_autotune)},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": round(achieved_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
            "traffic": K_HALF_TRAFFIC_ZONE7 if default_zone7 else None,
            "avg_launch_ms": round(half_ms, 3),
            "algorithmic_bytes_per_launch": int(avg_bytes),
        },
        "roofline_valu": {
            "bound": "valu-fp32",
            "kernel": kernel_name,
            "achieved": round(achieved_tflops, 2),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
            "algorithmic_flop_per_launch": int(flop),
        },
        "stage_ms": {
            "photometric": round(sum(t.photometric_ms for t in tms) / len(tms), 2),
            "geometric": round(sum(t.geometric_ms for t in tms) / len(tms), 2),
            "filter": round(sum(t.filter_ms for t in tms) / len(tms), 2),
        },
    }
    sess.close()
    if not args.no_cpu and not fast960:
        threads = min(os.cpu_count() or 1, 16)
        view = max(range(nviews), key=lambda v: npix_v[v])
        rate, dt, sample = cpu_baseline(inp, args.cpu_rows, threads, args.photometric_iters, view)
        job_px_iters = sum(npix_v) * (args.photometric_iters + 1 + args.geometric_rounds * (args.geometric_iters + 1) + 1)
        cpu_mpxs = job_mpx / (job_px_iters / rate)
        line["cpu_baseline"] = {
            "value": round(cpu_mpxs, 5),
            "unit": "Mpx/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{sample}: {dt:.1f} s, {rate:.0f} px-iter/s, scaled to the job's "
                      f"{job_px_iters / 1e6:.1f} M px-iters (the reference has no CPU PatchStereo; this is the C oracle, "
                      f"{platform.processor() or platform.machine()})",
        }
        log(f"cpu baseline: {rate:.0f} px-iter/s on {threads} threads ({dt:.1f} s)")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
