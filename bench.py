"""Throughput of the gfx950 PatchStereo engine on one MI355X.

Default workload (zone108, the configuration the metric is quoted on): the full zone of a
formation flight, 36 stations x 3 cameras = 108 views at 1920 x 1080 over one seeded fBm terrain
(200 x 200 m, +-15 m relief, random albedo; Gaussian noise sigma = 0.01, quantised to u8), each view
with its 4-8 nearest overlapping views as sources (seeded), window radius 5 step 2, depth range
15-120 m, FP32. Method: the learned initialiser (bf16 U-Net on the quarter-resolution image and a
32-depth plane sweep of two sources) + batched PatchStereo (one launch per half step over all 108
views) + a geometric-consistency round + read-out filter, all on the device
(c3d_pm_densify_learned). The schedule after the network is picked by the autotuner on 12 of the
views: the fastest candidate that keeps a geometric round and reaches the accuracy gate.

A step = that whole job with images resident in HBM (uploaded before the timed region) and states
kept on the device. value = sum of W*H over the 108 depth maps / step time (Mpx/s).

Measured next to it, in the same run:
  plain       plain PatchStereo on the same views: 7 photometric iterations from random init, one
              geometric round of 2 iterations, read-out (c3d_pm_densify) - the accuracy reference
              (<= 1 % relative depth error on >= 99 % of its valid pixels is the gate);
  random      the reference's random initialisation on the learned arm's own schedule (the control
              arm: what the network's start is worth), and the fastest random-init schedule that
              reaches the gate (chosen by the same autotuner);
  begin->download  one learned job timed from c3d_pm_begin (image upload, source tables) to the
              downloaded read-out, as the metric's Mpx/s is defined on the host boundary;
  fusion      agreement fusion (min views 3) of the learned and plain maps and the cloud comparison
              (median point distance, completeness within 5 cm);
  cpu_baseline  the CPU oracle (C, OpenMP) on a band of one view (the reference has no CPU
              PatchStereo), scaled to the plain job.

Other workloads (--workload):
  zone7      7 views 1920x1080, each with the other 6 as sources, 7 photometric iterations + one
             geometric round of 2, read-out (plain PatchStereo on one GPU, the single-reference
             configuration with noise sigma = 0.01).
  batched48  48 views of a formation flight, resolutions drawn from 960x540 / 1280x720 /
             1920x1080, 2-8 sources each, one batch per half-step, the whole schedule replayed as a
             hipGraph.
  fast960    the fast profile over 3 scales on 240 views of 960x540 (the reference's published
             densification setting).
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cumulus3d_amd import abi, pm, scenes  # noqa: E402

METRIC = "depth-map Mpx/s at <=1% rel. depth error vs plain PatchStereo; HBM GB/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per half-step launch, from rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the
# factors from a 512 MiB device copy in the same run); FETCH_SIZE counts L2 misses, which the
# Infinity Cache mostly serves: an upper bound on HBM reads.
# Keyed by workload and the half step's workgroup shape (rows, lanes per pixel), which changes the
# fetched bytes; other shapes report null:
#   zone7, (1, 1): 7 x 1920x1080, 6 sources, 7 + 1x2 iterations, averaged over its 18 launches
#          (profiles/k_half-zone7-7x1080p-6src.md)
#   zone108_1_1x1, (1, 1): the 108-view zone's learned job, 1 refinement + 1x1 geometric
#          iterations, averaged over its 4 launches (profiles/k_half-zone108-learned.md)
K_HALF_TRAFFIC = {("zone7", (1, 1)): 42.721e9, ("zone108_1_1x1", (1, 1)): 86.917e9}
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_SAMPLE = 28  # projective transform 9, reciprocal 1, bilinear 7, scale 1, accumulations 6 (rounded up)
WN_FAST = 6  # samples per window side at radius 5 step 2


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def shapes_text(shape_ms):
    """The tile autotuner's timings, one per shape of abi.TILE_SHAPES."""
    from cumulus3d_amd import abi
    return ", ".join(f"({r},{t}) {ms:.1f} ms" for (r, t), ms in zip(abi.TILE_SHAPES, shape_ms))


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


def host_cpu():
    """(threads this process may use, logical CPUs of the host, CPU model)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, os.cpu_count() or avail, model


# ---------------------------------------------------------------- workloads


def make_workload(width, height, nviews, noise, seed, window=(5, 2)):
    rig = scenes.oblique_rig(nviews, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0, seed=seed, noise_sigma=noise, noise_seed=seed + 1, device=True)
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
    sc = scenes.make_mixed_scene(cams, sizes, seed=seed, noise_sigma=noise, noise_seed=seed + 1, device=True)
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
    sc = scenes.make_mixed_scene(cams, [(960, 540)] * len(cams), seed=seed, noise_sigma=noise, noise_seed=seed + 1, device=True)
    src = scenes.nearest_sources(sc, [10] * len(cams))
    pmp, fl = abi.fast_profile()
    nl = pm.num_levels(540, 3, 64)
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl, levels=[levels_fn(v, nl) for v in sc.views])
    return sc, inp, src


def make_zone_workload(noise, seed, stations=36, device=True, size=(1920, 1080)):
    """The zone: `stations` formation stations x 3 cameras (108 views at 36) at 1920 x 1080, each
    view with its 4-8 nearest overlapping views as sources (seeded)."""
    rng = np.random.default_rng(seed)
    cams = scenes.formation_rig(stations, 3, spacing=8.0)
    sc = scenes.make_mixed_scene(cams, [size] * len(cams), seed=seed, noise_sigma=noise, noise_seed=seed + 1, device=device)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(4, 9, len(cams))])
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl)
    return sc, inp, src


# ---------------------------------------------------------------- roofline and CPU baseline


def half_step_work(inp, src, launches):
    """Average algorithmic bytes and FLOP per half-step launch over `launches` = [(level,
    geometric?)]. Per pixel and half step: S = 15 nsrc WN^2 bilinear samples (upper bound: 9
    propagated + 6 refinement hypotheses over every source), 28 FLOP each; compulsory HBM bytes
    58 + 2 (1 + nsrc), plus 8 nsrc for geometric iterations. A launch updates npix/2 pixels of
    every view."""
    wn = (2 * inp.pm.window_radius) // inp.pm.window_step + 1
    nv = len(src)
    tot_b = tot_f = 0.0
    for lv, geo in launches:
        for v in range(nv):
            h, w = inp.level_shape(v, lv)
            ns = len(src[v])
            tot_b += h * w / 2 * (58 + 2 * (1 + ns) + (8 * ns if geo else 0))
            tot_f += FLOP_PER_SAMPLE * 15 * ns * wn * wn * h * w / 2
    return tot_b / len(launches), tot_f / len(launches)


def roofline_objects(kernel, avg_bytes, flop, half_ms, traffic):
    achieved_gbs = avg_bytes / (half_ms / 1e3) / 1e9
    achieved_tflops = flop / (half_ms / 1e3) / 1e12
    return ({"bound": "hbm", "kernel": kernel, "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved_gbs / HBM_PEAK_GBS, 5), "traffic": traffic, "avg_launch_ms": round(half_ms, 3),
             "algorithmic_bytes_per_launch": int(avg_bytes)},
            {"bound": "valu-fp32", "kernel": kernel, "achieved": round(achieved_tflops, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
             "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4), "algorithmic_flop_per_launch": int(flop)})


def cpu_baseline(inp, rows, threads, photometric_iters, view=0):
    """The CPU oracle (C, OpenMP over rows) on a bounded sample: a band of `rows` rows of `view`,
    random-init photometric run (evaluation of the band and its 24-row reach, then the iterations
    on the band only). Returns (pixel-iterations per second (eval counted as one), seconds, text)."""
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


def cpu_baseline_object(inp, src, views, rows, photometric_iters, geometric_rounds, geometric_iters, job_mpx):
    """cpu_baseline for a plain job over `views` (all at the top scale of a single-level input).
    The sample's rate per pixel-iteration and source is scaled to the job's pixel-iterations x
    sources (the cost of a pixel-iteration grows with its view's source count); a geometric
    iteration is counted as a photometric one (it costs more: the figure is generous to the CPU)."""
    threads, ncpu, model = host_cpu()
    nsrc_med = sorted(len(src[v]) for v in views)[len(views) // 2]
    view = min(views, key=lambda v: (abs(len(src[v]) - nsrc_med), -np.prod(inp.level_shape(v, 0)), v))
    rate, dt, sample = cpu_baseline(inp, rows, threads, photometric_iters, view)
    rate_src = rate * len(src[view])
    per_px = photometric_iters + 1 + geometric_rounds * (geometric_iters + 1) + 1  # evals, iterations, read-out
    job_work = sum(np.prod(inp.level_shape(v, 0)) * per_px * len(src[v]) for v in views)
    cpu_s = job_work / rate_src
    log(f"cpu baseline: {rate:.0f} px-iter/s with {len(src[view])} sources on {threads} threads ({dt:.1f} s)")
    return {
        "value": round(job_mpx / cpu_s, 5),
        "unit": "Mpx/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample}, {len(src[view])} sources: {dt:.1f} s, {rate:.0f} px-iter/s; scaled per pixel-iteration and source "
                  f"to the plain job ({job_work / 1e9:.2f} G px-iter-sources, {cpu_s:.0f} s). The reference has no CPU "
                  f"PatchStereo: this is the C oracle (OpenMP, {threads} threads of {ncpu} logical CPUs, {model}).",
        "host_logical_cpus": ncpu,
        "cpu_model": model,
    }


# ---------------------------------------------------------------- zone (default)


def agreement(plain_outs, plain_raw, raws, plain_normals=None, normals=None):
    """Depth agreement (within 1 %) of `raws` with the plain raw depth on the plain method's valid
    pixels (read-out depth > 0), and normal agreement (within 2 deg) when normals are given."""
    ok_d = ok_n = tot = 0
    for i, (o, dp, dl) in enumerate(zip(plain_outs, plain_raw, raws)):
        m = o.depth > 0
        tot += int(m.sum())
        ok_d += int((np.abs(dl[m] - dp[m]) <= 0.01 * dp[m]).sum())
        if normals is not None:
            cosang = np.clip((plain_normals[i][m] * normals[i][m]).sum(-1), -1, 1)
            ok_n += int((cosang >= np.cos(np.radians(2.0))).sum())
    return ok_d / max(tot, 1), (ok_n / max(tot, 1) if normals is not None else None), tot


def timed(name, run, steps, warmup):
    """Runs `run()` warmup + steps times; returns (ms per step, timings of the timed steps)."""
    for i in range(warmup):
        run()
    device_sync()
    tms = []
    t = time.perf_counter()
    for i in range(steps):
        tms.append(run())
    device_sync()
    ms = (time.perf_counter() - t) / max(steps, 1) * 1e3
    log(f"{name}: {ms:.1f} ms per job ({steps} steps, {warmup} warmup)")
    return ms, tms


def confirm(rows, gate, run_full, name):
    """The fastest calibration candidate (rows = [(candidate, agreement, ms)]) whose agreement on
    every view of the job also reaches `gate`: candidates passing on the calibration views are
    tried fastest first on the whole job. Returns (candidate, full agreement, run_full's output,
    [(candidate, full agreement)] tried). When none passes, the most accurate one tried (or, when
    none passed calibration, the most accurate calibration candidate)."""
    tried = []
    best = None
    for cand, a, ms in sorted([r for r in rows if r[1] >= gate], key=lambda r: r[2]):
        acc, out = run_full(cand)
        tried.append((cand, acc))
        log(f"{name} {cand}: {acc:.4f} on every view (calibration {a:.4f}, {ms:.0f} ms)")
        if acc >= gate:
            return cand, acc, out, tried
        if best is None or acc > best[1]:
            best = (cand, acc, out)
    if best is None:
        cand = max(rows, key=lambda r: r[1])[0]
        acc, out = run_full(cand)
        tried.append((cand, acc))
        best = (cand, acc, out)
    return best[0], best[1], best[2], tried


def zone_main(args):
    from cumulus3d_amd import autotune, fusion, init_net

    wpath = os.path.join(ROOT, "cumulus3d_amd", "weights", "init_unet.npy")
    if not os.path.exists(wpath):
        print(f"{wpath} is missing: train it with python -m cumulus3d_amd.train_init", file=sys.stderr)
        sys.exit(1)
    t0 = time.perf_counter()
    stations = (args.views or 108) // 3
    sc, inp, src = make_zone_workload(0.01, seed=5, stations=stations)
    nviews = len(sc.views)
    views = list(range(nviews))
    job_mpx = sum(v.width * v.height for v in sc.views) / 1e6
    log(f"zone: {nviews} views, {job_mpx:.1f} Mpx rendered on the device in {time.perf_counter() - t0:.1f} s; "
        f"sources {min(len(s) for s in src)}-{max(len(s) for s in src)}")
    be = pm.HipPatchMatch()
    net = init_net.InitNet(be, 8, 272, 480)
    net.set_params(np.load(wpath, allow_pickle=False).astype(np.float32))
    plain_s = abi.Schedule(photometric_iters=args.photometric_iters, geometric_rounds=args.geometric_rounds,
                           geometric_iters=args.geometric_iters)
    calib = views[: min(12, nviews)]
    s = be.begin(inp)
    shape = (1, 1)  # the engine's default shape
    if args.tile_autotune:
        shape, shape_ms = s.autotune(calib, level=0)
        log(f"tile autotuner on {len(calib)} views: (rows, lanes per pixel) {shape}; " + shapes_text(shape_ms))
    gate = 0.99
    line_extra = {}
    if args.learned_only:
        ri, gr, gi, cm = args.learned_schedule or (1, 1, 1, 0.0)
        learned_s = init_net.LearnedSchedule(refine_iters=int(ri), geometric_rounds=int(gr), geometric_iters=int(gi), conf_min=cm)
    else:
        # the plain method's read-out on every view: the accuracy reference of every arm
        p_outs, _, p_raw, _ = s.densify(views, plain_s)
        rs_outs, _, rs_raw, _ = s.densify(views, dataclasses.replace(plain_s, first_run_id=101))
        pn = [o.normal for o in p_outs]
        acc_rs, nrm_rs, npx = agreement(p_outs, p_raw, rs_raw, pn, [o.normal for o in rs_outs])
        del rs_outs, rs_raw
        log(f"plain on {npx} valid pixels; plain with other random streams agrees on {acc_rs:.4f} (normals {nrm_rs:.4f})")

        def run_learned(c):
            outs, _, raw, _, _, _ = init_net.densify_learned(s, net, views, init_net.LearnedSchedule(*c[:3], conf_min=c[3]))
            return agreement(p_outs, p_raw, raw)[0], (outs, raw)

        def run_random(c):
            _, _, raw, _ = s.densify(views, abi.Schedule(photometric_iters=c[0], geometric_rounds=c[1], geometric_iters=c[2]))
            return agreement(p_outs, p_raw, raw)[0], raw

        if args.learned_schedule:
            ri, gr, gi, cm = args.learned_schedule
            lc = (int(ri), int(gr), int(gi), cm)
            acc_l, (l_outs, l_raw) = run_learned(lc)
            rc = (int(ri) + 1, int(gr), int(gi))
            acc_r, r_raw = run_random(rc)
            l_rows = r_rows = l_tried = r_tried = []
        else:
            # schedule autotuner on the calibration views, then confirmed on every view; the
            # random-init control arm goes through the same two steps
            t = time.perf_counter()
            ref = autotune.plain_reference(s, calib, plain_s)
            _, l_rows = autotune.tune_learned(s, net, calib, plain_s, gate=gate, reference=ref)
            _, r_rows = autotune.tune_random(s, calib, plain_s, gate=gate, reference=ref)
            log(f"schedule autotuner on {len(calib)} views ({time.perf_counter() - t:.1f} s): learned "
                + ", ".join(f"{c}: {a:.4f} {ms:.0f} ms" for c, a, ms in l_rows) + "; random "
                + ", ".join(f"{c}: {a:.4f} {ms:.0f} ms" for c, a, ms in r_rows))
            lc, acc_l, (l_outs, l_raw), l_tried = confirm(l_rows, gate, run_learned, "learned")
            rc, acc_r, r_raw, r_tried = confirm(r_rows, gate, run_random, "random init")
        learned_s = init_net.LearnedSchedule(refine_iters=lc[0], geometric_rounds=lc[1], geometric_iters=lc[2], conf_min=lc[3])
        random_s = abi.Schedule(photometric_iters=rc[0], geometric_rounds=rc[1], geometric_iters=rc[2])
        # the random-init twin of the learned schedule: the same iterations from the reference's start
        twin_s = abi.Schedule(photometric_iters=learned_s.refine_iters, geometric_rounds=learned_s.geometric_rounds,
                              geometric_iters=learned_s.geometric_iters)
        acc_t, _ = run_random((twin_s.photometric_iters, twin_s.geometric_rounds, twin_s.geometric_iters))
        nrm_l = agreement(p_outs, p_raw, l_raw, pn, [o.normal for o in l_outs])[1]
        log(f"chosen: learned {lc} ({acc_l:.4f} on every view), random init {rc} ({acc_r:.4f}); random init on the learned "
            f"schedule {acc_t:.4f}")
        line_extra["autotune"] = {
            "calibration_views": len(calib) if l_rows else 0,
            "learned": [[list(c), round(a, 5), round(ms, 1)] for c, a, ms in l_rows],
            "random": [[list(c), round(a, 5), round(ms, 1)] for c, a, ms in r_rows],
            "confirmed_on_every_view": {"learned": [[list(c), round(a, 5)] for c, a in l_tried],
                                        "random": [[list(c), round(a, 5)] for c, a in r_tried]}}

    # ---- timed arms (device-resident: images uploaded at begin, states stay in HBM)
    ms_l, tms_l = timed("learned", lambda: init_net.densify_learned(s, net, views, learned_s, download=False)[3],
                        args.steps, args.warmup)
    # roofline of the dominant kernel (the half step) from the learned arm's own launches
    launches = [(0, False)] * (2 * learned_s.refine_iters) + [(0, True)] * (2 * learned_s.geometric_rounds * learned_s.geometric_iters)
    avg_bytes, flop = half_step_work(inp, src, launches)
    half_ms = sum(t.half_kernel_ms for t in tms_l) / max(sum(t.half_launches for t in tms_l), 1)
    sched_key = f"zone{nviews}_{learned_s.refine_iters}_{learned_s.geometric_rounds}x{learned_s.geometric_iters}"
    roof, roof_valu = roofline_objects("k_half<5,2>", avg_bytes, flop, half_ms, K_HALF_TRAFFIC.get((sched_key, tuple(shape))))
    last = tms_l[-1]
    line = {
        "metric": METRIC,
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
        "data": "synthetic (seeded fBm terrain rendered on the device, sigma 0.01 noise, u8)",
        "config": {
            "workload": f"zone: {nviews} views 1920x1080 ({stations} stations x 3 cameras), 4-8 sources each, learned initialiser "
                        f"(bf16 U-Net at 480x270, 32-depth sweep) + {learned_s.refine_iters} refinement + "
                        f"{learned_s.geometric_rounds}x{learned_s.geometric_iters} geometric iterations + read-out, "
                        f"one launch per half step over all views; accuracy gate against plain PatchStereo "
                        f"({args.photometric_iters} + {args.geometric_rounds}x{args.geometric_iters})",
            "views": nviews,
            "megapixels": round(job_mpx, 3),
            "sources": [len(x) for x in src],
            "tile": {"rows": shape[0], "lanes_per_pixel": shape[1], "autotuned": bool(args.tile_autotune)},
            "learned_schedule": {"refine_iters": learned_s.refine_iters, "geometric_rounds": learned_s.geometric_rounds,
                                 "geometric_iters": learned_s.geometric_iters, "conf_min": learned_s.conf_min},
        },
        "roofline": roof,
        "roofline_valu": roof_valu,
        "stage_ms_learned": {"init": round(last.init_ms, 2), "photometric": round(last.photometric_ms, 2),
                             "geometric": round(last.geometric_ms, 2), "filter": round(last.filter_ms, 2)},
    }
    line.update(line_extra)
    if args.learned_only:
        s.close()
        print(json.dumps(line), flush=True)
        return
    ms_p, _ = timed("plain", lambda: s.densify(views, plain_s, download=False)[3], args.plain_steps, 1)
    ms_t, _ = timed("random init, learned schedule", lambda: s.densify(views, twin_s, download=False)[3], args.plain_steps, 1)
    ms_r, _ = timed("random init, tuned schedule", lambda: s.densify(views, random_s, download=False)[3], args.plain_steps, 1)

    # ---- fused clouds (min views 3) of the learned and plain maps, after the fast profile's
    # read-out post-processing (speckle removal and hole filling, densify.rs:455-462), and their
    # comparison; untimed
    if args.fusion:
        post = abi.PostParams.fast() if args.post else None
        if post is not None:
            del p_outs, l_outs
            p_outs, _, _, p_tm = s.densify(views, dataclasses.replace(plain_s, post=post))
            l_outs, _, _, l_tm, _, _ = init_net.densify_learned(s, net, views, dataclasses.replace(learned_s, post=post))
            line["post_processing"] = {"speckles": True, "fill_holes": True, "timed": False,
                                       "plain_ms": round(p_tm.post_ms, 2), "learned_ms": round(l_tm.post_ms, 2),
                                       "plain_removed": int(p_tm.post_removed), "plain_filled": int(p_tm.post_filled),
                                       "learned_removed": int(l_tm.post_removed), "learned_filled": int(l_tm.post_filled)}
        clouds = {}
        for name, outs in (("plain", p_outs), ("learned", l_outs)):
            keep = []
            fv = fusion.fusion_views(sc, [o.depth for o in outs], [o.normal for o in outs], [o.cost for o in outs], src, src, keep)
            t = time.perf_counter()
            clouds[name] = fusion.fuse(be, fv, fusion.FusionParams(), int(sum((o.depth > 0).sum() for o in outs)) + 1)
            log(f"{name} fusion: {len(clouds[name].positions)} points in {time.perf_counter() - t:.1f} s")
        del p_outs, l_outs
        t = time.perf_counter()
        med, comp = fusion.compare(clouds["learned"], clouds["plain"], 0.05)
        log(f"clouds: median learned->plain distance {med * 100:.2f} cm, completeness at 5 cm {comp:.4f} "
            f"({time.perf_counter() - t:.1f} s)")
        line["clouds"] = {"plain_points": len(clouds["plain"].positions), "learned_points": len(clouds["learned"].positions),
                          "median_distance_m": round(med, 5), "completeness_5cm": round(comp, 5), "min_views": 3,
                          "post_processing": post is not None}
        del clouds
    s.close()

    # ---- one learned job from c3d_pm_begin to the downloaded read-out (the host-boundary rate). The
    # caller's result arrays exist before the timer and have been written once, as those of a caller
    # that keeps them between jobs: the timed region holds no host allocation or first-touch page
    # faults, only the library's work, its uploads and its downloads.
    dst = init_net.output_buffers(inp, views)
    for o, c, r in zip(*dst):
        for a in (o.depth, o.normal, o.cost, c, r):
            a.fill(0)
    device_sync()
    t = time.perf_counter()
    with be.begin(inp) as s2:  # the backend's default tile shape (1, 1): no tuning work inside the timed region
        init_net.densify_learned(s2, net, views, learned_s, download=True, out=dst)
    b2d_s = time.perf_counter() - t
    del dst
    log(f"learned begin -> downloaded read-out: {b2d_s * 1e3:.1f} ms ({job_mpx / b2d_s:.2f} Mpx/s)")
    line.update({
        "accuracy_vs_plain": {"depth_within_1pct": round(acc_l, 5), "gate_depth_99pct": acc_l >= gate,
                              "normal_within_2deg": round(nrm_l, 5), "valid_plain_pixels": npx,
                              "plain_reseeded_depth_within_1pct": round(acc_rs, 5),
                              "plain_reseeded_normal_within_2deg": round(nrm_rs, 5)},
        "plain": {"mpx_s": round(job_mpx / (ms_p / 1e3), 3), "ms_per_job": round(ms_p, 2), "steps": args.plain_steps,
                  "schedule": [args.photometric_iters, args.geometric_rounds, args.geometric_iters]},
        "speedup_vs_plain": round(ms_p / ms_l, 3),
        "random_init_same_schedule": {"mpx_s": round(job_mpx / (ms_t / 1e3), 3), "ms_per_job": round(ms_t, 2),
                                      "depth_within_1pct": round(acc_t, 5), "gate_depth_99pct": acc_t >= gate,
                                      "schedule": [twin_s.photometric_iters, twin_s.geometric_rounds, twin_s.geometric_iters]},
        "random_init_tuned": {"mpx_s": round(job_mpx / (ms_r / 1e3), 3), "ms_per_job": round(ms_r, 2),
                              "depth_within_1pct": round(acc_r, 5), "gate_depth_99pct": acc_r >= gate,
                              "schedule": [random_s.photometric_iters, random_s.geometric_rounds, random_s.geometric_iters]},
        "speedup_vs_random_tuned": round(ms_r / ms_l, 3),
        "begin_to_download": {"mpx_s": round(job_mpx / b2d_s, 3), "ms": round(b2d_s * 1e3, 1)},
    })

    net.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_object(inp, src, views, args.cpu_rows, args.photometric_iters, args.geometric_rounds,
                                                   args.geometric_iters, job_mpx)
    be.close()
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- plain workloads


def plain_main(args):
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
    log(f"scene {nviews} views, {job_mpx:.1f} Mpx rendered on the device in {time.perf_counter() - t0:.1f} s")
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
    shape = (1, 1)  # the engine's default shape
    if args.tile_autotune:
        # the half step's workgroup shape (rows, lanes per pixel), outside the timed region
        shape, shape_ms = sess.autotune(views, level=inp.num_levels - 1)  # the finest scale dominates
        log(f"tile autotuner: (rows, lanes per pixel) {shape}; two iterations " + shapes_text(shape_ms))
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
    launches = []
    for lv in range(inp.num_levels):
        n_ph = 2 * (sched.photometric_iters if lv == top else sched.photometric_iters_upper)
        launches += [(lv, False)] * n_ph + [(lv, True)] * (2 * sched.geometric_rounds * sched.geometric_iters)
    avg_bytes, flop = half_step_work(inp, src, launches)
    half_ms = sum(t.half_kernel_ms for t in tms) / sum(t.half_launches for t in tms)
    default_zone7 = (not batched and nviews == 7 and (args.width, args.height) == (1920, 1080) and args.photometric_iters == 7
                     and args.geometric_rounds == 1 and args.geometric_iters == 2 and tuple(args.window) == (5, 2))
    # the windows built with register-held weights; every other one runs the any-radius build
    reg_windows = {(5, 2), (4, 2), (3, 2), (2, 2), (3, 1), (2, 1), (1, 1)}
    kernel_name = (f"k_half<{inp.pm.window_radius},{inp.pm.window_step}>"
                   if (inp.pm.window_radius, inp.pm.window_step) in reg_windows else f"k_half<any radius,{inp.pm.window_step}>")
    roof, roof_valu = roofline_objects(kernel_name, avg_bytes, flop, half_ms, K_HALF_TRAFFIC.get(("zone7", tuple(shape))) if default_zone7 else None)
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
        "metric": METRIC,
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
        "data": "synthetic (seeded fBm terrain rendered on the device, sigma 0.01 noise, u8)",
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
        "roofline": roof,
        "roofline_valu": roof_valu,
        "stage_ms": {
            "photometric": round(sum(t.photometric_ms for t in tms) / len(tms), 2),
            "geometric": round(sum(t.geometric_ms for t in tms) / len(tms), 2),
            "filter": round(sum(t.filter_ms for t in tms) / len(tms), 2),
        },
    }
    sess.close()
    if not args.no_cpu and not fast960 and inp.num_levels == 1:
        line["cpu_baseline"] = cpu_baseline_object(inp, src, views, args.cpu_rows, args.photometric_iters, args.geometric_rounds,
                                                   args.geometric_iters, job_mpx)
    print(json.dumps(line), flush=True)


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["zone108", "zone7", "batched48", "fast960"], default="zone108",
                    help="zone108: 108 views 1920x1080, learned init + batched PatchStereo + geometric round (default); "
                         "zone7: 7 views 1920x1080, plain, 6 sources each; batched48: 48 views of uneven resolution and "
                         "source count in one batch, hipGraph; fast960: fast profile over 3 scales, 240 views 960x540")
    ap.add_argument("--window", type=int, nargs=2, default=[5, 2], metavar=("RADIUS", "STEP"),
                    help="zone7: NCC window radius and step (5 2 = fast profile, 5 1 = quality profile's window)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--views", type=int, default=None)
    ap.add_argument("--photometric-iters", type=int, default=7)
    ap.add_argument("--geometric-rounds", type=int, default=1)
    ap.add_argument("--geometric-iters", type=int, default=2)
    ap.add_argument("--graph", type=int, default=None, help="1 = replay the schedule as one hipGraph (default: batched48, fast960)")
    ap.add_argument("--tile-autotune", type=int, default=None,
                    help="pick the half step's workgroup shape with c3d_pm_autotune before timing (1) or keep the "
                         "engine's default (1, 1) (0); default 1, 0 for zone108 (the autotuner times iterations from a "
                         "random start, where the shapes tie, not the learned job's refinement, where one-row "
                         "workgroups are faster)")
    ap.add_argument("--plain-steps", type=int, default=2, help="zone108: timed jobs of the plain and random-init arms")
    ap.add_argument("--learned-only", action="store_true", help="zone108: time the learned arm only (profiling runs)")
    ap.add_argument("--learned-schedule", type=float, nargs=4, default=None, metavar=("REFINE", "GROUNDS", "GITERS", "CONF"),
                    help="zone108: fixed learned schedule instead of the autotuner's choice (profiling runs)")
    ap.add_argument("--fusion", type=int, default=1, help="zone108: fuse and compare the clouds (1) or not (0)")
    ap.add_argument("--post", type=int, default=1,
                    help="zone108: the fast profile's speckle removal and hole filling before fusion (1) or none (0)")
    ap.add_argument("--cpu-rows", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of this many MiB (profiling calibration)")
    args = ap.parse_args(argv)
    if args.tile_autotune is None:
        args.tile_autotune = 0 if args.workload == "zone108" else 1
    return args


def main():
    args = parse_args()
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
    else:
        plain_main(args)


if __name__ == "__main__":
    main()
