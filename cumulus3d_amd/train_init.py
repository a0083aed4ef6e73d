"""Trains the learned initialiser (include/c3d_init.h) on plain-PatchStereo depth maps of seeded
synthetic terrain, all on the GPU, and saves the fp32 parameters.

  python -m cumulus3d_amd.train_init --views 2000 --out cumulus3d_amd/weights/init_unet.npy

Data: scenes of 3 oblique cameras (45 deg down, 25-35 m up, 4-8 m baseline, random heading) over
fBm terrain with random albedo, rendered at 480 x 270 (the quarter resolution of 1920 x 1080 at
f = 500 px, so f = 125 px) with Gaussian noise sigma 0.01. Labels: the engine's plain PatchStereo
(fast profile, 3 scales, geometric rounds and read-out filter, c3d_pm_densify) on the same views,
every view using the other two as sources; only pixels that pass the read-out filter are labelled.
Inputs: the plane-sweep features of each view against its two sources. Training: Adam on the
device (c3d_init_train_step), batches of 8 views.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

from . import abi, init_net, pm, scenes

WIDTH, HEIGHT, FOCAL = 480, 270, 125.0
NET_H = (HEIGHT + 7) // 8 * 8
DMIN, DMAX = 15.0, 120.0


def log(msg):
    print(f"[train_init] {msg}", file=sys.stderr, flush=True)


def scene(idx: int, width=WIDTH, height=HEIGHT, focal=FOCAL):
    rng = np.random.default_rng(1000 + idx)
    rig = scenes.oblique_rig(3, baseline=float(rng.uniform(4.0, 8.0)), altitude=float(rng.uniform(25.0, 35.0)),
                             ground_dist=30.0, yaw_deg=float(rng.uniform(0.0, 360.0)))
    return scenes.make_terrain_scene(rig, width, height, focal=focal, seed=7000 + idx, noise_sigma=0.01, noise_seed=9000 + idx)


def sweep_sources(sc, v):
    out = []
    for s in range(len(sc.views)):
        if s != v:
            r, t, _ = scenes.pair_geometry(sc, v, s)
            out.append((sc.views[s].gray, sc.views[s].k, r, t))
    return out


def label_scenes(be, scs):
    """Plain PatchStereo (fast profile, every scale) on a batch of 3-view scenes in one session."""
    n = 3 * len(scs)
    views, src = [], []
    for i, sc in enumerate(scs):
        for v in range(3):
            views.append(sc.views[v])
            src.append([3 * i + o for o in range(3) if o != v])
    big = scenes.DenseScene(views, [], None, None)
    h, w = views[0].gray.shape
    nl = pm.num_levels(min(w, h), 3, 64)
    levels = [be.build_pyramid(v.gray, v.k, nl) for v in views]
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(big, src, [(DMIN, DMAX)] * n, pmp, fl, levels=levels)
    with be.begin(inp) as s:
        outs, counts, raws, _ = s.densify(list(range(n)), abi.Schedule.fast())
    return outs


def pad(a, h):
    out = np.zeros((h,) + a.shape[1:], a.dtype)
    out[: a.shape[0]] = a
    return out


def build_dataset(be, nviews, first=0):
    data = []
    t0 = time.time()
    nsc = math.ceil(nviews / 3)
    for c0 in range(0, nsc, 16):
        scs = [scene(first + i) for i in range(c0, min(nsc, c0 + 16))]
        outs = label_scenes(be, scs)
        for i, sc in enumerate(scs):
            sc.gt_depth = sc.gt_normal = None  # only the images and cameras are kept
            for v in range(3):
                o = outs[3 * i + v]
                data.append((sc, v, pad(o.depth, NET_H), pad(o.normal, NET_H)))
        log(f"scenes {min(nsc, c0 + 16)}/{nsc} rendered and labelled ({time.time() - t0:.0f} s)")
    return data[:nviews]


def train(be, data, epochs, batch, lr, seed=1, log_every=50):
    net = init_net.InitNet(be, batch, NET_H, WIDTH, seed=seed)
    tp = init_net.TrainParams(lr=lr)
    rng = np.random.default_rng(seed)
    step = 0
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        for b0 in range(0, len(order) - batch + 1, batch):
            keep = []
            idx = order[b0:b0 + batch]
            views = [init_net.sweep_view(data[i][0].views[data[i][1]].gray, data[i][0].views[data[i][1]].k,
                                         sweep_sources(data[i][0], data[i][1]), DMIN, DMAX, keep) for i in idx]
            net.features(0, views)
            for j, i in enumerate(idx):
                net.set_labels(j, data[i][2], data[i][3], WIDTH, HEIGHT, DMIN, DMAX)
            lr_t = lr * (0.1 if ep == epochs - 1 and epochs > 1 else 1.0)
            tp.lr = lr_t
            losses, _ = net.train_step(batch, tp)
            tot = (losses[0] + tp.w_normal * losses[1] + tp.w_conf * losses[2]) / max(losses[3], 1.0)
            hist.append(float(tot))
            step += 1
            if step % log_every == 0:
                log(f"epoch {ep + 1}/{epochs} step {step}: loss {np.mean(hist[-log_every:]):.4f}")
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "init_unet.npy"))
    args = ap.parse_args()
    if not pm.is_available():
        print("no gfx950 device: training runs on the GPU only", file=sys.stderr)
        sys.exit(1)
    be = pm.HipPatchMatch()
    data = build_dataset(be, args.views)
    net, hist = train(be, data, args.epochs, args.batch, args.lr)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, net.get_params())
    log(f"saved {net.n} parameters to {args.out}; loss {np.mean(hist[:20]):.4f} -> {np.mean(hist[-20:]):.4f}")


if __name__ == "__main__":
    main()
