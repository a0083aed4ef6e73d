"""Trains the learned initialiser (include/c3d_init.h) on plain-PatchStereo depth maps of seeded
synthetic terrain, all on the GPU, and saves the fp32 parameters.

  python -m cumulus3d_amd.train_init --views 2400 --epochs 40 --out cumulus3d_amd/weights/init_unet.npy

Data: the zone workload's kind of scene. Formation flights over fBm terrain with random albedo:
per scene 4 stations x 3 cameras 45 deg down (25-35 m up, 6-10 m apart, random heading), rendered
at 1920 x 1080 on the device with Gaussian noise sigma 0.01 and quantised to u8, each view with its
4-8 nearest overlapping views as sources. The terrain seeds (20000 + scene) are disjoint from every
benchmark and test scene.

Labels: the engine's plain PatchStereo at 1920 x 1080 (7 photometric iterations from random
initialisation, one geometric round of 2, read-out; c3d_pm_densify), the method the learned path is
held to. They are reduced to the network's quarter resolution: a quarter-resolution pixel covers
a 4 x 4 block of full-resolution pixels (its centre maps onto the block's centre, as in the
device's plane initialisation), and it is labelled where all 16 pixels passed the read-out, with
the reciprocal of their mean inverse depth and their normalised mean normal.

Inputs: exactly what c3d_pm_densify_learned feeds the network. The quarter-resolution image is
the device pyramid of the view (two blur-and-halve steps, c3d_pm_build_pyramid). The plane-sweep
features are computed against the view's first two sources at the same resolution.
Training: Adam on the device (c3d_init_train_step), batches of 8 views, the learning rate divided
by 10 for the last 15 % of the epochs.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import abi, init_net, pm, scenes

FULL_W, FULL_H, FOCAL_FULL = 1920, 1080, 500.0
WIDTH, HEIGHT = FULL_W // 4, FULL_H // 4  # the network's input: the quarter-resolution pyramid level
NET_H = (HEIGHT + 7) // 8 * 8
DMIN, DMAX = 15.0, 120.0
STATIONS = 4
SEED0 = 20000
PLAIN = abi.Schedule(photometric_iters=7, geometric_rounds=1, geometric_iters=2)


def log(msg):
    print(f"[train_init] {msg}", file=sys.stderr, flush=True)


def scene(idx: int, device: bool = True):
    """Training scene `idx`: (DenseScene, sources per view)."""
    rng = np.random.default_rng(SEED0 + idx)
    cams = scenes.formation_rig(STATIONS, 3, spacing=float(rng.uniform(6.0, 10.0)), altitude=float(rng.uniform(25.0, 35.0)),
                                ground_dist=30.0, heading_deg=float(rng.uniform(0.0, 360.0)),
                                offset=(float(rng.uniform(-40.0, 40.0)), float(rng.uniform(-40.0, 40.0))))
    sc = scenes.make_mixed_scene(cams, [(FULL_W, FULL_H)] * len(cams), focal_1080=FOCAL_FULL, seed=SEED0 + idx, noise_sigma=0.01,
                                 noise_seed=SEED0 + 7919 * idx + 1, device=device)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(4, 9, len(cams))])
    return sc, src


def quarter_labels(depth: np.ndarray, normal: np.ndarray):
    """(H/4, W/4) labels of a full-resolution read-out: where all 16 pixels of a 4 x 4 block are
    valid, depth = 1 / mean(1 / d) and the normalised mean normal; 0 elsewhere."""
    h, w = depth.shape
    d = depth[: h // 4 * 4, : w // 4 * 4].reshape(h // 4, 4, w // 4, 4)
    n = normal[: h // 4 * 4, : w // 4 * 4].reshape(h // 4, 4, w // 4, 4, 3)
    valid = (d > 0).all(axis=(1, 3))
    inv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0).mean(axis=(1, 3))
    lab_d = np.where(valid, 1.0 / np.where(valid, inv, 1.0), 0.0).astype(np.float32)
    nm = n.mean(axis=(1, 3))
    nn = np.linalg.norm(nm, axis=-1, keepdims=True)
    lab_n = np.where(valid[..., None], nm / np.maximum(nn, 1e-12), 0.0).astype(np.float32)
    return lab_d, lab_n


def pad(a, h):
    out = np.zeros((h,) + a.shape[1:], a.dtype)
    out[: a.shape[0]] = a
    return out


class Sample:
    """One training view: quarter-resolution images and intrinsics of the view and its first two
    sources, the reference-to-source poses, and the quarter-resolution labels."""

    def __init__(self, gray, k, sources, depth, normal):
        self.gray, self.k, self.sources, self.depth, self.normal = gray, k, sources, depth, normal

    def sweep(self, keep):
        return init_net.sweep_view(self.gray, self.k, self.sources, DMIN, DMAX, keep)


def build_dataset(be, nviews, first=0):
    """Renders and labels scenes first, first + 1, ... until `nviews` views are collected."""
    data = []
    t0 = time.time()
    idx = first
    pmp, fl = abi.fast_profile()
    while len(data) < nviews:
        sc, src = scene(idx)
        inp = scenes.kernel_input(sc, src, [(DMIN, DMAX)] * len(sc.views), pmp, fl)
        with be.begin(inp) as s:
            outs, _, _, _ = s.densify(list(range(len(sc.views))), PLAIN)
        quarter = []
        for v in sc.views:
            lv = be.build_pyramid(v.gray, v.k, 3)[0]
            quarter.append((lv.gray, lv.k))
        for v in range(len(sc.views)):
            srcs = []
            for j in src[v][:2]:
                r, t, _ = scenes.pair_geometry(sc, v, j)
                srcs.append((quarter[j][0], quarter[j][1], r, t))
            lab_d, lab_n = quarter_labels(outs[v].depth, outs[v].normal)
            data.append(Sample(quarter[v][0], quarter[v][1], srcs, pad(lab_d, NET_H), pad(lab_n, NET_H)))
        idx += 1
        if idx % 10 == 0:
            log(f"scenes {idx - first} rendered and labelled, {len(data)} views ({time.time() - t0:.0f} s)")
    return data[:nviews]


def train(be, data, epochs, batch, lr, seed=1, log_every=100):
    net = init_net.InitNet(be, batch, NET_H, WIDTH, seed=seed)
    tp = init_net.TrainParams(lr=lr)
    rng = np.random.default_rng(seed)
    step = 0
    hist = []
    decay_at = int(round(0.85 * epochs)) if epochs > 1 else epochs
    for ep in range(epochs):
        order = rng.permutation(len(data))
        tp.lr = lr * (0.1 if ep >= decay_at else 1.0)
        for b0 in range(0, len(order) - batch + 1, batch):
            keep = []
            idx = order[b0:b0 + batch]
            net.features(0, [data[i].sweep(keep) for i in idx])
            for j, i in enumerate(idx):
                net.set_labels(j, data[i].depth, data[i].normal, WIDTH, HEIGHT, DMIN, DMAX)
            losses, _ = net.train_step(batch, tp)
            tot = (losses[0] + tp.w_normal * losses[1] + tp.w_conf * losses[2]) / max(losses[3], 1.0)
            hist.append(float(tot))
            step += 1
            if step % log_every == 0:
                log(f"epoch {ep + 1}/{epochs} step {step}: loss {np.mean(hist[-log_every:]):.4f}")
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=2400)
    ap.add_argument("--epochs", type=int, default=40)
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
