"""NumPy restatement of the learned initialiser's U-Net (forward, loss, backward, Adam) —
TEST INFRASTRUCTURE ONLY.

The checker for cumulus3d_amd/csrc/nn_kernels.hip + nn_net.cpp (layout and hyper-parameters in
include/c3d_init.h). No reference counterpart exists (the reference initialises PatchStereo at
random, crates/cuda/src/kernels/patchmatch.cu:661-692). With ``bf16=True`` the activations,
weights and the pre-activation gradients are rounded to bfloat16 (round to nearest even) at the
same points as the device; with ``bf16=False`` everything is float64, which the finite-difference
test uses to pin the analytic backward pass.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, k, relu, cin_real) in layer order; parameters: weights [cout][k*k][cin], then bias
SPEC = [(64, 32, 3, 1, 33), (32, 32, 3, 1, 32), (32, 64, 3, 1, 32), (64, 64, 3, 1, 64), (64, 128, 3, 1, 64),
        (128, 128, 3, 1, 128), (128, 128, 3, 1, 128), (128, 128, 3, 1, 128), (256, 128, 3, 1, 256), (192, 64, 3, 1, 192),
        (96, 32, 3, 1, 96), (32, 32, 1, 0, 32)]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2", "d1", "d0", "head"]


def layout():
    """[(name, woff, boff, cin, cout, k, relu)] and the parameter count."""
    out, off = [], 0
    for name, (cin, cout, k, relu, _) in zip(NAMES, SPEC):
        woff = off
        off += cout * k * k * cin
        boff = off
        off += cout
        out.append((name, woff, boff, cin, cout, k, relu))
    return out, off


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def bf16_bits(x):
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class Net:
    def __init__(self, params: np.ndarray, bf: bool = True):
        self.lay, n = layout()
        assert params.size == n
        self.p = np.asarray(params, np.float64)
        self.bf = bf

    def r(self, x):
        return bf16(x) if self.bf else np.asarray(x, np.float64)

    def weights(self, i):
        name, woff, boff, cin, cout, k, relu = self.lay[i]
        w = self.p[woff:woff + cout * k * k * cin].reshape(cout, k * k, cin)
        return self.r(w), self.p[boff:boff + cout]

    # ---------------------------------------------------------------- layers
    @staticmethod
    def _shift(x, dy, dx):
        """x[b, y + dy, x + dx, :] with zeros outside."""
        n, h, w, c = x.shape
        out = np.zeros_like(x)
        ys, ye = max(0, -dy), min(h, h - dy)
        xs, xe = max(0, -dx), min(w, w - dx)
        if ys < ye and xs < xe:
            out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
        return out

    def conv(self, i, x):
        name, woff, boff, cin, cout, k, relu = self.lay[i]
        w, b = self.weights(i)
        r = k // 2
        y = np.zeros(x.shape[:3] + (cout,))
        for t in range(k * k):
            dy, dx = t // k - r, t % k - r
            y += self._shift(x, dy, dx) @ w[:, t, :].T
        return y + b

    def conv_bwd(self, i, x, dz, need_dx=True):
        name, woff, boff, cin, cout, k, relu = self.lay[i]
        w, _ = self.weights(i)
        r = k // 2
        dw = np.zeros((cout, k * k, cin))
        dx = np.zeros(x.shape[:3] + (cin,)) if need_dx else None
        for t in range(k * k):
            dy, ddx = t // k - r, t % k - r
            xs = self._shift(x, dy, ddx)
            dw[:, t, :] = np.einsum("nhwo,nhwi->oi", dz, xs)
            if need_dx:
                dx += self._shift(dz @ w[:, t, :], -dy, -ddx)
        db = dz.sum(axis=(0, 1, 2))
        return dw, db, dx

    def act(self, i, x):
        z = self.conv(i, x)
        relu = self.lay[i][6]
        return self.r(np.maximum(z, 0.0)) if relu else z

    def pool(self, x):
        n, h, w, c = x.shape
        s = x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2]
        if self.bf:
            # the device sums the four bf16 values in fp32
            s = (x[:, 0::2, 0::2].astype(np.float32) + x[:, 0::2, 1::2].astype(np.float32) + x[:, 1::2, 0::2].astype(np.float32)
                 + x[:, 1::2, 1::2].astype(np.float32))
        return self.r(0.25 * s)

    @staticmethod
    def up(x):
        return x.repeat(2, axis=1).repeat(2, axis=2)

    @staticmethod
    def up_bwd(d):
        return d[:, 0::2, 0::2] + d[:, 0::2, 1::2] + d[:, 1::2, 0::2] + d[:, 1::2, 1::2]

    @staticmethod
    def pool_bwd(d):
        return 0.25 * d.repeat(2, axis=1).repeat(2, axis=2)

    # ---------------------------------------------------------------- network
    def forward(self, feat):
        A = {}
        A["feat"] = self.r(feat)
        A["e0a"] = self.act(0, A["feat"])
        A["e0b"] = self.act(1, A["e0a"])
        A["p1"] = self.pool(A["e0b"])
        A["e1a"] = self.act(2, A["p1"])
        A["e1b"] = self.act(3, A["e1a"])
        A["p2"] = self.pool(A["e1b"])
        A["e2a"] = self.act(4, A["p2"])
        A["e2b"] = self.act(5, A["e2a"])
        A["p3"] = self.pool(A["e2b"])
        A["e3a"] = self.act(6, A["p3"])
        A["e3b"] = self.act(7, A["e3a"])
        A["cat2"] = np.concatenate([self.up(A["e3b"]), A["e2b"]], axis=3)
        A["d2"] = self.act(8, A["cat2"])
        A["cat1"] = np.concatenate([self.up(A["d2"]), A["e1b"]], axis=3)
        A["d1"] = self.act(9, A["cat1"])
        A["cat0"] = np.concatenate([self.up(A["d1"]), A["e0b"]], axis=3)
        A["d0"] = self.act(10, A["cat0"])
        A["out"] = self.conv(11, A["d0"])
        return A

    def loss(self, out, label, rng, valid, w_normal, w_conf, conf_tol):
        """Loss sums (depth, normal, conf, count) and d loss / d out (channels 0..4)."""
        n, h, w, _ = out.shape
        vw, vh = valid
        o = out[..., :5].astype(np.float64)
        d_lab = label[..., 0].astype(np.float64)
        ok = np.zeros((n, h, w), bool)
        ok[:, :vh, :vw] = True
        ok &= d_lab > 0
        ia = (1.0 / rng[:, 1])[:, None, None]
        ib = (1.0 / rng[:, 0])[:, None, None]
        s0 = 1.0 / (1.0 + np.exp(-o[..., 0]))
        idp = ia + s0 * (ib - ia)
        idt = np.where(ok, 1.0 / np.where(ok, d_lab, 1.0), 0.0)
        e = idp - idt
        ld = np.abs(e) / (ib - ia)
        g = np.zeros((n, h, w, 32))
        g[..., 0] = np.sign(e) * s0 * (1 - s0)
        v = o[..., 1:4]
        nv = np.sqrt((v * v).sum(-1) + 1e-6)
        nn = v / nv[..., None]
        lab_n = label[..., 1:4].astype(np.float64)
        dot = (nn * lab_n).sum(-1)
        ln = 1.0 - dot
        g[..., 1:4] = -w_normal * (lab_n - dot[..., None] * nn) / nv[..., None]
        c = 1.0 / (1.0 + np.exp(-o[..., 4]))
        tgt = (np.abs(1.0 / idp - d_lab) <= conf_tol * d_lab).astype(np.float64)
        lc = -(tgt * np.log(np.maximum(c, 1e-7)) + (1 - tgt) * np.log(np.maximum(1 - c, 1e-7)))
        g[..., 4] = w_conf * (c - tgt)
        g[~ok] = 0.0
        sums = np.array([ld[ok].sum(), ln[ok].sum(), lc[ok].sum(), ok.sum()], np.float64)
        return sums, g

    def backward(self, A, dout):
        """Parameter gradient (flat, layout()) of the loss whose output gradient is dout."""
        g = np.zeros_like(self.p)

        def put(i, dw, db):
            name, woff, boff, cin, cout, k, relu = self.lay[i]
            g[woff:woff + dw.size] = dw.reshape(-1)
            g[boff:boff + cout] = db

        def relu_dz(a, da):
            return self.r(np.where(a > 0, da, 0.0))

        dz = self.r(dout)
        dw, db, dd0 = self.conv_bwd(11, A["d0"], dz)
        put(11, dw, db)
        dw, db, dcat0 = self.conv_bwd(10, A["cat0"], relu_dz(A["d0"], dd0))
        put(10, dw, db)
        dd1 = self.up_bwd(dcat0[..., :64])
        de0b = dcat0[..., 64:96].copy()
        dw, db, dcat1 = self.conv_bwd(9, A["cat1"], relu_dz(A["d1"], dd1))
        put(9, dw, db)
        dd2 = self.up_bwd(dcat1[..., :128])
        de1b = dcat1[..., 128:192].copy()
        dw, db, dcat2 = self.conv_bwd(8, A["cat2"], relu_dz(A["d2"], dd2))
        put(8, dw, db)
        de3b = self.up_bwd(dcat2[..., :128])
        de2b = dcat2[..., 128:256].copy()
        dw, db, de3a = self.conv_bwd(7, A["e3a"], relu_dz(A["e3b"], de3b))
        put(7, dw, db)
        dw, db, dp3 = self.conv_bwd(6, A["p3"], relu_dz(A["e3a"], de3a))
        put(6, dw, db)
        de2b = de2b + self.pool_bwd(dp3)
        dw, db, de2a = self.conv_bwd(5, A["e2a"], relu_dz(A["e2b"], de2b))
        put(5, dw, db)
        dw, db, dp2 = self.conv_bwd(4, A["p2"], relu_dz(A["e2a"], de2a))
        put(4, dw, db)
        de1b = de1b + self.pool_bwd(dp2)
        dw, db, de1a = self.conv_bwd(3, A["e1a"], relu_dz(A["e1b"], de1b))
        put(3, dw, db)
        dw, db, dp1 = self.conv_bwd(2, A["p1"], relu_dz(A["e1a"], de1a))
        put(2, dw, db)
        de0b = de0b + self.pool_bwd(dp1)
        dw, db, de0a = self.conv_bwd(1, A["e0a"], relu_dz(A["e0b"], de0b))
        put(1, dw, db)
        dw, db, _ = self.conv_bwd(0, A["feat"], relu_dz(A["e0a"], de0a), need_dx=False)
        put(0, dw, db)
        return g


def total_loss(sums, w_normal, w_conf):
    return sums[0] + w_normal * sums[1] + w_conf * sums[2]


def adam(w, g, m, v, t, lr, b1, b2, eps, count):
    """One Adam step (in place on copies); the gradient is divided by max(1, count)."""
    g = g / max(1.0, count)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


def sweep(ref, rk, srcs, dmin, dmax, H, W):
    """Plane-sweep features [H][W][64] (float64; the device stores bf16). srcs: list of
    (gray, k, r[9], t[3])."""
    rh, rw = ref.shape
    f = np.zeros((H, W, 64))
    ys, xs = np.mgrid[0:rh, 0:rw].astype(np.float64)
    rx, ry = (xs - rk[2]) / rk[0], (ys - rk[3]) / rk[1]
    ia, ib = 1.0 / dmax, 1.0 / dmin
    ir = ref.astype(np.float64)
    for k in range(32):
        d = 1.0 / (ia + (k + 0.5) / 32.0 * (ib - ia))
        X, Y, Z = rx * d, ry * d, np.full_like(rx, d)
        cost = np.zeros_like(rx)
        cnt = np.zeros_like(rx)
        for (g, sk, r, t) in srcs:
            r = np.asarray(r, np.float64).reshape(3, 3)
            xs_ = r[0, 0] * X + r[0, 1] * Y + r[0, 2] * Z + t[0]
            ys_ = r[1, 0] * X + r[1, 1] * Y + r[1, 2] * Z + t[1]
            zs_ = r[2, 0] * X + r[2, 1] * Y + r[2, 2] * Z + t[2]
            front = zs_ > 1e-6
            zz = np.where(front, zs_, 1.0)
            u = sk[0] * xs_ / zz + sk[2]
            v = sk[1] * ys_ / zz + sk[3]
            sh, sw = g.shape
            inside = front & (u >= 0) & (v >= 0) & (u <= sw - 1) & (v <= sh - 1)
            uc, vc = np.clip(u, 0, sw - 1), np.clip(v, 0, sh - 1)
            x0, y0 = np.floor(uc).astype(int), np.floor(vc).astype(int)
            x1, y1 = np.minimum(x0 + 1, sw - 1), np.minimum(y0 + 1, sh - 1)
            ax, ay = uc - x0, vc - y0
            gg = g.astype(np.float64)
            top = gg[y0, x0] + ax * (gg[y0, x1] - gg[y0, x0])
            bot = gg[y1, x0] + ax * (gg[y1, x1] - gg[y1, x0])
            val = top + ay * (bot - top)
            cost += np.where(inside, np.abs(ir - val), 0.0)
            cnt += inside
        f[:rh, :rw, k] = np.where(cnt > 0, cost / (255.0 * np.maximum(cnt, 1)), 1.0)
    f[:rh, :rw, 32] = ir / 255.0
    return f
