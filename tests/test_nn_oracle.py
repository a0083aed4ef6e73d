"""Pins of the NumPy restatement of the learned initialiser (oracle/nn_oracle.py), CPU only:

* the analytic backward pass against central finite differences of the float64 loss, for
  parameters of every layer (a dropped term, a transposed weight or a wrong skip connection fails);
* one Adam step from zero moments is a signed step of size lr (closed form);
* the plane-sweep cost is exactly 0 at every depth for a source identical to the reference, and at
  the true depth for a source shifted by an exact integer disparity.
"""
import numpy as np

from oracle import nn_oracle as nno

WN, WC = 0.5, 0.1


def _problem(seed=0, h=8, w=8):
    rng = np.random.default_rng(seed)
    lay, n = nno.layout()
    p = np.zeros(n)
    for (name, woff, boff, cin, cout, k, relu), (_, _, _, _, cin_real) in zip(lay, nno.SPEC):
        sd = np.sqrt(2.0 / (k * k * cin_real))
        p[woff:woff + cout * k * k * cin] = rng.normal(0, sd, cout * k * k * cin)
        p[boff:boff + cout] = rng.normal(0, 0.05, cout)
    feat = np.zeros((1, h, w, 64))
    feat[..., :33] = rng.uniform(0, 1, (1, h, w, 33))
    label = np.zeros((1, h, w, 4))
    label[..., 0] = rng.uniform(20, 60, (1, h, w))
    nrm = rng.normal(size=(1, h, w, 3))
    nrm[..., 2] = -np.abs(nrm[..., 2]) - 1
    label[..., 1:] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    label[0, 0, 0, 0] = 0.0  # an unlabelled pixel
    rng_ = np.array([[15.0, 120.0]])
    return p, feat, label, rng_


def _loss(p, feat, label, rng_):
    net = nno.Net(p, bf=False)
    A = net.forward(feat)
    sums, g = net.loss(A["out"], label, rng_, (8, 8), WN, WC, 0.01)
    return nno.total_loss(sums, WN, WC), net, A, g


def test_backward_matches_finite_differences():
    p, feat, label, rng_ = _problem()
    L0, net, A, g = _loss(p, feat, label, rng_)
    grad = net.backward(A, g)
    lay, _ = nno.layout()
    rng = np.random.default_rng(1)
    checked = 0
    for (name, woff, boff, cin, cout, k, relu) in lay:
        for idx in list(rng.integers(woff, woff + cout * k * k * cin, 3)) + [boff + int(rng.integers(0, min(cout, 5)))]:
            if name == "head" and idx < boff and (idx - woff) // (k * k * cin) >= 5:
                continue
            h = 1e-5
            pp, pm_ = p.copy(), p.copy()
            pp[idx] += h
            pm_[idx] -= h
            fd = (_loss(pp, feat, label, rng_)[0] - _loss(pm_, feat, label, rng_)[0]) / (2 * h)
            assert abs(fd - grad[idx]) <= 1e-5 + 1e-4 * abs(fd), (name, idx, fd, grad[idx])
            checked += 1
    assert checked >= 40
    # the unused head outputs get no gradient
    name, woff, boff, cin, cout, k, relu = lay[-1]
    assert np.all(grad[woff + 5 * cin:boff] == 0) and np.all(grad[boff + 5:boff + cout] == 0)


def test_adam_first_step_closed_form():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 0.0]) * 10
    w1, m, v = nno.adam(w, g, np.zeros(3), np.zeros(3), 1, 0.01, 0.9, 0.999, 1e-12, 10.0)
    assert np.allclose(w1, w - 0.01 * np.sign(g), atol=1e-9)


def test_sweep_closed_forms():
    rng = np.random.default_rng(2)
    ref = rng.integers(0, 256, (20, 40)).astype(np.uint8)
    k = [50.0, 50.0, 19.5, 9.5]
    eye = np.eye(3).reshape(9)
    f = nno.sweep(ref, k, [(ref, k, eye, [0.0, 0.0, 0.0])], 10.0, 100.0, 24, 40)
    assert np.all(f[1:19, 1:39, :32] < 1e-9) and np.allclose(f[:20, :, 32], ref / 255.0) and np.all(f[20:] == 0)
    # source = reference shifted left by 4 px: x_src = x - fx b / d; b is chosen so that the
    # disparity is exactly 4 px at the centre depth of one bin
    src = np.zeros_like(ref)
    src[:, :-4] = ref[:, 4:]
    ia, ib = 1 / 100.0, 1 / 10.0
    kk = int(round((1 / 12.5 - ia) / (ib - ia) * 32 - 0.5))
    d = 1.0 / (ia + (kk + 0.5) / 32.0 * (ib - ia))
    # translation making the disparity at depth d exactly 4 px: fx * b / d = 4
    b = 4 * d / 50.0
    f = nno.sweep(ref, k, [(src, k, eye, [-b, 0.0, 0.0])], 10.0, 100.0, 20, 40)
    # (row 0 and column 4 project onto the image border, where rounding may fall just outside)
    assert np.all(f[1:, 5:, kk] < 1e-9)
    assert f[1:, 5:, (kk + 8) % 32].mean() > 0.1


def _torch_net(p, feat):
    """The U-Net written independently with torch.nn.functional (NCHW, float64): 3x3 convolutions
    with zero padding, 2x2 average pooling, nearest upsampling, the upsampled decoder input first
    and the encoder skip second, a 1x1 head. Parameters in the flat layout of include/c3d_init.h:
    weights [cout][k*k][cin] (tap t = (dy + r) k + (dx + r)), then bias, layer by layer."""
    import torch
    import torch.nn.functional as F

    lay, _ = nno.layout()
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)

    def conv(i, x, relu=True):
        name, woff, boff, cin, cout, k, r = lay[i]
        w = pt[woff:woff + cout * k * k * cin].reshape(cout, k, k, cin).permute(0, 3, 1, 2)
        y = F.conv2d(x, w, pt[boff:boff + cout], padding=k // 2)
        return F.relu(y) if r else y

    x = torch.tensor(feat, dtype=torch.float64).permute(0, 3, 1, 2)
    e0 = conv(1, conv(0, x))
    e1 = conv(3, conv(2, F.avg_pool2d(e0, 2)))
    e2 = conv(5, conv(4, F.avg_pool2d(e1, 2)))
    e3 = conv(7, conv(6, F.avg_pool2d(e2, 2)))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")  # noqa: E731
    d2 = conv(8, torch.cat([up(e3), e2], 1))
    d1 = conv(9, torch.cat([up(d2), e1], 1))
    d0 = conv(10, torch.cat([up(d1), e0], 1))
    return conv(11, d0).permute(0, 2, 3, 1), pt


def test_forward_matches_torch_functional():
    """The NumPy forward pass (float64) equals the same network in torch.nn.functional: pins the
    weight layout, the padding, the pooling, the upsampling and the order of the skip
    concatenations, which the finite-difference test (forward against itself) cannot."""
    p, feat, _, _ = _problem(seed=3, h=16, w=16)
    A = nno.Net(p, bf=False).forward(feat)
    out, _ = _torch_net(p, feat)
    assert np.allclose(A["out"], out.detach().numpy(), rtol=1e-9, atol=1e-10)
    # the comparison can fail: layer d1 reading its input channels in another order (its skip
    # concatenated first) is told apart
    q = p.copy()
    lay, _ = nno.layout()
    name, woff, boff, cin, cout, k, r = lay[9]
    w = q[woff:woff + cout * k * k * cin].reshape(cout, k * k, cin)
    w[:] = np.concatenate([w[:, :, 128:], w[:, :, :128]], axis=2)
    assert not np.allclose(nno.Net(q, bf=False).forward(feat)["out"], out.detach().numpy(), rtol=1e-6)


def test_backward_matches_torch_autograd():
    """The analytic parameter gradient equals torch autograd through the torch network and a torch
    statement of the loss (L1 on range-normalised inverse depth, 1 - cosine of the normal, binary
    cross-entropy of the confidence against |1/idp - d| <= tol d, over labelled valid pixels)."""
    import torch

    p, feat, label, rng_ = _problem(seed=4)
    _, net, A, g = _loss(p, feat, label, rng_)
    grad = net.backward(A, g)
    out, pt = _torch_net(p, feat)
    lab = torch.tensor(label, dtype=torch.float64)
    ok = lab[..., 0] > 0
    ia, ib = 1.0 / rng_[0, 1], 1.0 / rng_[0, 0]
    idp = ia + torch.sigmoid(out[..., 0]) * (ib - ia)
    d_lab = torch.where(ok, lab[..., 0], torch.ones_like(lab[..., 0]))
    ld = (idp - 1.0 / d_lab).abs() / (ib - ia)
    v = out[..., 1:4]
    nn_ = v / torch.sqrt((v * v).sum(-1, keepdim=True) + 1e-6)
    ln = 1.0 - (nn_ * lab[..., 1:4]).sum(-1)
    c = torch.sigmoid(out[..., 4])
    tgt = ((1.0 / idp - lab[..., 0]).abs() <= 0.01 * lab[..., 0]).double()
    lc = -(tgt * torch.log(c.clamp_min(1e-7)) + (1 - tgt) * torch.log((1 - c).clamp_min(1e-7)))
    loss = ld[ok].sum() + WN * ln[ok].sum() + WC * lc[ok].sum()
    loss.backward()
    tg = pt.grad.numpy()
    assert np.allclose(grad, tg, rtol=1e-7, atol=1e-10), float(np.abs(grad - tg).max())
