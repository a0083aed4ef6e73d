"""Pins the CPU oracle (oracle/pm_oracle.c) independently of its own formulas.

* integer hashing and RNG: published FNV-1a-64 and splitmix64 vectors;
* math twins: the reference's own unit-test values (crates/dense/src/math.rs:164-212,
  image.rs:316-326, upsample.rs:266-276, neighbors.rs:193-198);
* elementary functions: libm within 2 ulp;
* geometry: the homography maps a pixel to the brute-force projection of its 3-D plane point;
* NCC cost: closed forms (identical textures under the true plane -> 0, constant patch -> 2,
  projection outside -> 2) and FP32-vs-FP64 agreement to 1e-4;
* each check is chosen so that a plausible slip (sign, index, transposed operand, dropped term)
  fails it; tests/test_oracle_mutations.py plants such slips to confirm.
"""
import ctypes as C
import math

import numpy as np
import pytest

from cumulus3d_amd import abi, scenes

M64 = (1 << 64) - 1


def test_fnv1a_published_vectors(oracle_lib):
    L = oracle_lib.lib()
    assert L.pmo_hash_bytes(b"", 0) == 0xCBF29CE484222325
    assert L.pmo_hash_bytes(b"a", 1) == 0xAF63DC4C8601EC8C
    assert L.pmo_hash_bytes(b"foobar", 6) == 0x85944171F73967E8
    assert scenes.hash_bytes(b"foobar") == 0x85944171F73967E8


def test_splitmix64_published_vectors(oracle_lib):
    L = oracle_lib.lib()
    # splitmix64 next() from state 0 and from 1234567 (Vigna's reference generator)
    for seed, want in [(0, [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]),
                       (1234567, [6457827717110365317, 3203168211198807973, 9817491932198370423])]:
        s = seed
        for w in want:
            s = (s + 0x9E3779B97F4A7C15) & M64
            assert L.pmo_mix64(s) == w
        st = C.c_uint64(seed)
        for w in want:
            f = L.pmo_rng_next(C.byref(st))
            assert f == np.float32(((w >> 40) + 1) / 16777216.0)


def test_rng_state_definition(oracle_lib):
    L = oracle_lib.lib()
    mix = scenes.mix64
    seed, key = 0x005EED0FDE75, scenes.hash_bytes(b"synth_0_0.png")
    for (lvl, run, t, half, pix) in [(0, 1, 0xFFFF, 0, 0), (2, 7, 3, 1, 123456), (1, 0, 0, 0, 2**40 + 5)]:
        tag = (lvl << 48) ^ (run << 32) ^ (t << 16) ^ half
        want = mix(seed ^ mix(key ^ mix(tag)) ^ pix)
        assert L.pmo_rng_state(seed, key, lvl, run, t, half, pix) == want


def rel(a, b):
    return abs((a - b) / b)


def test_math_reference_values(oracle_lib):
    """crates/dense/src/math.rs:164-197."""
    L = oracle_lib.lib()
    assert rel(L.pmo_emission_norm(0.6), 1.330950) < 1e-5
    assert rel(L.pmo_emission(0.0, 0.6), 1.330950) < 1e-5
    assert rel(L.pmo_emission(0.9, 0.6), 0.432096) < 1e-5
    assert rel(L.pmo_emission(1.0, 0.6), 0.331875) < 1e-5
    assert rel(L.pmo_emission(2.0, 0.6), 0.005145) < 1e-4
    assert abs(L.pmo_erf(0.5) - 0.5204998778130465) < 1e-12
    assert abs(L.pmo_erf(2.0) - 0.9953222650189527) < 1e-12
    assert abs(L.pmo_erf(-1.0) + 0.8427007929497149) < 1e-12
    a1 = math.radians(1.0)
    assert L.pmo_triangulation_prior(1.0, a1) == 0.0
    assert abs(L.pmo_triangulation_prior(math.cos(math.radians(0.5)), a1) - 0.437507) < 1e-5
    assert L.pmo_triangulation_prior(math.cos(math.radians(2.0)), a1) == 1.0
    assert abs(L.pmo_incident_prior(1.0, 0.9) - 1.0) < 1e-12
    assert abs(L.pmo_incident_prior(-0.3, 0.9) - 0.539408) < 1e-6
    ident = (C.c_double * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    assert abs(L.pmo_resolution_prior(ident, 40.0, 30.0, 5.0) - 100.0 / 121.0) < 1e-9
    v = (C.c_double * 4)(1.0, 2.0, 3.0, 4.0)
    assert abs(L.pmo_quantile_sorted(v, 4, 0.75) - 3.25) < 1e-12


def test_plane_transfer_reference(oracle_lib):
    """math.rs:199-212: ray-plane intersection."""
    L = oracle_lib.lib()
    l = math.hypot(0.2, 1.0)
    n = (C.c_double * 3)(-0.2 / l, 0.0, -1.0 / l)
    hit = lambda r: 10.0 / (r[2] + 0.2 * r[0])
    rq, rp = [0.1, -0.05, 1.0], [-0.07, 0.12, 1.0]
    d = C.c_double()
    assert L.pmo_plane_transfer(hit(rq), n, (C.c_double * 3)(*rq), (C.c_double * 3)(*rp), C.byref(d)) == 1
    assert rel(d.value, hit(rp)) < 1e-12
    # parallel plane rejected
    assert L.pmo_plane_transfer(5.0, (C.c_double * 3)(1, 0, 0), (C.c_double * 3)(*rq), (C.c_double * 3)(0, 1, 1), C.byref(d)) == 0


def test_bilinear_reference(oracle_lib):
    """image.rs:316-326."""
    L = oracle_lib.lib()
    g = np.array([[x + 10.0 * y for x in range(5)] for y in range(5)], np.float32)
    p = g.ctypes.data_as(C.POINTER(C.c_float))
    assert abs(L.pmo_bilinear_clamped(p, 5, 5, 1.25, 2.5) - 26.25) < 1e-5
    assert abs(L.pmo_bilinear_clamped(p, 5, 5, -3.0, 0.0) - 0.0) < 1e-6


def test_levels_and_pyramid_reference(oracle_lib):
    """upsample.rs:266-276."""
    L = oracle_lib.lib()
    assert L.pmo_num_levels(539, 3, 64) == 3
    assert L.pmo_num_levels(240, 3, 64) == 2
    assert L.pmo_num_levels(90, 3, 64) == 1
    g = np.full((7, 9), 100, np.uint8)
    out = np.zeros((3, 4), np.uint8)
    outs = (C.POINTER(C.c_uint8) * 1)(out.ctypes.data_as(C.POINTER(C.c_uint8)))
    k = (C.c_double * 4)(10.0, 10.0, 4.0, 3.0)
    ko = (C.c_float * 8)()
    L.pmo_build_pyramid(g.ctypes.data_as(C.POINTER(C.c_uint8)), 9, 7, k, 2, outs, ko)
    assert abs(ko[4 + 2] - 1.75) < 1e-6 and abs(ko[4 + 0] - 5.0) < 1e-6
    assert (out == 100).all()


def test_depth_range_reference():
    """neighbors.rs:193-198: the depths 1..100 give the range (1.5, 125)."""
    import oracle

    view = scenes.DenseView("v", 4, 4, np.array([10.0, 10.0, 1.5, 1.5]), np.eye(3), np.zeros(3), np.zeros((4, 4), np.uint8))
    pts = [scenes.ScenePoint(np.array([0.0, 0.0, float(d)]), [0]) for d in range(100, 0, -1)]
    (a, b), = oracle.depth_ranges(scenes.DenseScene([view], pts, None, None))
    assert abs(a - 1.5) < 1e-12 and abs(b - 125.0) < 1e-12


def _ulp_diff(a, b):
    ia = np.frombuffer(np.float32(a).tobytes(), np.int32)[0]
    ib = np.frombuffer(np.float32(b).tobytes(), np.int32)[0]
    return abs(int(ia) - int(ib))


def test_expf_matches_libm(oracle_lib):
    L = oracle_lib.lib()
    xs = np.concatenate([np.linspace(-85.9, 88.0, 20001), np.linspace(-1, 1, 2001)]).astype(np.float32)
    worst = 0
    for x in xs:
        got = L.pmo_expf(float(x))
        want = np.float32(math.exp(float(x)))
        worst = max(worst, _ulp_diff(got, want))
    assert worst <= 2
    assert L.pmo_expf(-100.0) == 0.0


def test_sincos_matches_libm(oracle_lib):
    L = oracle_lib.lib()
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for x in np.linspace(-4.0, 4.0, 8001).astype(np.float32):
        L.pmo_sincosf(float(x), C.byref(s), C.byref(c))
        worst = max(worst, abs(s.value - math.sin(float(x))), abs(c.value - math.cos(float(x))))
    assert worst < 2e-7


def test_rcp_sample_within_one_ulp_of_division(oracle_lib):
    """The device's per-sample reciprocal against IEEE float32 division: every float of [1, 2)
    (the seed and both steps act on the significand alone) and 10^6 values over [2^-40, 2^40]
    within 1 ulp; at least 90 % correctly rounded."""
    z = np.arange(0x3F800000, 0x40000000, dtype=np.uint32).view(np.float32)
    rng = np.random.default_rng(5)
    z = np.concatenate([z, np.exp2(rng.uniform(-40, 40, 10**6)).astype(np.float32)])
    got = oracle_lib.rcp_sample(z)
    want = np.float32(1.0) / z
    d = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert d.max() <= 1
    assert (d == 0).mean() > 0.9


def _scene_and_input(w=160, h=96, seed=3):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, w, h, focal=500.0 * w / 640.0, seed=seed)
    pm, fl = abi.fast_profile()
    src = [[1, 2], [0, 2], [0, 1]]
    inp = scenes.kernel_input(sc, src, [(20.0, 80.0)] * 3, pm, fl)
    return sc, inp


def test_level_pair_homography_projects_plane_points(oracle_lib):
    """H = K_j (R + t n^T/delta) K^-1 maps a reference pixel to the projection of its plane point."""
    import oracle

    sc, inp = _scene_and_input()
    v, s = 0, 1
    ref, src = sc.views[v], sc.views[s]
    L = oracle.lib()
    kr = (C.c_float * 4)(*ref.k)
    kj = (C.c_float * 4)(*src.k)
    r, t, _ = scenes.pair_geometry(sc, v, s)
    A = (C.c_float * 9)()
    b = (C.c_float * 3)()
    L.pmo_level_pair(kr, kj, (C.c_float * 9)(*r), (C.c_float * 3)(*t), A, b)
    A = np.array(A[:], float).reshape(3, 3)
    b = np.array(b[:], float)
    rng = np.random.default_rng(0)
    for _ in range(50):
        x, y = rng.uniform(0, ref.width - 1), rng.uniform(0, ref.height - 1)
        d = rng.uniform(20, 60)
        n = -np.array([rng.normal(0, 0.3), rng.normal(0, 0.3), 1.0])
        n /= np.linalg.norm(n)
        ray = np.array([(x - ref.k[2]) / ref.k[0], (y - ref.k[3]) / ref.k[1], 1.0])
        if n @ ray >= 0:
            n = -n
        X = ray * d
        delta = d * (n @ ray)
        kinv = np.array([[1 / ref.k[0], 0, -ref.k[2] / ref.k[0]], [0, 1 / ref.k[1], -ref.k[3] / ref.k[1]], [0, 0, 1]])
        H = A + np.outer(b, n @ kinv) / delta
        p = H @ np.array([x, y, 1.0])
        Xs = r.reshape(3, 3).astype(float) @ X + t
        u = src.k[0] * Xs[0] / Xs[2] + src.k[2]
        vv = src.k[1] * Xs[1] / Xs[2] + src.k[3]
        assert abs(p[0] / p[2] - u) < 1e-3 and abs(p[1] / p[2] - vv) < 1e-3


def _identity_input(tex, src_tex=None):
    """Two identical cameras (source = reference): the fronto-parallel plane at any depth maps every
    pixel onto itself, so NCC of identical textures is exactly 1 and the cost 0."""
    h, w = tex.shape
    src_tex = tex if src_tex is None else src_tex
    f = 200.0
    k = [f, f, (w - 1) / 2.0, (h - 1) / 2.0]
    v = abi.ViewData(key=1, levels=[abi.LevelData(tex, k)], sources=[1],
                     pairs=[(np.eye(3, dtype=np.float32).reshape(9), np.zeros(3, np.float32), np.zeros(3, np.float32))],
                     depth_min=1.0, depth_max=10.0)
    v2 = abi.ViewData(key=2, levels=[abi.LevelData(src_tex, k)], sources=[0],
                      pairs=[(np.eye(3, dtype=np.float32).reshape(9), np.zeros(3, np.float32), np.zeros(3, np.float32))],
                      depth_min=1.0, depth_max=10.0)
    pm, fl = abi.fast_profile()
    return abi.KernelInput([v, v2], pm, fl)


def test_ncc_closed_forms(oracle_lib):
    import oracle

    rng = np.random.default_rng(1)
    tex = rng.integers(0, 256, (40, 48)).astype(np.uint8)
    inp = _identity_input(tex)
    o = oracle.Oracle(inp, threads=1)
    plane = [0.0, 0.0, -1.0, 5.0]
    for (x, y) in [(20, 20), (10, 8), (30, 25)]:
        assert abs(o.ncc_cost_f64(0, 0, 0, x, y, plane)) < 1e-12
        assert abs(o.ncc_cost_f32(0, 0, 0, x, y, plane)) < 2e-6
    # a normal facing away from the camera: no valid homography -> 2
    assert o.ncc_cost_f32(0, 0, 0, 20, 20, [0.0, 0.0, 1.0, 5.0]) == 2.0
    # inverted texture: rho = -1 exactly, cost 1 - rho = 2 (a dropped mean term would clamp to 0)
    oi = oracle.Oracle(_identity_input(tex, (255 - tex).astype(np.uint8)), threads=1)
    for (x, y) in [(20, 20), (10, 8)]:
        assert abs(oi.ncc_cost_f64(0, 0, 0, x, y, plane) - 2.0) < 1e-9
        assert abs(oi.ncc_cost_f32(0, 0, 0, x, y, plane) - 2.0) < 5e-5
    # constant patch -> 2 (variance below 1e-5)
    flat = np.full((40, 48), 90, np.uint8)
    o2 = oracle.Oracle(_identity_input(flat), threads=1)
    assert o2.ncc_cost_f32(0, 0, 0, 20, 20, plane) == 2.0
    assert o2.ncc_cost_f64(0, 0, 0, 20, 20, plane) == 2.0
    # shifted texture: the plane maps the centre outside the source -> 2 (translate by t = (-1, 0, 0))
    v = inp.views[0]
    v.pairs = [(np.eye(3, dtype=np.float32).reshape(9), np.array([-100.0, 0, 0], np.float32), np.array([100.0, 0, 0], np.float32))]
    inp3 = abi.KernelInput([v, inp.views[1]], inp.pm, inp.filter)
    o3 = oracle.Oracle(inp3, threads=1)
    assert o3.ncc_cost_f32(0, 0, 0, 20, 20, plane) == 2.0


def test_ncc_fp32_agrees_with_fp64(oracle_lib):
    import oracle

    sc, inp = _scene_and_input()
    o = oracle.Oracle(inp, threads=1)
    rng = np.random.default_rng(2)
    diffs = []
    for _ in range(400):
        x, y = int(rng.integers(8, 152)), int(rng.integers(8, 88))
        d = float(sc.gt_depth[0][y, x]) * rng.uniform(0.95, 1.05)
        n = sc.gt_normal[0][y, x].astype(float) + rng.normal(0, 0.1, 3)
        n /= np.linalg.norm(n)
        c32 = o.ncc_cost_f32(0, 0, 0, x, y, [*n, d])
        c64 = o.ncc_cost_f64(0, 0, 0, x, y, [*n, d])
        if c64 < 2.0:
            diffs.append(abs(c32 - c64))
    assert len(diffs) > 300
    assert max(diffs) < 1e-4


def test_ncc_fp64_brute_force(oracle_lib):
    """FP64 cost equals a direct numpy evaluation of the reference formula (patchmatch.cu:199-290)."""
    import oracle

    sc, inp = _scene_and_input()
    o = oracle.Oracle(inp, threads=1)
    ref, src = sc.views[0], sc.views[1]
    Rg = ref.gray.astype(float) / 255.0
    Sg = src.gray.astype(float) / 255.0
    r, t, _ = scenes.pair_geometry(sc, 0, 1)
    r = r.astype(float).reshape(3, 3)
    t = t.astype(float)
    Kr = np.array([[ref.k[0], 0, ref.k[2]], [0, ref.k[1], ref.k[3]], [0, 0, 1.0]])
    Ks = np.array([[src.k[0], 0, src.k[2]], [0, src.k[1], src.k[3]], [0, 0, 1.0]])
    Kr = Kr.astype(np.float32).astype(float)
    Ks = Ks.astype(np.float32).astype(float)
    offs = np.arange(-5, 6, 2)

    def at(img, x, y):
        if x < 0 or y < 0 or x >= img.shape[1] or y >= img.shape[0]:
            return 0.0
        return img[y, x]

    def sample(u, v):
        x0, y0 = math.floor(u), math.floor(v)
        ax, ay = u - x0, v - y0
        p00, p10, p01, p11 = at(Sg, x0, y0), at(Sg, x0 + 1, y0), at(Sg, x0, y0 + 1), at(Sg, x0 + 1, y0 + 1)
        top = p00 + ax * (p10 - p00)
        bot = p01 + ax * (p11 - p01)
        return top + ay * (bot - top)

    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(30):
        x, y = int(rng.integers(10, 150)), int(rng.integers(10, 86))
        d = float(sc.gt_depth[0][y, x])
        n = sc.gt_normal[0][y, x].astype(float)
        H = Ks @ (r + np.outer(t, n) / (d * (n @ np.linalg.inv(Kr) @ [x, y, 1.0]))) @ np.linalg.inv(Kr)
        pc = H @ [x, y, 1.0]
        uc, vc = pc[0] / pc[2], pc[1] / pc[2]
        if not (0 <= uc <= src.width - 1 and 0 <= vc <= src.height - 1):
            assert o.ncc_cost_f64(0, 0, 0, x, y, [*n, d]) == 2.0
            continue
        ic = Rg[y, x]
        ws, iks, svs = [], [], []
        for dy in offs:
            for dx in offs:
                ik = at(Rg, x + dx, y + dy)
                w = math.exp(-(dx * dx + dy * dy) / (2 * 25.0) - (ic - ik) ** 2 / (2 * 0.04))
                p = H @ [x + dx, y + dy, 1.0]
                ws.append(w)
                iks.append(ik)
                svs.append(sample(p[0] / p[2], p[1] / p[2]) if p[2] > 1e-8 else 0.0)
        ws, iks, svs = map(np.array, (ws, iks, svs))
        W = ws.sum()
        mu, mus = (ws * iks).sum() / W, (ws * svs).sum() / W
        var, vars_ = (ws * iks * iks).sum() / W - mu * mu, (ws * svs * svs).sum() / W - mus * mus
        if var < 1e-5 or vars_ < 1e-5:
            continue
        want = min(max(1 - ((ws * iks * svs).sum() / W - mu * mus) / math.sqrt(var * vars_), 0), 2)
        got = o.ncc_cost_f64(0, 0, 0, x, y, [*n, d])
        assert abs(got - want) < 1e-6
        checked += 1
    assert checked > 20


def test_pyramid_ramp_closed_form(oracle_lib):
    """build_pyramid (upsample.rs:78-102) of a ramp: the symmetric 3-tap blur leaves a linear ramp
    unchanged away from the clamped edges, so level 1 is the 2x2 mean 4i + 1 of the ramp 2x; the
    same along columns. A transposed pass, a wrong averaging cell or a shifted cell fails it."""
    import oracle

    h, w = 40, 64
    ramp_x = np.tile((2 * np.arange(w)).astype(np.uint8), (h, 1))
    lv = oracle.build_pyramid(ramp_x, [100.0, 100.0, 31.5, 19.5], 3)
    assert [l[0].shape for l in lv] == [(10, 16), (20, 32), (40, 64)]
    l1 = lv[1][0].astype(int)
    assert (l1[:, 1:-1] == 4 * np.arange(1, 31) + 1).all()
    assert (l1 == l1[0]).all()  # rows identical
    ramp_y = np.tile((2 * np.arange(h)).astype(np.uint8)[:, None], (1, w))
    l1y = oracle.build_pyramid(ramp_y, [100.0, 100.0, 31.5, 19.5], 2)[0][0].astype(int)
    assert (l1y[1:-1, :] == (4 * np.arange(1, 19) + 1)[:, None]).all()
    # the intrinsics follow f/2, (c + 0.5)/2 - 0.5
    assert lv[0][1] == [25.0, 25.0, 7.5, 4.5]


def _two_level_input(low_gray, high_gray, lk, hk):
    v = abi.ViewData(key=1, levels=[abi.LevelData(low_gray, lk), abi.LevelData(high_gray, hk)], sources=[], pairs=[],
                     depth_min=1.0, depth_max=10.0)
    pmp, fl = abi.fast_profile()
    return abi.KernelInput([v], pmp, fl)


def _ray(k, x, y):
    return np.array([(x - k[2]) / k[0], (y - k[3]) / k[1], 1.0])


def test_upsample_reproduces_plane(oracle_lib):
    """upsample.rs:278-307: a tilted plane upsampled from 16 x 12 to 32 x 24 matches the plane
    along every high-resolution ray (1e-5), and the median filter keeps it."""
    import oracle

    hk = [40.0, 40.0, 15.5, 11.5]
    lk = [20.0, 20.0, (15.5 + 0.5) * 0.5 - 0.5, (11.5 + 0.5) * 0.5 - 0.5]
    n = np.array([0.3, -0.2, -1.0])
    n /= np.linalg.norm(n)
    hit = lambda r: -5.0 / (n @ r)
    o = oracle.Oracle(_two_level_input(np.full((12, 16), 50, np.uint8), np.full((24, 32), 50, np.uint8), lk, hk), threads=1)
    low = abi.ViewState(16, 12)
    for y in range(12):
        for x in range(16):
            low.depth[y, x] = hit(_ray(lk, x, y))
            low.normal[y, x] = n
    up = o.upsample(1, [0], [low], 1.0, 0.04)[0]
    for y in range(24):
        for x in range(32):
            t = hit(_ray(hk, x, y))
            assert abs((up.depth[y, x] - t) / t) < 1e-5
    m = o.median_filter(1, [0], [up])[0]
    assert abs((m.depth.reshape(-1)[100] - up.depth.reshape(-1)[100]) / up.depth.reshape(-1)[100]) < 1e-5


def test_upsample_weights_and_median_fp64(oracle_lib):
    """Joint bilateral weights (patchmatch.cu:741-800) and the 5 x 5 lower median
    (patchmatch.cu:803-849) against an independent FP64 evaluation: fronto-parallel planes transfer
    their depth unchanged, so the upsampled depth is the weighted mean of the 5 x 5 low-resolution
    depths around round((x - 0.5) / 2) with weights exp(-ds / 2 s_s^2 - (I_p - I_q)^2 / 2 s_c^2),
    and the median filter output is the lower median of the 5 x 5 depths."""
    import oracle

    rng = np.random.default_rng(5)
    lo_g = rng.integers(0, 256, (12, 16)).astype(np.uint8)
    hi_g = rng.integers(0, 256, (24, 32)).astype(np.uint8)
    hk = [40.0, 40.0, 15.5, 11.5]
    lk = [20.0, 20.0, 7.5, 5.5]
    o = oracle.Oracle(_two_level_input(lo_g, hi_g, lk, hk), threads=1)
    low = abi.ViewState(16, 12)
    low.depth[...] = rng.uniform(5.0, 9.0, (12, 16)).astype(np.float32)
    low.normal[...] = np.array([0.0, 0.0, -1.0], np.float32)
    ss, sc = 1.0, 0.3
    up = o.upsample(1, [0], [low], ss, sc)[0]
    worst = 0.0
    for y in range(24):
        for x in range(32):
            px, py = (x - 0.5) * 0.5, (y - 0.5) * 0.5
            qx0, qy0 = int(np.floor(px + 0.5)) if px >= 0 else -int(np.floor(-px + 0.5)), \
                int(np.floor(py + 0.5)) if py >= 0 else -int(np.floor(-py + 0.5))
            ip = hi_g[y, x] / 255.0
            sw = sd = 0.0
            for qy in range(qy0 - 2, qy0 + 3):
                for qx in range(qx0 - 2, qx0 + 3):
                    if 0 <= qx < 16 and 0 <= qy < 12:
                        ds = (qx - px) ** 2 + (qy - py) ** 2
                        wq = math.exp(-ds / (2 * ss * ss) - (ip - lo_g[qy, qx] / 255.0) ** 2 / (2 * sc * sc))
                        sw += wq
                        sd += wq * float(low.depth[qy, qx])
            worst = max(worst, abs(up.depth[y, x] - sd / sw) / (sd / sw))
    assert worst < 2e-6
    d = rng.uniform(5.0, 9.0, (24, 32)).astype(np.float32)
    st = abi.ViewState(32, 24)
    st.depth[...] = d
    st.normal[...] = np.array([0.0, 0.0, -1.0], np.float32)
    m = o.median_filter(1, [0], [st])[0]
    for y in range(24):
        for x in range(32):
            win = np.sort(d[max(0, y - 2):y + 3, max(0, x - 2):x + 3].reshape(-1))
            assert m.depth[y, x] == win[(win.size - 1) // 2]
