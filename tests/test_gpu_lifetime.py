"""Lifetime of the backend handle and snapshot scale checks, through the C ABI.

A Session borrows its backend for its whole lifetime in the reference (kernel.rs:244), so Rust
cannot drop the backend first; a C caller can. c3d_pm_destroy therefore defers the release while
sessions or learned-initialiser networks are open, refuses new ones, and the last one to close
frees the backend. A geometric run reads the snapshot with the sizes of the run's level, so a
snapshot of another level is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from cumulus3d_amd import abi, init_net, pm, scenes

pytestmark = pytest.mark.gpu


def _three_view(w, h, seed):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, w, h, focal=500.0 * w / 640.0, seed=seed)
    pmp, fl = abi.fast_profile()
    return sc, scenes.kernel_input(sc, [[1, 2], [0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl)


def test_destroy_before_session_and_network_defers_the_release():
    L = init_net._declare(pm.load())
    _, inp = _three_view(96, 64, seed=3)
    be = C.c_void_p()
    assert L.c3d_pm_create(0, None, C.byref(be)) == abi.C3D_OK
    s = C.c_void_p()
    assert L.c3d_pm_begin(be, C.byref(inp.c), C.byref(s)) == abi.C3D_OK
    net = C.c_void_p()
    assert L.c3d_init_create(be, 1, 16, 24, 1, C.byref(net)) == abi.C3D_OK
    # wrong order: the backend first, while a session and a network hold it
    assert L.c3d_pm_destroy(be) == 2
    s2, net2 = C.c_void_p(), C.c_void_p()
    assert L.c3d_pm_begin(be, C.byref(inp.c), C.byref(s2)) == abi.C3D_INVALID_ARGUMENT
    assert b"destroyed" in L.c3d_pm_last_error()
    assert L.c3d_init_create(be, 1, 16, 24, 1, C.byref(net2)) == abi.C3D_INVALID_ARGUMENT
    # the open session still works on its backend's stream
    st = [abi.ViewState(96, 64) for _ in range(3)]
    arr = abi.states_array(st)
    assert L.c3d_pm_run(s, abi.views_array([0, 1, 2]), 3, arr, C.byref(abi.run_params(0, False, True, 1, 1, False)), None) == abi.C3D_OK
    L.c3d_init_destroy(net)  # touches the backend's device and stream: still alive
    L.c3d_pm_end(s)  # the last holder: frees the backend
    # a fresh backend in the right order is released at once
    be3 = C.c_void_p()
    assert L.c3d_pm_create(0, None, C.byref(be3)) == abi.C3D_OK
    s3 = C.c_void_p()
    assert L.c3d_pm_begin(be3, C.byref(inp.c), C.byref(s3)) == abi.C3D_OK
    L.c3d_pm_end(s3)
    assert L.c3d_pm_destroy(be3) == 0


def test_snapshot_of_another_level_is_refused():
    sc, _ = _three_view(128, 96, seed=4)
    pmp, fl = abi.fast_profile()
    levels = []
    for v in sc.views:
        k = [float(q) for q in v.k]
        half = np.ascontiguousarray(v.gray[::2, ::2])
        kh = [k[0] * 0.5, k[1] * 0.5, (k[2] + 0.5) * 0.5 - 0.5, (k[3] + 0.5) * 0.5 - 0.5]
        levels.append([abi.LevelData(half, kh), abi.LevelData(v.gray, k)])
    inp = scenes.kernel_input(sc, [[1, 2], [0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl, levels=levels)
    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        st = [abi.ViewState(128, 96) for _ in range(3)]
        s.run([0, 1, 2], st, abi.run_params(1, False, True, 1, 1, False))
        coarse = abi.DepthSnapshot(7, 0, [np.ones((48, 64), np.float32)] * 3)
        with pytest.raises(pm.PmError) as e:
            s.run([0, 1, 2], st, abi.run_params(1, True, False, 1, 2, False), coarse)
        assert e.value.status == abi.C3D_INVALID_ARGUMENT and "snapshot of level 0" in e.value.message
        with pytest.raises(pm.PmError) as e:
            s.evaluate(1, [0, 1, 2], st, True, coarse)
        assert e.value.status == abi.C3D_INVALID_ARGUMENT
        # the same id at the right level is accepted (the id cache is keyed on the level too)
        fine = abi.DepthSnapshot(7, 1, [x.depth.copy() for x in st])
        s.run([0, 1, 2], st, abi.run_params(1, True, False, 1, 2, False), fine)
    be.close()
