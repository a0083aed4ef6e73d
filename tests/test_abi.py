"""The C ABI boundary on the CPU: the engine library loads, exports every entry point that
include/c3d_pm.h declares, its POD layouts match the ctypes mirror (as a C compiler lays them out),
and without a gfx950 device it reports NOT_AVAILABLE instead of falling back to anything."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from cumulus3d_amd import abi, init_net, pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "c3d_pm.h")
INIT_HEADER = os.path.join(ROOT, "include", "c3d_init.h")
FUSION_HEADER = os.path.join(ROOT, "include", "c3d_fusion.h")
SCENE_HEADER = os.path.join(ROOT, "include", "c3d_scene.h")


def declared_functions(header=HEADER, prefix="c3d_"):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_header_declares_the_trait_surface():
    names = declared_functions()
    for n in ["c3d_pm_create", "c3d_pm_name", "c3d_pm_begin", "c3d_pm_run", "c3d_pm_evaluate", "c3d_pm_filter",
              "c3d_pm_upsample", "c3d_pm_median", "c3d_pm_end", "c3d_pm_destroy", "c3d_pm_last_error"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(pm.library_path())
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(pm.EXPORTS) == declared_functions()


def test_library_exports_the_initialiser():
    L = C.CDLL(pm.library_path())
    names = declared_functions(INIT_HEADER, "c3d_(?:init|pm)_")
    assert not [n for n in names if not hasattr(L, n)]
    assert sorted(init_net.EXPORTS) == names


def test_library_exports_fusion_and_scene_preparation():
    from cumulus3d_amd import neighbors

    L = C.CDLL(pm.library_path())
    for header, exports in ((FUSION_HEADER, ["c3d_fuse_consistency"]), (SCENE_HEADER, neighbors.EXPORTS)):
        names = declared_functions(header)
        assert names and not [n for n in names if not hasattr(L, n)]
        assert sorted(exports) == names


STRUCTS = {
    "c3d_pm_options": abi.c3d_pm_options,
    "c3d_level_image": abi.c3d_level_image,
    "c3d_pair_geometry": abi.c3d_pair_geometry,
    "c3d_kernel_view": abi.c3d_kernel_view,
    "c3d_pm_params": abi.c3d_pm_params,
    "c3d_filter_params": abi.c3d_filter_params,
    "c3d_kernel_input": abi.c3d_kernel_input,
    "c3d_view_state": abi.c3d_view_state,
    "c3d_depth_snapshot": abi.c3d_depth_snapshot,
    "c3d_run_params": abi.c3d_run_params,
    "c3d_post_params": abi.c3d_post_params,
    "c3d_schedule": abi.c3d_schedule,
    "c3d_timing": abi.c3d_timing,
}


STRUCTS.update({
    "c3d_sweep_source": init_net.c3d_sweep_source,
    "c3d_sweep_view": init_net.c3d_sweep_view,
    "c3d_init_train_params": init_net.c3d_init_train_params,
    "c3d_learned_schedule": init_net.c3d_learned_schedule,
})


def _more_structs():
    from cumulus3d_amd import fusion, neighbors

    return {
        "c3d_fusion_view": fusion.c3d_fusion_view,
        "c3d_fusion_params": fusion.c3d_fusion_params,
        "c3d_point_cloud": fusion.c3d_point_cloud,
        "c3d_sparse_scene": neighbors.c3d_sparse_scene,
    }


def test_struct_layouts_match_the_header():
    STRUCTS.update(_more_structs())
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', f'#include "{INIT_HEADER}"',
             f'#include "{FUSION_HEADER}"', f'#include "{SCENE_HEADER}"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for name, cls in STRUCTS.items():
        assert int(got[f"{name} size"]) == C.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert int(got[f"{name}.{fname}"]) == getattr(cls, fname).offset, (name, fname)


def test_no_device_means_not_available():
    if pm.is_available():
        pytest.skip("a gfx950 device is present")
    with pytest.raises(pm.PmError) as e:
        pm.HipPatchMatch(0)
    assert e.value.status == abi.C3D_NOT_AVAILABLE


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.delenv("C3D_PM_LIBRARY", raising=False)
    monkeypatch.setattr(pm, "_LIB_DIR", str(tmp_path))
    monkeypatch.setattr(pm, "_libs", {})
    with pytest.raises(pm.PmError) as e:
        pm.load()
    assert e.value.status == abi.C3D_NOT_AVAILABLE


def test_library_path_can_name_another_build(monkeypatch, tmp_path):
    """$C3D_PM_LIBRARY selects another build of the engine; a path that does not exist fails as a
    missing in-tree build does, naming the path."""
    import shutil

    monkeypatch.setattr(pm, "_libs", {})
    monkeypatch.setenv("C3D_PM_LIBRARY", str(tmp_path / "nowhere.so"))
    with pytest.raises(pm.PmError) as e:
        pm.load()
    assert e.value.status == abi.C3D_NOT_AVAILABLE and "nowhere.so" in e.value.message
    monkeypatch.delenv("C3D_PM_LIBRARY")
    in_tree = pm.library_path()
    other = tmp_path / "copy.so"
    shutil.copy(in_tree, other)
    monkeypatch.setattr(pm, "_libs", {})
    monkeypatch.setenv("C3D_PM_LIBRARY", str(other))
    assert pm.library_path() == str(other)
    L = pm.load()
    assert L._name == str(other) and hasattr(L, "c3d_pm_create")
