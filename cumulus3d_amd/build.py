"""Builds the native libraries of cumulus3d_amd in-tree (cumulus3d_amd/lib/):

* libc3d_pm.so        the gfx950 PatchStereo engine and its learned initialiser (HIP kernels + C ABI of
                      include/c3d_pm.h and include/c3d_init.h)
* libc3d_scenes.so    the synthetic terrain renderer used by tests and benchmarks (host C)
* libc3d_render.so    the same renderer on the device (HIP), for the large benchmark workloads

All HIP code is compiled for gfx950 only, with -ffp-contract=off (every multiply-add is written out)
and without packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return r.stdout


def _engine(name: str, defines: list[str]) -> str:
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)
    objs = []
    jobs = []
    for src in ("pm_kernels.hip", "pm_session.cpp", "nn_kernels.hip", "nn_net.cpp", "fusion.hip", "postproc.hip", "neighbors.hip"):
        obj = os.path.join(LIB, "obj", f"{name}_{os.path.splitext(src)[0]}.o")
        jobs.append([HIPCC, *HIP_FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(4) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(LIB, f"lib{name}.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_scenes() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libc3d_scenes.so")
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-Wall",
          os.path.join(CSRC, "terrain.c"), "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_render() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libc3d_render.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "terrain.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> list[str]:
    os.makedirs(LIB, exist_ok=True)
    stale = os.path.join(LIB, "libc3d_pm_repro.so")
    if os.path.exists(stale):
        os.remove(stale)
    with ThreadPoolExecutor(3) as ex:
        futs = [ex.submit(_engine, "c3d_pm", []), ex.submit(build_scenes), ex.submit(build_render)]
        return [f.result() for f in futs]


if __name__ == "__main__":
    for p in build_all():
        print(p)
