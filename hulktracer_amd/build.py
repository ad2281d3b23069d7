"""Build libhulk.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m hulktracer_amd.build   (or hulktracer_amd.build.build())
The shared library lands in hulktracer_amd/lib/libhulk.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libhulk.so"

HIP_SOURCES = ["hulk_kernels.hip", "hulk_api.hip", "hulk_mlp.hip", "hulk_kdbuild.hip"]
CPP_SOURCES = ["kdtree_sah.cpp", "scenes.cpp", "loaders.cpp", "loader_aff.cpp", "dataset.cpp"]

# -ffp-contract=off: no implicit multiply-add fusion, so the kernels keep
# the reference's operation order bit for bit.  Packed FP32 ops are turned
# off for the device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libhulk.so")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = objdir / (src + ".o")
        jobs.append([cc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = objdir / (src + ".o")
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", f"-I{ROOT / 'include'}",
                     "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for j in jobs]
    errs = []
    for j, p in zip(jobs, procs):
        out, err = p.communicate()
        if p.returncode != 0:
            errs.append(f"$ {' '.join(j)}\n{out}{err}")
        elif verbose and err.strip():
            sys.stderr.write(err)
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("libhulk build failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
