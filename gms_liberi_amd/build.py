"""Build the HIP library in-tree with hipcc for gfx950.

    python -m gms_liberi_amd.build            # -> gms_liberi_amd/libliberi_hip.so

The kernels of the 55 integral classes are split over several translation
units (one per class group) that compile in parallel; the host orchestration
is one more unit.  Every HIP unit is compiled with packed FP32 disabled
(`-target-feature -packed-fp32-ops`).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libliberi_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", str(PKG.parent / "include")]

# class-kernel translation units: LIBERI_CLASS_GROUP selects which classes a unit instantiates
KERNEL_GROUPS = 8


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    jobs.append([HIPCC, *COMMON, "-c", str(CSRC / "api.cpp"), "-o", str(OBJ / "api.o")])
    jobs.append([HIPCC, *COMMON, "-DLIBERI_CLASS_GROUP=-1", "-c", str(CSRC / "eri_kernels.hip"), "-o",
                 str(OBJ / "eri_common.o")])
    jobs.append([HIPCC, *COMMON, "-c", str(CSRC / "noalloc.hip"), "-o", str(OBJ / "noalloc.o")])
    jobs.append([HIPCC, *COMMON, "-DLIBERI_CLASS_GROUP=-3", "-c", str(CSRC / "eri_kernels.hip"), "-o",
                 str(OBJ / "eri_fused.o")])
    for g in range(KERNEL_GROUPS):
        jobs.append([HIPCC, *COMMON, f"-DLIBERI_CLASS_GROUP={g}", f"-DLIBERI_CLASS_GROUPS={KERNEL_GROUPS}", "-c",
                     str(CSRC / "eri_kernels.hip"), "-o", str(OBJ / f"eri_g{g}.o")])
    nproc = min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with ThreadPoolExecutor(nproc) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
