"""Builds boris2_amd/libboris_mi355x.so (HIP kernels + C-ABI) for gfx950 with hipcc.

Run as ``python -m boris2_amd.build`` or through ``__graft_entry__.build()``.  Sources are
compiled in parallel into boris2_amd/_build/ and linked in place; the .so is what the
ctypes layer loads.  Every HIP translation unit is built with packed-FP32 instruction
selection disabled (the ``-packed-fp32-ops`` target feature).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libboris_mi355x.so"
ARCH = os.environ.get("BM_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", str(PKG.parent / "include"),
]
SOURCES = ["bm_demag.hip", "bm_mlayer.hip", "bm_batch.hip", "bm_adjoint.hip", "bm_row_rkdp.hip", "bm_row_rk4.hip", "bm_row_heun.hip", "bm_ode.hip", "bm_common.cpp", "bm_tensor.cpp", "bm_ovf2.cpp"]


def _compile(src: Path, obj: Path) -> tuple[Path, str]:
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{err}")
    return obj, err


def build(verbose: bool = False) -> Path:
    """Compile every source (always) and link the shared library."""
    BUILD.mkdir(exist_ok=True)
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for s in SOURCES:
            src = CSRC / s
            obj = BUILD / (src.stem + ".o")
            jobs.append(ex.submit(_compile, src, obj))
        objs = []
        for j in jobs:
            obj, msg = j.result()
            if verbose and msg.strip():
                print(msg)
            objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
