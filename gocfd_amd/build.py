"""Build libgocfd_euler2d.so in-tree with hipcc for gfx950.

Run as ``python -m gocfd_amd.build``. The library is written next to this
file, so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgocfd_euler2d.so")

SOURCES = ["dfr_setup.cpp", "capi_common.cpp", "euler2d_capi.cpp", "ensemble_capi.cpp", "adjoint_runner.cpp",
           "kernels_common.hip", "adjoint_common.hip", "noalloc.hip"] + [f"stage_n{n}.hip" for n in range(5)] + [
           f"stage_multi_f{f}.hip" for f in range(4)] + [
           f"adjoint_n{n}.hip" for n in range(5)]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 vector ops are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result",
]


def _compile(src: str, build_dir: str = BUILD, defines=()) -> str:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + [f"-D{d}" for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile every source and link the shared library; returns its path.
    `defines` are extra preprocessor definitions (e.g. GOCFD_TILE=32) and
    `out` the library path, for building variants beside the default."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build gocfd_amd")
    build_dir = BUILD if out == LIB else os.path.join(os.path.dirname(os.path.abspath(out)), "_build")
    os.makedirs(build_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 8)) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines), SOURCES))
    LIBP = out
    tmp = LIBP + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIBP)
    if verbose:
        print(f"built {LIBP}")
    return LIBP


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
