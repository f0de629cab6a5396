"""Build libflipgpu.so (HIP kernels + C-ABI) for gfx950 with hipcc.

Usage: python -m flipsolver2d_amd.build   (or flipsolver2d_amd.build.build())
Objects go to flipsolver2d_amd/_build/, the library to flipsolver2d_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libflipgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["fg_particles.hip", "fg_grid.hip", "fg_pcg.hip", "fg_mg.hip", "fg_init.hip", "fg_nbflip.hip", "fg_visc.hip",
           "fg_api.hip", "fg_serve.cpp"]
# -ffp-contract=off: every float expression rounds as written (the kernels
# are compared bit for bit with the CPU restatement).  Packed FP32 ops are
# disabled for all device code.
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_facade()
    if verbose:
        print(f"built {LIB}, {FACADE}, {DEMO}")
    return LIB


CPP = os.path.join(PKG, "cpp")
FACADE = os.path.join(LIBDIR, "libflipsolver2d_hip.so")
DEMO = os.path.join(LIBDIR, "facade_demo")
CXX = os.environ.get("CXX", "g++")


def _build_facade() -> None:
    """The reference-named C++ API (host-only C++ over the C-ABI) and its
    demo driver, linked against libflipgpu.so next to them."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CPP]
    link = ["-L", LIBDIR, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    steps = [
        [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc
        + [os.path.join(CPP, "flipsolver2d_hip.cpp"), os.path.join(CPP, "json_scene.cpp")]
        + link + ["-lflipgpu", "-o", FACADE],
        [CXX, "-O2", "-std=c++17", "-Wall"] + inc + [os.path.join(CPP, "facade_demo.cpp")]
        + link + ["-lflipsolver2d_hip", "-lflipgpu", "-o", DEMO],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ API build failed:\n{' '.join(cmd)}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
