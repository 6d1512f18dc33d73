"""Builds liblbm355.so (HIP, gfx950) and the sim_tcpc executable in-tree.

    python -m nonnewlbm_amd.build          # library + CLI
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
LIB = PKG / "liblbm355.so"
EXE = PKG / "bin" / "sim_tcpc"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LBM355_ARCH", "gfx950")

# packed-FP32 instructions are disabled for every device compile
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(target: Path, sources: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(s.stat().st_mtime > t for s in sources)


def build_library(force: bool = True) -> Path:
    srcs = [CSRC / "lbm355_kernels.hip", CSRC / "lbm355_host.cpp", CSRC / "lbm355_pool.cpp"]
    deps = srcs + [CSRC / "lbm355_internal.h", ROOT / "include" / "lbm355.h"]
    if not force and not _stale(LIB, deps):
        return LIB
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for s in srcs:
        o = objdir / (s.stem + ".o")
        lang = ["-x", "hip"]
        _run([HIPCC, *lang, *DEVICE_FLAGS, "-c", str(s), "-o", str(o)])
        objs.append(str(o))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cli(force: bool = True) -> Path:
    src = CSRC / "sim_tcpc_main.cpp"
    if not src.exists():
        return EXE
    deps = [src, LIB, ROOT / "include" / "lbm355.h"]
    if not force and not _stale(EXE, deps):
        return EXE
    EXE.parent.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", str(src), "-o", str(EXE), f"-L{PKG}", "-llbm355",
          f"-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,$ORIGIN"])
    return EXE


def build_all(force: bool = True) -> None:
    build_library(force)
    build_cli(force)


if __name__ == "__main__":
    build_all(force="--if-stale" not in sys.argv)
