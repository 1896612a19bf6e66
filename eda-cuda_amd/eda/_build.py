# SPDX-License-Identifier: Apache-2.0
"""Builds libeda_hip.so from eda-cuda_amd/csrc for gfx950 with hipcc (no CMake, no JIT cache).

Numerics policy of every translation unit: -ffp-contract=off (no fused multiply-add that the source
does not spell out), no fast-math, denormals kept; packed FP32 instructions disabled."""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD = ROOT / "_build"
LIB = PKG / "libeda_hip.so"

SOURCES = ["vhgw.hip", "reduce.hip", "select.hip", "flags.hip", "ccl.hip", "orient.hip", "element.hip", "driver.hip", "noalloc.hip", "adjoint.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fgpu-flush-denormals-to-zero=false" if False else "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("eda: hipcc not found (ROCm is required to build libeda_hip.so)")
    return path


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{CSRC}", f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"eda: compiling {src.name} failed:\n{proc.stderr[-6000:]}")


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = [BUILD / (Path(s).stem + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, CSRC / s, o) for s, o in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"eda: linking libeda_hip.so failed:\n{proc.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"eda: built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
