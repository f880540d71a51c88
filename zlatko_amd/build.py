"""Builds libccr_hip.so (HIP, gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host-only C++ sources go through
hipcc as plain C++17.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libccr_hip.so"
ARCH = os.environ.get("CCR_HIP_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "amc.hip", "mlp.hip", "mlp_host.cpp", "capi.cpp"]
HOST_SOURCES = ["plan.cpp", "draws.cpp", "app.cpp"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-parameter"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, verbose: bool) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", *COMMON, *HIP_FLAGS, "-c", str(path), "-o", str(obj)]
    elif src in HOST_SOURCES:
        cmd = [hipcc(), *COMMON, "-c", str(path), "-o", str(obj)]
    else:
        cmd = [hipcc(), "-x", "hip", *COMMON, *HIP_FLAGS, "-c", str(path), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), HIP_SOURCES + HOST_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
