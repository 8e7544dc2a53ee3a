"""Builds libdiaphane_amd.so in-tree with hipcc for gfx950.

Run as ``python -m diaphane_amd.build`` (or through ``__graft_entry__.build``).
Objects go to ``diaphane_amd/_build/``; the library lands next to this file,
where the Python binding loads it from.
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
LIBRARY = HERE / "libdiaphane_amd.so"
ARCH = os.environ.get("DIAPHANE_AMD_ARCH", "gfx950")

# Every float expression keeps its written order: no a*b+c contraction and
# IEEE single-precision division, which is what makes the kernels
# bit-identical to the reference's CPU sweep. Packed FP32 is disabled.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]
DEVICE = [
    "-x", "hip", f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]

SOURCES = [
    ("kernels.hip", DEVICE),
    ("kernels_pair.hip", DEVICE),
    ("kernels_stack.hip", DEVICE),
    ("noalloc.hip", DEVICE),
    ("scene.cpp", HOST),
    ("scene_ron.cpp", HOST),
    ("timeline.cpp", HOST),
    ("engine.cpp", HOST),
    ("batch.cpp", HOST),
    ("server.cpp", HOST),
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build diaphane_amd")


def build(verbose: bool = False) -> Path:
    compiler = hipcc()
    BUILD.mkdir(exist_ok=True)
    commands, objects = [], []
    for name, flags in SOURCES:
        src = CSRC / name
        obj = BUILD / (src.stem + ".o")
        commands.append([compiler, *flags, *COMMON, "-c", str(src), "-o", str(obj)])
        objects.append(str(obj))
    # The translation units compile in parallel (the two kernel units take
    # minutes each; the host units seconds).
    with ThreadPoolExecutor(max_workers=min(len(commands), os.cpu_count() or 1)) as pool:
        for cmd in commands:
            if verbose:
                print(" ".join(cmd), flush=True)
        list(pool.map(lambda cmd: subprocess.run(cmd, check=True), commands))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [compiler, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objects, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
