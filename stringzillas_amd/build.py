"""Builds ``libstringzillas_amd.so`` in-tree with hipcc for gfx950.

Every HIP translation unit is compiled with the packed-FP32 target feature disabled, then linked into one shared
library next to this file. Usage: ``python -m stringzillas_amd.build`` (or ``python stringzillas_amd/build.py``).
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libstringzillas_amd.so")
SOURCES = ["dp_kernels.hip", "runtime.hip", "batch.hip", "abi.hip", "server.hip", "tape_file.hip", "fingerprints.hip",
           "noalloc.hip"]
ARCH = os.environ.get("SZS_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and os.path.exists(candidate):
            return candidate
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build stringzillas_amd")


COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(source: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(source)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-c", os.path.join(CSRC, source), "-o", obj]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"compiling {source} failed:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    noise = [line for line in result.stderr.splitlines() if "packed-fp32-ops" not in line and line.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objects = list(pool.map(_compile, SOURCES))
    temporary = LIBRARY + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", temporary, "-lpthread"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"linking failed:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    os.replace(temporary, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build()
