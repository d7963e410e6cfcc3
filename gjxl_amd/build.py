"""Builds libgjxl_amd.so (gfx950) from gjxl_amd/csrc with hipcc.

Usage: python -m gjxl_amd.build   (or gjxl_amd.build.build_library()).
Object files go to gjxl_amd/_build/; the shared library lands beside this
file so that it travels with the source tree.  build_caller() compiles the
C++ test caller (tests/plain_path_caller.cpp, plain g++ against
include/gjxl_amd.h) into tests/_bin/.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
SOURCE_DIR = PACKAGE_DIR / "csrc"
BUILD_DIR = PACKAGE_DIR / "_build"
LIBRARY = PACKAGE_DIR / "libgjxl_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMPILE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++20",
    # Two roundings for a*b+c unless the source says fma: the reference's
    # CPU arithmetic has no contraction.
    "-ffp-contract=off",
    "-fno-fast-math",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I", str(REPO_DIR / "include"),
]


def _compile(source: Path) -> Path:
    obj = BUILD_DIR / (source.stem + ".o")
    cmd = [HIPCC, *COMPILE_FLAGS, "-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            f"hipcc failed for {source.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build_library(jobs: int | None = None) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(SOURCE_DIR.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
           *map(str, objects), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


CALLER_SOURCE = REPO_DIR / "tests" / "plain_path_caller.cpp"
CALLER = REPO_DIR / "tests" / "_bin" / "plain_path_caller"


def build_caller() -> Path:
    """A host-only C++ program that links libgjxl_amd.so through its C
    header; the rpath finds the library beside the package."""
    CALLER.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", str(REPO_DIR / "include"),
           str(CALLER_SOURCE), "-o", str(CALLER), "-L", str(PACKAGE_DIR), "-lgjxl_amd",
           "-Wl,-rpath,$ORIGIN/../../gjxl_amd"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"caller build failed:\n{proc.stdout}\n{proc.stderr}")
    return CALLER


if __name__ == "__main__":
    path = build_library()
    build_caller()
    print(path)
    sys.exit(0)
