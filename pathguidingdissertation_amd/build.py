"""Build the gfx950 product library ``libpgtrace.so`` in-tree.

Host C++ (scene preparation, BVH build, procedural scenes) is compiled with g++,
device code and the C ABI with hipcc for gfx950. FMA contraction is off on both
so that intersection arithmetic rounds exactly as the reference's does.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpgtrace.so"
EXAMPLE = PKG / "pgd_render"  # example host program over the C++ mirror

HOST_SOURCES = ["host_prep.cpp", "scenes.cpp", "metrics.cpp", "scene_loader.cpp", "jpeg_decode.cpp"]
HIP_SOURCES = ["kernels.hip", "guiding.hip", "mlp.hip", "batch.hip", "bvh_build.hip", "api.hip"]

CXXFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-Wall",
            "-Wno-unused-function"]
HIPFLAGS = ["--offload-arch=gfx950", "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
            "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> Path:
    """Compile every source (in parallel) and link ``libpgtrace.so`` (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        cmds.append([os.environ.get("CXX", "g++"), *CXXFLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        cmds.append([hipcc, *HIPFLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1), 16))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_run, cmd) for cmd in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *map(str, objs), "-lz", "-o", str(tmp)])
    os.replace(tmp, LIB)
    # example host program over include/pgd/pgd.hpp, linked against the library
    ex_tmp = EXAMPLE.with_suffix(".tmp")
    _run([os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", str(PKG / "examples" / "pgd_render.cpp"),
          "-o", str(ex_tmp), f"-L{PKG}", "-lpgtrace", f"-Wl,-rpath,$ORIGIN"])
    os.replace(ex_tmp, EXAMPLE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
