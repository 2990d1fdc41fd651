"""Build libgs3d.so (HIP kernels + C++ host library + C ABI) for gfx950.

Every source is compiled as HIP by hipcc, in parallel, into
wgpu-3dgs-viewer_amd/_build/, then linked into wgpu-3dgs-viewer_amd/libgs3d.so.
Run directly (``python wgpu-3dgs-viewer_amd/build.py``) or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgs3d.so")
ARCH = os.environ.get("GS3D_OFFLOAD_ARCH", "gfx950")

SOURCES = ["preprocess.hip", "radix_sort.hip", "tiles.hip", "composite.hip", "viewer.cpp", "batch.cpp",
           "capi.cpp", "synth.cpp", "ply.cpp", "spz.cpp"]

# -ffp-contract=off: the splat arithmetic follows a fixed operation order
# (csrc/math.hpp) so that depth keys and visibility match the f32 oracle bit
# for bit. Packed FP32 instructions are disabled for all device code.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, build_dir: str = BUILD, defines=()) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + FLAGS + [f"-D{d}" for d in defines]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile and link; `defines` (NAME=VALUE) set tuning macros such as
    GS3D_BIN_ITEMS, `out` another library path (load it with
    GS3D_LIBRARY)."""
    build_dir = BUILD if out == LIB else out + ".build"
    os.makedirs(build_dir, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines), SOURCES))
    tmp = out + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=LIB, help="library path (default: the package's libgs3d.so)")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="NAME=VALUE tuning define")
    args = ap.parse_args()
    build(verbose=True, out=os.path.abspath(args.out), defines=args.defines)
    sys.exit(0)
