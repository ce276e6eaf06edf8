"""Builds libfhhip.so (the C-ABI of include/fh_hair.h) for gfx950 with hipcc.

Every translation unit is compiled with -ffp-contract=off: the chain's discrete
results (tile membership, depth bins, grid cells, i32 grid values) must not
depend on where the compiler chose to fuse a multiply-add.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "lib"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = LIB_DIR / "libfhhip.so"
SOURCES = ["fh_api.hip", "fh_sim.hip", "fh_raster.hip", "fh_shading.hip", "fh_post.hip", "fh_hwraster.hip", "fh_meshes.hip",
           "fh_uniforms.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link libfhhip.so. Always recompiles."""
    BUILD_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = BUILD_DIR / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    # --no-undefined: a missing definition fails the build, not the first dlopen
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", *map(str, objs), "-o",
          str(tmp)])
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
