"""Builds libpyre_hip.so (gfx950) in-tree with hipcc.

    python -m pyre_amd.build            # incremental (per-object mtime)
    python -m pyre_amd.build --force    # rebuild everything
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpyre_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYRE_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "bvh_build.hip", "trace.hip", "integrator.hip", "nrc.hip", "gen.cpp", "scene_io.cpp", "film_io.cpp"]
HEADERS = ["internal.hpp", "pmath.hpp", "shade.hpp", "traverse.hpp", "nrc.hpp", "preset.hpp"]
INCLUDE_HEADERS = ["pyre_hip.h", "pyre_gen.h"]

# -ffp-contract=off: the hot path reproduces the reference's unfused f32
# arithmetic. Packed-FP32 instructions are disabled for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Per-source device flags. nrc.hip: MFMA results in VGPRs, since its kernels
# round and ReLU every product, which would otherwise first copy each result
# out of the accumulation registers (about 160 moves per 64-sample chunk of a
# training step).
EXTRA = {"nrc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    deps = [path] + [os.path.join(CSRC, h) for h in HEADERS] + \
        [os.path.join(ROOT, "include", h) for h in INCLUDE_HEADERS]
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, *EXTRA.get(src, []), "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    try:
        build(force=a.force)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
