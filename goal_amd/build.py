"""Build the gfx950 shared library libgoal_amd.so in-tree with hipcc.

Usage: python goal_amd/build.py [-v] [output.so]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgoal_amd.so"

ARCH = os.environ.get("GOAL_AMD_ARCH", "gfx950")

# Minimum waves per SIMD requested for the element kernel (register cap).
ELEM_WAVES = int(os.environ.get("GOAL_ELEM_WAVES", "3"))
# Elements per workgroup of the matrix-core (neo-Hookean ten-node, 4-point) element kernel.
MF_G = int(os.environ.get("GOAL_MF_G", "16"))
# 0 builds the vector-ALU element-matrix phase for every kernel instead.
MF = int(os.environ.get("GOAL_MF", "1"))

COMMON = [
    "-O3",
    f"-DGOAL_ELEM_WAVES={ELEM_WAVES}",
    f"-DGOAL_MF_G={MF_G}",
    f"-DGOAL_MF={MF}",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    f"-I{ROOT / 'include'}",
    f"-I{SRC}",
]
# Device code: CDNA4 only, hardware FP64 atomics, no packed-FP32 instruction
# selection (some of its forms have given wrong results on this hardware).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]

SOURCES = [
    "goal_mesh.cpp",
    "goal_adapt.cpp",
    "goal_meshio.cpp",
    "goal_expr.cpp",
    "goal_amg.cpp",
    "goal_capi.cpp",
    "goal_assembly.hip",
    "goal_util.hip",
    "goal_graph.hip",
    "goal_dwr.hip",
    "goal_solve.hip",
    "goal_learn.hip",
]


def hipcc() -> str:
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return h


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    cmd = [hipcc()] + COMMON + HIP_FLAGS + ["-x", "hip", "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)


def build(verbose: bool = False, out: Path | None = None) -> Path:
    lib = Path(out) if out else LIB
    BUILD.mkdir(exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        futs = []
        for s in SOURCES:
            src = SRC / s
            obj = BUILD / (src.stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj, verbose))
        for f in futs:
            f.result()
    tmp = lib.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "-v"]
    print(build(verbose="-v" in sys.argv, out=Path(args[0]) if args else None))
