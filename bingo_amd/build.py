"""Builds bingo_amd/libbingo.so for gfx950 (hipcc; no GPU needed).

    python -m bingo_amd.build

Kernels and the host runtime are compiled by hipcc, the OpenMP generators by
g++; everything links into one in-tree shared library that the C ABI in
include/bingo.h describes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbingo.so")
CLI = os.path.join(HERE, "bingo_cli")
ARCH = os.environ.get("BINGO_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 ops have given silent wrong results on these parts
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    # the host pass warns that it does not know the packed-fp32 feature: harmless
    noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise and verbose:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("kernels.hip", "api.cpp", "apps.cpp", "server.cpp"):
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    gen_obj = os.path.join(BUILD, "gen.o")
    _run(["g++", "-O3", "-march=x86-64-v2", "-fPIC", "-fopenmp", "-std=c++17", "-Wall",
          f"-I{os.path.join(ROOT, 'include')}", "-c", os.path.join(CSRC, "gen.cpp"), "-o", gen_obj], verbose)
    objs.append(gen_obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    # the reference's command line (src/main.cpp) over the C ABI
    cli_tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", os.path.join(CSRC, "cli.cpp"),
          "-o", cli_tmp, f"-L{HERE}", "-lbingo", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(cli_tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
