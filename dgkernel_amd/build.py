"""Build libdgk.so (HIP, gfx950) in-tree, and the CPU oracle used by the tests.

Usage: python -m dgkernel_amd.build [--only lib|oracle] [-j N]

The library is compiled with hipcc for gfx950 only; every translation unit
gets ``-Xclang -target-feature -Xclang -packed-fp32-ops`` so that the
compiler never forms packed FP32 instructions. Objects go to ``_build/``.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "dgkernel_amd"
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libdgk.so"
ORACLE_SRC = ROOT / "oracle" / "dgk_oracle.cpp"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = [l for l in (res.stderr or "").splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "not a recognized feature" not in l
             and "warning generated" not in l and "warnings generated" not in l]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = [BUILD / (s.stem + ".o") for s in sources]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)] for s, o in zip(sources, objs)]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-shared", "-fPIC", "-Wall", "-o", str(tmp),
          str(ORACLE_SRC)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["lib", "oracle"], default=None)
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args(argv)
    if a.only in (None, "oracle"):
        print("built", build_oracle())
    if a.only in (None, "lib"):
        print("built", build_lib(a.j))
    return 0


if __name__ == "__main__":
    sys.exit(main())
