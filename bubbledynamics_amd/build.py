"""Build recipes for the HIP library (gfx950) and the FP64 CPU oracle.

The library is compiled in-tree so that the shared object travels with the
source tree; every call recompiles (the build is a few seconds per file).
"""
from __future__ import annotations

import concurrent.futures as _cf
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "bubbledynamics_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libbubbledynamics.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
BUILD_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # reproducible FP64 parity with the oracle: no implicit FMA contraction
    "-ffp-contract=off",
    # never emit packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]
CXX = os.environ.get("CXX", "g++")
ORACLE_FLAGS = ["-O3", "-march=x86-64-v3", "-mtune=generic", "-ffp-contract=off",
                "-fno-fast-math", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_library(verbose: bool = False) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [BUILD_DIR / (s.stem + ".o") for s in srcs]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)] for s, o in zip(srcs, objs)]
    with _cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST_DIR = PKG / "host"
HOST_PROGRAMS = ("gilmore_threshold", "westervelt_dropin", "temperature_spiral")


def build_host(verbose: bool = False) -> list[Path]:
    """C host programs over the C ABI (bubbledynamics_amd/host/*.c), linked
    against the in-tree library: bd_<name> next to their sources."""
    rocm = Path(HIPCC).resolve().parent.parent
    out = []
    for name in HOST_PROGRAMS:
        exe = HOST_DIR / f"bd_{name}"
        tmp = exe.with_suffix(".tmp")
        _run([os.environ.get("CC", "gcc"), "-std=c11", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
              f"-I{rocm / 'include'}", f"-I{ROOT / 'include'}", str(HOST_DIR / f"{name}.c"), "-o", str(tmp),
              f"-L{PKG}", "-lbubbledynamics", f"-L{rocm / 'lib'}", "-lamdhip64", "-lm",
              "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{rocm / 'lib'}"])
        os.replace(tmp, exe)
        out.append(exe)
        if verbose:
            print(f"built {exe}")
    return out


def build_oracle(verbose: bool = False) -> Path:
    srcs = sorted(ORACLE_DIR.glob("*.cpp"))
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([CXX, *ORACLE_FLAGS, *map(str, srcs), "-o", str(tmp), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    if verbose:
        print(f"built {ORACLE_LIB}")
    return ORACLE_LIB


if __name__ == "__main__":
    build_oracle(verbose=True)
    build_library(verbose=True)
    build_host(verbose=True)
