"""Build recipe for the engine library and the CPU oracle.

Compiles every HIP source for gfx950 with hipcc, the host-only C++ with g++,
links pramaan_amd/lib/libpramaan_hip.so, and builds oracle/liboracle.so
(the CPU checker used by tests, smoke() and the bench's CPU baseline) plus,
when the reference sources are present, oracle/_ref (see oracle/Makefile).
Everything is rebuilt on every call.

Usage: python -m pramaan_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "pramaan_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libpramaan_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
ARCH = os.environ.get("PRAMAAN_ARCH", "gfx950")

HIP_SOURCES = ["engine.hip", "launchers.hip", "batch.hip", "gnn.hip", "learned.hip", "presolve.hip"]
HOST_SOURCES = ["host/gen.cpp", "host/gen_rolling.cpp", "host/gnn_features.cpp"]
CXX_SOURCES = ["cxx/model_ir.cpp", "cxx/pdhg_solver.cpp", "cxx/backend_hip.cpp", "cxx/test_abi.cpp",
               "cxx/certificate.cpp", "cxx/precision_ladder.cpp", "cxx/presolve.cpp"]
CXX_LIB = PKG / "lib" / "libpramaan_cxx.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed FP32 instructions are kept out of every device build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _run(cmd: list[str], what: str) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed: {what}\n$ {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def _obj_for(src: str) -> Path:
    return OBJ / (src.replace("/", "_") + ".o")


def build_library(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    steps = []
    for src in HIP_SOURCES:
        steps.append(([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(_obj_for(src))], src))
    for src in HOST_SOURCES:
        if (CSRC / src).exists():
            steps.append(([CXX, *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(_obj_for(src))], src))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_run, cmd, what) for cmd, what in steps]
        for f in futs:
            f.result()
    objs = [str(_obj_for(s)) for s in HIP_SOURCES + HOST_SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-L/opt/rocm/lib",
          "-lrocprofiler-sdk-roctx"], "link")
    os.replace(tmp, LIB)
    return LIB


def build_cxx() -> Path:
    """The C++ interface layer (include/pramaan/...) over the C-ABI."""
    srcs = [str(CSRC / s) for s in CXX_SOURCES]
    tmp = CXX_LIB.with_suffix(".so.tmp")
    _run([CXX, *CXX_FLAGS, "-DPRAMAAN_ENABLE_HIP", "-shared", *srcs, "-o", str(tmp), f"-L{LIB.parent}", "-lpramaan_hip",
          "-Wl,-rpath,$ORIGIN"], "libpramaan_cxx")
    os.replace(tmp, CXX_LIB)
    return CXX_LIB


def build_oracle() -> Path:
    _run([CC, "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          str(ORACLE_DIR / "pdhg_oracle.c"), "-o", str(ORACLE_LIB), "-lm"], "oracle")
    return ORACLE_LIB


def build_reference_oracle() -> bool:
    """Builds oracle/_ref from the reference's CPU sources when present."""
    mk = ORACLE_DIR / "Makefile"
    if not Path("/root/reference").is_dir() or not mk.exists() or shutil.which("make") is None:
        return False
    _run(["make", "-B", "-s", "-C", str(ORACLE_DIR), "-j8", "ref", "reftests"], "oracle/_ref")
    return True


def build_all(jobs: int = 8) -> None:
    build_library(jobs)
    build_cxx()
    build_oracle()
    build_reference_oracle()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build_all(args.jobs)
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
