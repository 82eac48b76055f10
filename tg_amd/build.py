"""Builds the native parts of tg_amd in-tree.

    python -m tg_amd.build            # libtgx.so (gfx950) + oracle/liboracle.so

libtgx.so is the product: HIP kernels for gfx950 plus the C ABI declared in
include/tgx.h, with the host-only generators of include/tgx_gen.h linked in.
Device code is compiled with -ffp-contract=off (no FMA contraction), IEEE
fp32 division and square root and denormals preserved, which the bit-exact
visibility and SVO paths rely on; packed fp32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libtgx.so"
ORACLE_DIR = ROOT / "oracle"

ARCH = os.environ.get("TGX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["tgx_runtime.hip", "tgx_visibility.hip", "tgx_svo.hip", "tgx_gi.hip", "tgx_nrc.hip", "tgx_gen_dev.hip", "tgx_frame.hip", "tgx_shade.hip", "tgx_dynamic.hip", "tgx_prims.hip"]
C_SOURCES = ["tgx_gen.c", "tgx_io.c"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CC_FLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        jobs.append(([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], str(obj)))
    for src in C_SOURCES:
        obj = BUILD / (Path(src).stem + ".c.o")
        jobs.append((["gcc", *CC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], str(obj)))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        list(pool.map(lambda j: _run(j[0]), jobs))
    objs = [o for _, o in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp), "-lm"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run(["make", "-C", str(ORACLE_DIR), "-s"])
    return ORACLE_DIR / "liboracle.so"


def main(argv: list[str]) -> int:
    what = argv[1:] or ["lib", "oracle"]
    if "lib" in what:
        build_lib()
    if "oracle" in what:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
