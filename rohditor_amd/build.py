"""Builds librhd.so in-tree with hipcc for gfx950.

Every source is compiled with -ffp-contract=off (the reference never
contracts a*b+c), correctly rounded f32 division and square root, and packed
FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "librhd.so"
OBJ = HERE / "_build"
ARCH = os.environ.get("RHD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{HERE.parent / 'include'}", f"-I{OBJ}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed for {src.name}")
    return obj


def log2_table() -> Path:
    """Generates the log2 exception list of csrc/libm_f32.hpp from the host
    libm (gen_log2_table.cc), which csrc/libm.hip compiles in."""
    OBJ.mkdir(exist_ok=True)
    exe = OBJ / "gen_log2_table"
    inc = OBJ / "log2f_exceptions.inc"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", f"-I{CSRC}", str(CSRC / "gen_log2_table.cc"), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("compiling gen_log2_table failed")
    res = subprocess.run([str(exe), str(inc)], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gen_log2_table failed")
    return inc


def build(jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    log2_table()
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of librhd.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
