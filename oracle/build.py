"""Builds the CPU oracle (test infrastructure) and, where the reference tree is present, the
reference harness in oracle/_ref/ (its src/devtypes.h compiled as it lies; nothing is copied).

  oracle/liboracle.so               the CPU restatement (ws_oracle.cpp)
  oracle/libcpu_baseline.so         multithreaded CPU restatements of the measured queries (the
                                    benchmark's CPU baseline; cpu_baseline.cpp)
  oracle/_ref/libref_devtypes.so    reference expression semantics (ref_devtypes_harness.cpp)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_SRC = "/root/reference/src"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_oracle() -> str:
    out = os.path.join(HERE, "liboracle.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "ws_oracle.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpu_baseline() -> str:
    # x86-64-v3 (AVX2) rather than -march=native: the library is built here and runs on the GPU
    # box's host, whose CPU model differs
    out = os.path.join(HERE, "libcpu_baseline.so")
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread",
          os.path.join(HERE, "cpu_baseline.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_ref() -> str | None:
    if not os.path.exists(os.path.join(REF_SRC, "devtypes.h")):
        return None
    os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
    out = os.path.join(HERE, "_ref", "libref_devtypes.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", REF_SRC,
          os.path.join(HERE, "ref_devtypes_harness.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build() -> None:
    build_oracle()
    build_cpu_baseline()
    build_ref()


if __name__ == "__main__":
    build()
    print("built oracle", "and reference harness" if build_ref() else "(no reference tree)", file=sys.stderr)
