"""Build librmw_hip.so (gfx950) in-tree with hipcc, plus the test helpers.

Usage: python -m rigidmultiblobswall_amd.build [--jobs N]

Every HIP translation unit under csrc/ is compiled for gfx950 only and
linked into one shared library next to this file.  Packed FP32 VALU
instructions are disabled (-packed-fp32-ops): on this hardware pool some
forms of them have produced silently wrong results.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librmw_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(REPO, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_library(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not sources:
        raise RuntimeError("no HIP sources found")
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_test_helpers() -> str:
    """Host-only build of the pair-term header for CPU checks (tests only)."""
    src = os.path.join(REPO, "tests", "helpers", "pair_terms_host.cpp")
    out = os.path.join(REPO, "tests", "helpers", "libpair_terms_host.so")
    if os.path.exists(src):
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, src])
    return out


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) is built from oracle/Makefile."""
    _run(["make", "-B", "-C", os.path.join(REPO, "oracle")])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    lib = build_library(args.jobs, args.verbose)
    print("built", lib)


if __name__ == "__main__":
    sys.exit(main())
