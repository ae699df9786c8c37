"""Builds libcns_hip.so and the cns_run driver in-tree with hipcc for gfx950.

    python -m cns_nospec_amd.build [--verbose]

The library is one shared object: HIP kernels plus the extern "C" entry
points declared in include/*.h.  cns_run (csrc/cns_run.cpp) is the reference
program's main loop on top of it, linked against it with an $ORIGIN rpath.
Device code is built with -ffp-contract=off so that every kernel keeps the
reference's rounding (bit-identical results), and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libcns_hip.so")
DRIVER = os.path.join(PKG, "cns_run")
ARCH = "gfx950"

SOURCES = ["cns_kernels.hip", "cns_stencil.hip", "cns_abi.hip", "cns_batch.hip",
           "cns_batch_engine.hip", "cns_io.cpp", "cns_stencil_pipe.hip", "cns_domain.hip"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(ln for ln in (res.stdout + res.stderr).splitlines()
                     if "packed-fp32-ops" not in ln)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs)


def build(verbose: bool = False, extra=()) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    common = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-pthread",
              "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              f"-I{INCLUDE}", f"-I{CSRC}", "-Wno-unused-result"]
    tmp = LIB + ".tmp"
    _run(common + ["-fPIC", "-shared", "-o", tmp, *srcs, *extra], verbose)
    os.replace(tmp, LIB)
    tmp = DRIVER + ".tmp"
    _run(common + ["-o", tmp, os.path.join(CSRC, "cns_run.cpp"), f"-L{PKG}", "-lcns_hip",
                   "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
