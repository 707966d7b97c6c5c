"""Builds libbsn_hip.so (the HIP kernels and the C ABI) for gfx950 in-tree,
then libbsn_io.so (the snapshot and checkpoint files, plain C on the HDF5
library of the image, BSN_HDF5_ROOT, default /opt/conda).

    python binary-supernova_amd/build.py

The libraries are always rebuilt when this runs.  Flags:
  -ffp-contract=off   no a*b+c contraction outside code that opts in, so the
                      FP64 path reproduces the reference's rounding;
  -packed-fp32-ops    disabled (packed FP32 VALU ops are not emitted).
libbsn_io.so is C so that it needs no C++ runtime from the HDF5 installation;
its run path holds that installation's lib directory and its own directory.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbsn_hip.so")
IO_LIB = os.path.join(HERE, "libbsn_io.so")
HDF5_ROOT = os.environ.get("BSN_HDF5_ROOT", "/opt/conda")
SOURCES = ["bsn_kernels.hip", "bsn_poisson.hip", "bsn_fft.hip", "bsn_bodies.hip",
           "bsn_slabs.hip", "bsn_api.cpp"]
ARCH = os.environ.get("BSN_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-I", CSRC, "-I", os.path.join(HERE, "..", "include"),
           "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES] + [
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    build_io(verbose)
    return LIB


def build_io(verbose: bool = False) -> str:
    tmp = IO_LIB + ".tmp"
    inc = os.path.join(HERE, "..", "include")
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-I", inc, "-idirafter", os.path.join(HDF5_ROOT, "include"),
           "-o", tmp, os.path.join(CSRC, "bsn_io.c"),
           "-L", HERE, "-lbsn_hip", "-Wl,-rpath,$ORIGIN",
           "-L", os.path.join(HDF5_ROOT, "lib"), "-lhdf5",
           "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed ({r.returncode}):\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build(verbose=True))
