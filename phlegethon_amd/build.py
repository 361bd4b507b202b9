"""Build of the HIP library libphlx.so for gfx950 (in-tree, no JIT cache).

Flags: -ffp-contract=off keeps the reference's no-FMA operation order so the
FP64 path agrees with the CPU restatement to round-off; packed FP32 ops are
disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OUT = os.path.join(OUT_DIR, "libphlx.so")
SOURCES = ["phlx_sel_fp64.hip", "phlx_sel_mixed.hip", "phlx_sel_fp64_r5.hip",
           "phlx_sel_mixed_r5.hip", "phlx_kernels.hip", "phlx_helm.hip", "phlx_mlp.hip",
           "phlx_diffusion.hip", "phlx_api.hip", "phlx_noalloc.hip"]
ARCH = os.environ.get("PHLX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-command-line-argument"]
OBJ_DIR = os.path.join(HERE, "_build")


def _run(cmd, what, verbose):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed: {what}")
    if verbose:
        sys.stderr.write(r.stderr)


def build(verbose: bool = False) -> str:
    """Compiles each source to an object in parallel (no cross-unit device
    references, so no relocatable device code) and links libphlx.so."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), f"--offload-arch={ARCH}"] + FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, os.path.basename(c[-1]), verbose), cmds))
    tmp = OUT + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs,
         "link libphlx.so", verbose)
    os.replace(tmp, OUT)
    return OUT


H5_OUT = os.path.join(OUT_DIR, "libphlx_h5.so")
H5_PREFIX = os.environ.get("PHLX_HDF5_PREFIX", "/opt/conda")


def build_h5(verbose: bool = False):
    """Host-only library of the reference-layout HDF5 output (phlx_h5.cpp),
    built with g++ against the image's HDF5 1.10; None when HDF5 is absent
    (the output module then reports it at use)."""
    inc = os.path.join(H5_PREFIX, "include")
    lib = os.path.join(H5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = H5_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{inc}", "-o", tmp,
           os.path.join(CSRC, "phlx_h5.cpp"), f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libphlx_h5.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, H5_OUT)
    return H5_OUT


if __name__ == "__main__":
    print(build(verbose=True))
