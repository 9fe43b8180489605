"""Build libgoosefem_amd.so (HIP for gfx950) in-tree.

    python -m goosefem_amd.build          # or goosefem_amd.build.build()

Every translation unit under csrc/ is compiled by hipcc into _build/, then linked into
goosefem_amd/libgoosefem_amd.so.  Flags:
  -ffp-contract=off      no implicit FMA: FP64 kernels keep the reference's expression order
  -packed-fp32-ops off   no packed FP32 instructions (plain FP32 ALU ops only)
  code object v5         loadable by both the ROCm 7.2 runtime and the one bundled with torch
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgoosefem_amd.so")
ARCH = os.environ.get("GF_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))
    )


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


H5_SRC = os.path.join(HERE, "csrc_h5", "gf_h5.cpp")
H5_LIB = os.path.join(HERE, "libgoosefem_amd_h5.so")
HDF5_DIR = os.environ.get("GF_HDF5_DIR", "/opt/conda")


def build_h5(verbose: bool = True) -> str | None:
    """The HDF5 output library (host C++, over the HDF5 C library in $GF_HDF5_DIR, /opt/conda by
    default, linked with an rpath to it).  Skipped, with a message, where HDF5 is absent."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"HDF5 not found under {HDF5_DIR}: {H5_LIB} not built")
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = H5_LIB + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{inc}", H5_SRC, "-o", tmp, f"-L{libdir}",
           "-lhdf5", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {H5_SRC}:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print(f"built {H5_LIB}")
    return H5_LIB


def build(verbose: bool = True) -> str:
    """Compile all sources and link the shared library (and the HDF5 output library).  Always
    rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
