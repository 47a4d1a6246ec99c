"""Build the native libraries of the package in-tree (cup3d_amd/lib/).

  libcup3d_poisson.so   HIP kernels for gfx950 + C++ host driver + C ABI
                        (include/cup3d_poisson.h), built with hipcc
  libcup3d_workloads.so synthetic grids and right-hand sides (g++, OpenMP)
  example_projection    a caller of the header-only C++ PoissonSolverBase
                        adapter (cup3d_amd/host) on the standalone stand-in types
  libcup3d_h5.so        block fields in the reference's multi-block HDF5 layout
                        (include/cup3d_h5.h; host only, the image's HDF5)

Run: python -m cup3d_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for all device code (they have given
# silently wrong results on this hardware pool).
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INCLUDE, "-I", CSRC]

POISSON_SOURCES = ["kernels.hip", "mixed.hip", "cnn.hip", "projection.hip", "topology_gpu.hip",
                   "diffusion.hip", "noalloc.hip", "topology.cpp", "solver.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_poisson(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile_hip, POISSON_SOURCES))
    out = os.path.join(LIB, "libcup3d_poisson.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_workloads():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libcup3d_workloads.so")
    _run([CXX, "-std=c++17", "-O3", "-fopenmp", "-fPIC", "-shared", "-Wall",
          os.path.join(CSRC, "workloads.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


HDF5 = os.environ.get("CUP3D_HDF5", "/opt/conda")  # the image's serial HDF5 1.10


def build_h5():
    """libcup3d_h5.so: block fields in the reference's multi-block HDF5 layout
    (host only); skipped, with a message, where no HDF5 is installed."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        sys.stderr.write(f"libcup3d_h5.so not built: no HDF5 under {HDF5}\n")
        return None
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libcup3d_h5.so")
    _run([CXX, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
          "-I", os.path.join(HDF5, "include"), os.path.join(CSRC, "h5io.cpp"), "-o", out + ".tmp",
          "-L", os.path.join(HDF5, "lib"), "-lhdf5", f"-Wl,-rpath,{os.path.join(HDF5, 'lib')}"])
    os.replace(out + ".tmp", out)
    return out


def build_example():
    out = os.path.join(LIB, "example_projection")
    _run([CXX, "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, "-I", HOST,
          os.path.join(HOST, "example_projection.cpp"), "-o", out + ".tmp",
          "-L", LIB, "-lcup3d_poisson", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [p for p in (build_workloads(), build_poisson(), build_example(), build_h5()) if p]


if __name__ == "__main__":
    for p in build_all():
        print(p)
