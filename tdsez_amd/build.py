"""Build the HIP extension ``tdsez_amd/libtdz.so`` for gfx950 with hipcc.

The library holds the HIP kernels, the C++ host code and the C ABI of
``include/tdz.h``; it is built in-tree so that it travels with the sources.
Each source compiles to its own object (in parallel; kernels are launched only
from host code of their own file) and the objects link into one library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtdz.so")
OBJ_DIR = os.path.join(HERE, "_build")
SOURCES = ["tdz_kernels.hip", "tdz_learn.hip", "tdz_galerkin.hip", "tdz_noalloc.hip", "tdz_api.cpp", "tdz_basis.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (known wrong results on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = [[HIPCC] + FLAGS + ["-c", s, "-o", o] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), jobs))
    failed = [r for r in results if r.returncode != 0]
    for r in results:
        if r.returncode != 0 or verbose:
            sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed building libtdz.so")
    tmp = LIB + ".tmp"
    r = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libtdz.so")
    os.replace(tmp, LIB)
    return LIB


HDF5_ROOT = os.environ.get("TDZ_HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(HERE, "libtdz_io.so")


def build_io(verbose=False):
    """Host library of include/tdz_io.h (HDF5 checkpoints and traces); returns None
    when no HDF5 installation is found (the propagator does not need it)."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        sys.stderr.write(f"tdsez_amd: no hdf5.h under {inc}; libtdz_io.so not built\n")
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + inc, os.path.join(HERE, "csrc", "tdz_io.cpp"), "-L" + lib, "-Wl,-rpath," + lib, "-lhdf5",
           "-o", tmp]
    r = _run(cmd, verbose)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libtdz_io.so")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_io(verbose=True))
