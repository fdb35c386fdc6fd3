"""Builds liblgv_hip.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m langevina_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblgv_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LGV_OFFLOAD_ARCH", "gfx950")

SOURCES = ["lgv_kernels.hip", "lgv_api.hip", "lgv_em.hip"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # every rounding follows the reference expressions: no contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


BIN = os.path.join(PKG, "bin")
DRIVER = os.path.join(BIN, "lgv_superpions")
HDF5_INCLUDE = os.environ.get("LGV_HDF5_INCLUDE", "/opt/conda/include")


# offline post-processing tools, named as ModelG/x2k.py looks them up
# ($MODELGPATH/x2k.exe, $MODELGPATH/x2k_rotated.exe)
TOOLS = {"x2k.exe": "x2k.cpp", "x2k_rotated.exe": "x2k_rotated.cpp"}


def _host_program(out, src, link):
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host"),
           "-idirafter", HDF5_INCLUDE, os.path.join(PKG, "host", src), "-o", tmp, *link]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build of {src} failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build_driver(verbose=False):
    """The C++ run driver and the x2k tools (host code only): HDF5 headers at
    compile time, libhdf5 dlopen'ed at run time, liblgv_hip.so (driver only)
    found through $ORIGIN."""
    os.makedirs(BIN, exist_ok=True)
    jobs = [(DRIVER, "superpions.cpp", ["-L", PKG, "-llgv_hip", "-Wl,-rpath,$ORIGIN/..", "-ldl"])]
    jobs += [(os.path.join(BIN, exe), src, ["-ldl"]) for exe, src in TOOLS.items()]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda j: _host_program(*j), jobs))
    if verbose:
        for out, _, _ in jobs:
            print("built", out)
    return DRIVER


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if os.path.exists(os.path.join(HDF5_INCLUDE, "hdf5.h")):
        build_driver(verbose)
    elif verbose:
        print("hdf5.h not found: run driver not built")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
