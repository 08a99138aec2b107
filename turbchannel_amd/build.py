"""Build the HIP engine in-tree: turbchannel_amd/libturbchannel_amd.so (gfx950),
and the command-line solver turbchannel_amd/turbch that drives it.

Every source is compiled by hipcc for gfx950 only, with packed FP32
instructions disabled, then linked into one shared library whose exported
symbols are the C ABI of include/turbchannel.h. The driver is host C++ that
uses only that C ABI; it finds the library next to itself.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libturbchannel_amd.so")
DRIVER = os.path.join(HERE, "turbch")
DRIVER_SRC = os.path.join(HERE, "driver", "turbch.cpp")
SOURCES = ["host_setup.cpp", "kernels_stencil.hip", "kernels_spectral.hip", "kernels_reduce.hip", "kernels_mlp.hip",
           "engine.cpp", "mlp.cpp", "imex.cpp"]

ARCH = os.environ.get("TC_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


# The z-solve's row loops are unrolled in full (each intermediate row has a
# place fixed at compile time); their unrolled size exceeds the default limit
# of a pragma-directed unroll.
SRC_FLAGS = {"kernels_spectral.hip": ["-mllvm", "-pragma-unroll-threshold=65536"]}


def _compile(src):
    out = os.path.join(OBJ, src + ".o")
    cmd = [hipcc(), *FLAGS, *SRC_FLAGS.get(src, []), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB, "and", DRIVER)
    return LIB


def build_driver():
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no host C++ compiler for the driver")
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    tmp = DRIVER + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
           "-o", tmp, "-L", HERE, "-lturbchannel_amd", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{rocm_lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the driver failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
