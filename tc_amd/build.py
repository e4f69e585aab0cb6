"""Build the gfx950 engine libtcx.so in-tree with hipcc (no JIT, no cache).

Every HIP source is compiled for gfx950 only, with packed FP32 instructions
disabled, and linked into tc_amd/libtcx.so next to this file.  Object files
go to tc_amd/_build/ (git-ignored) and are compiled in parallel.
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
LIB = os.path.join(HERE, "libtcx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", CSRC, "-I", INCLUDE, "-Wall", "-Wno-unused-parameter"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# paths relative to tc_amd/
SOURCES = [
    "csrc/kernels_fft.hip",
    "csrc/kernels_hf2x.hip",
    "csrc/kernels_hf2x_xy.hip",
    "csrc/kernels_tc2x.hip",
    "csrc/kernels_local.hip",
    "csrc/kernels_tc3x.hip",
    "csrc/kernels_tc3xy.hip",
    "csrc/kernels_tc3z.hip",
    "csrc/kernels_davidson.hip",
    "csrc/tcx_noalloc.hip",
    "csrc/kernels_zw.hip",
    "csrc/tcx_api.cpp",
    "csrc/tcx_tc.cpp",
    "csrc/tcx_tc3x.cpp",
    "csrc/tcx_davidson.cpp",
    "csrc/tcx_adjoint.cpp",
    "csrc/tcx_zeroweight.cpp",
    "csrc/tcx_jastrow.cpp",
    "csrc/tcx_hf2x_xy.cpp",
    "host/calc_hamiltonian.cpp",
    "host/scf.cpp",
    "host/host_api.cpp",
]
ROOT = os.path.dirname(HERE)
SHIM_TEST_SRC = os.path.join(ROOT, "tests", "host_shim_test.cpp")
SHIM_TEST_BIN = os.path.join(ROOT, "tests", "_bin", "host_shim_test")
SCF_TEST_SRC = os.path.join(ROOT, "tests", "scf_host_test.cpp")
SCF_TEST_BIN = os.path.join(ROOT, "tests", "_bin", "scf_host_test")
FFT_TEST_SRC = os.path.join(ROOT, "tests", "fft_reg_test.cpp")
FFT_TEST_BIN = os.path.join(ROOT, "tests", "_bin", "fft_reg_test")


def _compile(src: str) -> str:
    path = os.path.join(HERE, src)
    obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *DEVICE, *COMMON, "-c", path, "-o", obj]
    else:
        # host-only translation unit: the HIP runtime headers need the platform macro
        cmd = [HIPCC, *COMMON, "-x", "c++", "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c", path,
               "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_shim_test(verbose)
    build_fft_test(verbose)
    return LIB


def build_shim_test(verbose: bool = True) -> str:
    """C++ programs driving the host side (calc_hamiltonian::* and the SCF loop), run by GPU tests."""
    os.makedirs(os.path.dirname(SHIM_TEST_BIN), exist_ok=True)
    for src, out in ((SHIM_TEST_SRC, SHIM_TEST_BIN), (SCF_TEST_SRC, SCF_TEST_BIN)):
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", src, "-o", out, "-L", HERE, "-ltcx",
               "-Wl,-rpath,$ORIGIN/../../tc_amd"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"host test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        if verbose:
            print(f"built {out}")
    return SHIM_TEST_BIN


def build_fft_test(verbose: bool = True) -> str:
    """Host program checking the register FFT templates against a direct DFT (a CPU test)."""
    os.makedirs(os.path.dirname(FFT_TEST_BIN), exist_ok=True)
    cmd = [HIPCC, "-x", "c++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", CSRC,
           FFT_TEST_SRC, "-o", FFT_TEST_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"fft test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {FFT_TEST_BIN}")
    return FFT_TEST_BIN


if __name__ == "__main__":
    build()
    sys.exit(0)
