"""Build the HIP library (libmhd_gpu.so) and the Fortran driver in-tree for gfx950.

Translation units and their floating-point contracts:
  mhd_gpu.hip        C-ABI, update/CT/fill/dt kernels   binary64, -ffp-contract=off
  mhd_ops_fp64.hip   binary64 flux / EMF / reconstruct  -ffp-contract=off
  mhd_ops_fp32.hip   fp32 flux / EMF (mixed precision)  -ffp-contract=fast
  mhd_surrogate.hip  learned reconstruction (bf16 MFMA) -ffp-contract=fast
  mhd_fused.hip      fused fp32 flux + binary64 update  -ffp-contract=fast (update: contraction off)
Packed-fp32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmhd_gpu.so")
DRIVER = os.path.join(PKG, "mhd3d_gpu.x")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
ARCH = os.environ.get("MHD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-I" + os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
UNITS = {
    "mhd_gpu.hip": ["-ffp-contract=off"],
    "mhd_ops_fp64.hip": ["-ffp-contract=off"],
    "mhd_ops_fp32.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "mhd_surrogate.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "mhd_fused.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout[-6000:] + r.stderr[-6000:])
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return r


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)

    def one(unit):
        src = os.path.join(CSRC, unit)
        obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
        _run([HIPCC] + COMMON + UNITS[unit] + ["-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(one, list(UNITS)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_driver(verbose=False):
    """the Fortran driver (ISO_C_BINDING over libmhd_gpu.so)."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(BUILD, exist_ok=True)
    mod = os.path.join(FSRC, "mhd_gpu_mod.f90")
    drv = os.path.join(FSRC, "mhd3d_gpu.f90")
    if not (os.path.exists(mod) and os.path.exists(drv)):
        return None
    tmp = DRIVER + ".tmp"
    _run([FLANG, "-O2", "-ffp-contract=off", "-J", BUILD, mod, drv, "-o", tmp, "-L" + PKG, "-lmhd_gpu", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + PKG])
    os.replace(tmp, DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build_library(verbose=True)
    build_driver(verbose=True)
