"""Build recipe for libmfa_hip (gfx950) and the Fortran driver.

Compiles in-tree so the shared library travels with the repository:
    python -m mfa_amd.build            # library + Fortran driver
The HIP code is built with -ffp-contract=off (the FP64 forms follow the
reference's operation order) and with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mfa_amd")
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "libmfa_hip.so")
DRIVER = os.path.join(PKG, "mfa_hip_prog")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose=True):
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    host_o = os.path.join(objdir, "mfa_host.o")
    dev_o = os.path.join(objdir, "mfa_device.o")
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-c",
          os.path.join(CSRC, "mfa_host.cpp"), "-o", host_o])
    noalloc_o = os.path.join(objdir, "mfa_noalloc.o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "mfa_device.hip"), "-o", dev_o])
    # the allocation-free entry point and its file-scope scratch: a code
    # object of its own, loaded only when that entry point runs
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "mfa_noalloc.hip"), "-o", noalloc_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", dev_o, noalloc_o, host_o, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    """Fortran driver (md_main_hip.f90) linked against libmfa_hip."""
    srcs = ["mfa_hip_bind.f90", "mfa_inputs_mod.f90", "md_main_hip.f90"]
    if not all(os.path.exists(os.path.join(FSRC, s)) for s in srcs):
        return None
    if not os.path.exists(FLANG):
        print("amdflang not found; Fortran driver not built", file=sys.stderr)
        return None
    objdir = os.path.join(PKG, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(objdir, s.replace(".f90", ".o"))
        _run([FLANG, "-O2", "-c", "-module-dir", objdir, os.path.join(FSRC, s), "-o", o])
        objs.append(o)
    _run([FLANG, "-O2"] + objs + ["-o", DRIVER, "-L" + PKG, "-lmfa_hip",
                                   "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return DRIVER


def main():
    build_library()
    build_driver()


if __name__ == "__main__":
    main()
