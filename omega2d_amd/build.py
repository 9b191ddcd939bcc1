"""Build recipe for the gfx950 library (and the CPU checkers beside it).

    python -m omega2d_amd.build            # library + C++ host test driver + oracle

The library is compiled in-tree with hipcc for gfx950 only:
  omega2d_amd/lib/libomega2d_hip.so   (C ABI declared in include/omega2d_hip.h)
Packed-FP32 instruction selection is disabled for every HIP translation unit.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "omega2d_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libomega2d_hip.so")
HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "host_driver.cpp")
HOST_TEST_BIN = os.path.join(LIBDIR, "o2d_host_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIPFLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
LINKLIBS = ["-lrocsolver", "-lrocblas"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    return err


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIPFLAGS, "-c", src, "-o", obj])
    return obj


def build_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, *LINKLIBS])
    os.replace(tmp, LIB)
    return LIB


def build_host_driver():
    """The C++ host mirror of the reference interface, exercised by a small driver."""
    if not os.path.exists(HOST_TEST_SRC):
        return None
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
          "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
          HOST_TEST_SRC, "-o", HOST_TEST_BIN, "-L" + LIBDIR, "-lomega2d_hip",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    return HOST_TEST_BIN


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])


def main():
    lib = build_library()
    print("built", lib)
    drv = build_host_driver()
    if drv:
        print("built", drv)
    build_oracle()
    print("built oracle")


if __name__ == "__main__":
    sys.exit(main())
