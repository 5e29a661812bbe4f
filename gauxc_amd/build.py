"""Build recipe for the native libraries (no CMake).

  libgauxc_amd_device.so  hipcc, gfx950: HIP kernels + the device C ABI
                          (include/gauxc_amd/device.h)
  libgauxc_amd.so         g++: GauXC-compatible C++ host library and C API
                          (include/gauxc/...), linked to the device library

Both land in gauxc_amd/lib/ (in-tree, so they travel with the source tree).
Usage: python -m gauxc_amd.build   (or gauxc_amd.build.build()).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
DEV_SRC = os.path.join(HERE, "csrc", "device")
HOST_SRC = os.path.join(HERE, "csrc", "host")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # Packed-FP32 VALU forms stay disabled: the compiler still vectorises
    # ordinary float code, just without v_pk_*_f32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + INC,
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wextra", "-I" + INC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    dev_srcs = sorted(f for f in os.listdir(DEV_SRC) if f.endswith(".hip"))
    host_srcs = sorted(f for f in os.listdir(HOST_SRC) if f.endswith(".cxx"))
    jobs_list = []
    for f in dev_srcs:
        o = os.path.join(OBJ, f + ".o")
        jobs_list.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(DEV_SRC, f), "-o", o])
    for f in host_srcs:
        o = os.path.join(OBJ, f + ".o")
        jobs_list.append([CXX, *HOST_FLAGS, "-c", os.path.join(HOST_SRC, f), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, jobs_list):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    dev_lib = os.path.join(LIB, "libgauxc_amd_device.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", dev_lib,
          *[os.path.join(OBJ, f + ".o") for f in dev_srcs]])
    host_lib = os.path.join(LIB, "libgauxc_amd.so")
    _run([CXX, "-shared", "-fPIC", "-fopenmp", "-o", host_lib,
          *[os.path.join(OBJ, f + ".o") for f in host_srcs],
          "-L" + LIB, "-lgauxc_amd_device", "-Wl,-rpath,$ORIGIN"])
    return dev_lib, host_lib


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
