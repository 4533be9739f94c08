"""Build recipe for libmcpet.so (gfx950 HIP + host C++), the mcgpu_pet CLI and the
oracle's CPU libraries.

Usage:  python mcgpu-pet_amd/build.py [--no-oracle] [-j N]

Everything is built in-tree: mcgpu-pet_amd/lib/, mcgpu-pet_amd/bin/, oracle/*.so.
The HIP flags keep IEEE semantics (no contraction, correctly rounded f32 divide and
sqrt, denormals preserved) so that the device arithmetic is reproducible on the host.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-I" + CSRC,
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + INC, "-I" + CSRC]
# Oracle: portable x86-64-v3 (AVX2 + FMA) so it runs on the host of any GPU box.
ORACLE_FLAGS = ["-O3", "-std=gnu11", "-fPIC", "-shared", "-march=x86-64-v3", "-ffp-contract=off"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit("build failed: " + os.path.basename(cmd[-1] if cmd else "?"))
    return r


def build_lib(jobs):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    units = [
        ([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, "kernels.hip"), "-o", os.path.join(BUILD, "kernels.o")]),
        ([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, "coinc.hip"), "-o", os.path.join(BUILD, "coinc.o")]),
        ([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, "runtime.cpp"), "-o", os.path.join(BUILD, "runtime.o")]),
        ([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, "server.cpp"), "-o", os.path.join(BUILD, "server.o")]),
        ([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, "unet.hip"), "-o", os.path.join(BUILD, "unet.o")]),
        ([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, "host_io.cpp"), "-o", os.path.join(BUILD, "host_io.o")]),
    ]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(run, units))
    so = os.path.join(LIB, "libmcpet.so")
    tmp = so + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
         os.path.join(BUILD, "kernels.o"), os.path.join(BUILD, "coinc.o"), os.path.join(BUILD, "runtime.o"),
         os.path.join(BUILD, "server.o"), os.path.join(BUILD, "unet.o"),
         os.path.join(BUILD, "host_io.o"), "-lz"])
    os.replace(tmp, so)
    cli = os.path.join(BIN, "mcgpu_pet")
    run([CXX, *HOST_FLAGS, "-o", cli + ".tmp", os.path.join(CSRC, "cli.cpp"), "-L" + LIB, "-lmcpet",
         "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(cli + ".tmp", cli)
    return so


def build_oracle():
    src = os.path.join(ORACLE, "pet_oracle.c")
    outs = []
    for name, extra in (("liboracle_libm.so", []), ("liboracle_fmath.so", ["-DORACLE_SHARED_FMATH"])):
        out = os.path.join(ORACLE, name)
        run([CC, *ORACLE_FLAGS, *extra, "-o", out + ".tmp", src, "-lz", "-lm", "-lpthread"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-lib", action="store_true")
    ap.add_argument("-j", type=int, default=3)
    a = ap.parse_args(argv)
    if not a.no_lib:
        print(build_lib(a.j))
    if not a.no_oracle:
        for o in build_oracle():
            print(o)


if __name__ == "__main__":
    main()
