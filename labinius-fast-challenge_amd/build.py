"""Build recipe of the MI355X prover: liblabinius_gpu.so and the labinius-fast executable.

hipcc compiles every source for gfx950 only. Packed FP32 instructions are disabled for all HIP
code (-target-feature -packed-fp32-ops), and floating point is not contracted, so the challenge
norm keeps the reference's two roundings and one add per root. Outputs land in this package's
build/ directory, inside the tree (git-ignored, so the history stays source-only).

    python labinius-fast-challenge_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "build")
LIB = os.path.join(OUT, "liblabinius_gpu.so")
CLI = os.path.join(OUT, "labinius-fast")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-ffp-contract=off", "-Wno-unused-result"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-mpclmul", "-msse4.1", "-ffp-contract=off", "-fno-fast-math",
              "-D__HIP_PLATFORM_AMD__"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in ("lbn_device.hip",):
        obj = os.path.join(OUT, src + ".o")
        run([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("lbn_host.cpp", "lbn_blake3.cpp", "lbn_sched.cpp"):
        obj = os.path.join(OUT, src + ".o")
        run(["g++", *HOST_FLAGS, "-I/opt/rocm/include", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    run(["g++", "-O2", "-std=c++17", "-o", CLI, os.path.join(CSRC, "labinius_fast.cpp"), "-L" + OUT,
         "-llabinius_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    if verbose:
        print("built", LIB, CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
