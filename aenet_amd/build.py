"""Build libaenet_amd.so (HIP for gfx950) in-tree.

    python -m aenet_amd.build        # or: python aenet_amd/build.py

It also builds aenet_amd/_lib/predict.x, the predict.x-compatible
command-line driver (csrc/predict.cpp).

Host C++ is compiled with -ffp-contract=off so that the FP64 plan
arithmetic follows the reference's evaluation order; device code keeps
contraction for the FP32 descriptor/network math (the FP64 neighbor test
switches it off locally).  Packed-FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libaenet_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("AENET_AMD_ARCH", "gfx950")

DEVICE_SOURCES = ["kernels.hip", "atom_cheb.hip", "atom_bp.hip", "train.hip", "mc.hip", "noalloc.hip"]
HOST_SOURCES = ["model.cpp", "plan.cpp", "engine.cpp", "mc.cpp", "capi.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics"]
# host-only sources: plain g++ against the HIP runtime API
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-ffp-contract=off", "-fno-fast-math", "-Wall", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, jobs=8):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    cmds = []
    objs = []
    for src in DEVICE_SOURCES:
        o = os.path.join(obj_dir, src + ".o")
        cmds.append([HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in HOST_SOURCES:
        o = os.path.join(obj_dir, src + ".o")
        cmds.append([CXX, *COMMON, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    # a failed compile must stop the build: linking would pick up the
    # object file an earlier build left behind
    for o in objs:
        if os.path.exists(o):
            os.remove(o)
    procs = []
    rc = []
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        procs.append(subprocess.Popen(c))
        if len(procs) >= jobs:
            rc.append(procs.pop(0).wait())
    rc += [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp, *objs,
          "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")], verbose)
    os.replace(tmp, LIB)
    # predict.x: the batched command-line driver, linked against the library
    exe = os.path.join(OUT_DIR, "predict.x")
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(os.path.dirname(HERE), "include"),
          os.path.join(CSRC, "predict.cpp"), "-o", exe + ".tmp", "-L" + OUT_DIR, "-laenet_amd",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
