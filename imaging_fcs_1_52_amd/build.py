"""Builds the native libraries of the package in-tree.

libimfcs_hip.so  HIP kernels for gfx950 + C ABI (include/imfcs.h) + JNI exports
libimfcs_sim.so  seeded particle simulator (host, OpenMP)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IMFCS_OFFLOAD_ARCH", "gfx950")

HIP_LIB = os.path.join(HERE, "libimfcs_hip.so")
SIM_LIB = os.path.join(HERE, "libimfcs_sim.so")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip(verbose=True):
    objs = []
    for src in ("imfcs_kernels.hip", "imfcs_corr.hip", "imfcs_corr_mfma.hip", "imfcs_gls.hip", "imfcs_sim.hip", "imfcs_mlp.hip", "imfcs_capi.hip"):
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    shim_obj = os.path.join(CSRC, "_build", "jni_shim.o")
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, "jni_shim.cpp"), "-o", shim_obj])
    objs.append(shim_obj)
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, HIP_LIB)


def build_sim():
    tmp = SIM_LIB + ".tmp"
    _run(["g++", "-O3", "-fPIC", "-fopenmp", "-shared", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "simulate.cpp"), "-o", tmp])
    os.replace(tmp, SIM_LIB)


def build_all():
    build_sim()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "sim" in what:
            build_sim()
        if "hip" in what:
            build_hip()
