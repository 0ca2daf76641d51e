"""Build recipe of the native library (hipcc, gfx950).

Compiles milton_amd/csrc/*.hip|*.cpp into milton_amd/libmilton_pt.so (the C ABI
of include/milton_pt.h and include/milton_pt_scenes.h) and the Milton plugin
facade into milton_amd/libmilton_pt_plugin.so.  Run as a module or call
build(); it always recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmilton_pt.so")
PLUGIN = os.path.join(HERE, "libmilton_pt_plugin.so")
DRIVER = os.path.join(HERE, "mpt_plugin_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPT_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions stay disabled: the compiler would otherwise form
# v_pk_* FP32 ops from float2-shaped code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["mpt.hip", "scenegen.cpp", "serve.cpp", "tune.cpp", "scenefile.cpp"]
# Host-only sources (no HIP): compiled by the host compiler, which optimises
# the kd builder's loops better than hipcc's host pass does (about 2x).
HOST_SOURCES = ["kdbuild.cpp", "filmio.cpp"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-x", "hip", *srcs, *COMMON, "-pthread", "-x", "none", *objs, "-o", tmp, "-lz"])
    os.replace(tmp, LIB)
    plugin_src = os.path.join(CSRC, "plugin.cpp")
    if os.path.exists(plugin_src):
        tmp = PLUGIN + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
              plugin_src, "-o", tmp, "-L", HERE, "-lmilton_pt", "-Wl,-rpath,$ORIGIN", "-ldl"])
        os.replace(tmp, PLUGIN)
        tmp = DRIVER + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              os.path.join(CSRC, "plugin_driver.cpp"), "-o", tmp, "-L", HERE, "-lmilton_pt",
              "-Wl,-rpath,$ORIGIN", "-ldl"])
        os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
