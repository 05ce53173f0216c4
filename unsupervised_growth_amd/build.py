"""Build the engine library (libhvc_engine.so) in-tree for gfx950.

    python -m unsupervised_growth_amd.build

hipcc compiles the kernels and the host C ABI into one shared object next to
this file.  FMA contraction is disabled (-ffp-contract=off) so the FP64 path
evaluates the reference's expression trees operation by operation; packed FP32
instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhvc_engine.so")
SOURCES = ["hvc_kernels.hip", "hvc_kernels_f32.hip", "hvc_noalloc.hip", "hvc_engine.cpp", "hvc_topology.cpp", "hvc_driver.cpp", "hvc_chain.cpp"]
ARCH = os.environ.get("HVC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=False, lib=LIB, defines=()):
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_obj_" + src.replace(".", "_") + ".o")
        cmd = [hipcc(), "-x", "hip", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
               "-ffp-contract=off", "-fno-fast-math",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value", "-I", CSRC] + ["-D" + d for d in defines] + [
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", lib] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    for o in objs:
        os.remove(o)
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
