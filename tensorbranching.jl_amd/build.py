"""Builds libtbh.so (HIP, gfx950) and libtbh_workload.so (host) in-tree.

Run as a module or call build(); outputs go to tensorbranching.jl_amd/lib/.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TBH_OFFLOAD_ARCH", "gfx950")
# packed fp32 instructions are disabled for all device code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-command-line-argument", f"-I{INC}", f"-I{CSRC}"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-pthread",
              "-Wall", f"-I{INC}"]

TBH_SOURCES = ["tbh_kernels.hip", "tbh_plan.cpp", "tbh_runtime.cpp", "tbh_serve.cpp",
               "tbh_regions.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in TBH_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = os.path.join(LIB, "libtbh.so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, os.path.join(LIB, "libtbh.so"))
    tmp = os.path.join(LIB, "libtbh_workload.so.tmp")
    _run(["g++", *HOST_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "workload.cpp")])
    os.replace(tmp, os.path.join(LIB, "libtbh_workload.so"))
    if verbose:
        print("built", os.listdir(LIB))


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
