"""Builds the gfx950 step library (libbepu_gpu.so) in-tree with hipcc.

Usage: python -m bepuphysics1_amd.build [--jobs N]

Flags that matter for parity with the reference's FP32 arithmetic:
  -ffp-contract=off                       no fused multiply-add contraction
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libbepu_gpu.so")
ARCH = os.environ.get("BEPU_OFFLOAD_ARCH", "gfx950")

SOURCES = ["k_bodies.hip", "k_broadphase.hip", "k_bvh.hip", "k_grid.hip", "k_narrowphase.hip", "k_convex.hip",
           "k_island.hip", "k_solver.hip", "k_sleep.hip", "k_ccd.hip", "k_joint.hip", "k_mesh.hip", "hull.cpp",
           "mesh_tree.cpp", "world.cpp", "noalloc.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_one(src, obj):
    path = os.path.join(CSRC, src)
    cmd = [hipcc(), "-c", "-x", "hip", "--offload-arch=" + ARCH] + COMMON + [path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = [os.path.join(OUT_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda a: compile_one(*a), zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "--offload-arch=" + ARCH, "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
