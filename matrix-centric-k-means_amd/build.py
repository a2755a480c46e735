"""Build the gfx950 engine library, the drop-in Kmeans class and the CLI.

    python matrix-centric-k-means_amd/build.py [--jobs N]

Outputs (in-tree, git-ignored, shipped with the tree):
    matrix-centric-k-means_amd/lib/libkmc.so     C ABI of include/kmc.h
    matrix-centric-k-means_amd/bin/gpukmeans     CLI with the reference's flags
    matrix-centric-k-means_amd/bin/raftbench     the explicit-Lloyd benchmark program

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["kmc_lloyd.hip", "kmc_api.hip", "kmc_kkmeans.hip", "kmc_kk_api.hip",
               "kmc_batch.hip", "kmc_static.hip", "kmc_probe.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(jobs=4, verbose=False):
    for d in (LIB, BIN, OBJ):
        os.makedirs(d, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libkmc.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, so)
    # host C++: the drop-in Kmeans class and the gpukmeans CLI (g++, no HIP headers)
    host_srcs = [os.path.join(CSRC, f) for f in ("kmeans.cpp", "gpukmeans.cpp")]
    if all(os.path.exists(s) for s in host_srcs):
        exe = os.path.join(BIN, "gpukmeans")
        _run([CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + INCLUDE] + host_srcs +
             ["-L" + LIB, "-lkmc", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"])
        os.replace(exe + ".tmp", exe)
    # the explicit-Lloyd benchmark program of the reference's harness
    rb = os.path.join(CSRC, "raftbench.cpp")
    if os.path.exists(rb):
        exe = os.path.join(BIN, "raftbench")
        _run([CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + INCLUDE, rb,
              "-L" + LIB, "-lkmc", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"])
        os.replace(exe + ".tmp", exe)
    return so


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    if a.clean:
        shutil.rmtree(OBJ, ignore_errors=True)
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    sys.exit(main())
