"""Builds libgls.so (HIP kernels + C ABI) and the gpuls shell for gfx950, in-tree.

Usage:  python -m culs_amd.build [--jobs N]
Outputs: culs_amd/libgls.so, culs_amd/gpuls  (object files under culs_amd/_build/)
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          "-I" + SRC, "-I" + os.path.join(HERE, "..", "include")]

LIB_SOURCES = ["strash.hip", "rewrite.hip", "batch.hip", "balance.hip", "sim.hip", "capi.cpp", "host_io.cpp",
               "cpu_strash.cpp", "server.cpp", "tune.cpp", "netbridge.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + COMMON + lang + ["-c", os.path.join(SRC, src), "-o", obj])
    return obj


def build(jobs=None, verbose=True):
    os.makedirs(OUT, exist_ok=True)
    jobs = jobs or min(8, len(LIB_SOURCES))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    lib = os.path.join(HERE, "libgls.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, lib)
    shell = os.path.join(HERE, "gpuls")
    if os.path.exists(os.path.join(SRC, "gpuls.cpp")):
        _run([HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"),
              os.path.join(SRC, "gpuls.cpp"), "-o", shell + ".tmp", "-L" + HERE, "-lgls",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(shell + ".tmp", shell)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
