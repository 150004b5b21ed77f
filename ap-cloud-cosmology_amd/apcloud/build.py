"""Builds libapcloud.so (HIP kernels for gfx950 + the C ABI + the C++ host
classes) in-tree with hipcc, and the oracle with make.

Every HIP source is compiled with packed-FP32 instructions disabled.  The
topology kernels (keys, selection, neighbour search, mover) are compiled
with -ffp-contract=off so their FP64 arithmetic matches the reference's
(x86-64, no FMA) bit for bit.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)            # ap-cloud-cosmology_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
HOST = os.path.join(ROOT, "host")
INCLUDE = os.path.join(REPO, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libapcloud.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

NO_CONTRACT = {"apc_sort.hip", "apc_select.hip", "apc_neigh.hip", "apc_mover.hip"}

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), p.stdout, p.stderr))
    return p


def _compile(src):
    name = os.path.basename(src)
    obj = os.path.join(BUILD, name + ".o")
    if src.endswith(".hip"):
        flags = list(COMMON)
        if name in NO_CONTRACT:
            flags.append("-ffp-contract=off")
        cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}",
               f"-I{HOST}", "-c", src, "-o", obj]
    _run(cmd)
    return obj


def sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = []
    if os.path.isdir(HOST):
        host = sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                      if f.endswith(".cpp") and f != MAIN)
    return hip + host


MAIN = "apcloud_main.cpp"
EXE = os.path.join(PKG, "apcloud")


def build_driver():
    """The `apcloud <file.param> [-ksp_* ...]` executable over libapcloud.so."""
    obj = _compile(os.path.join(HOST, MAIN))
    tmp = EXE + ".tmp"
    _run([HIPCC, "-o", tmp, obj, f"-L{PKG}", "-lapcloud", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, EXE)


def build_library(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB, "and", EXE)
    return LIB


def build_oracle(verbose=True):
    oracle = os.path.join(REPO, "oracle")
    _run(["make", "-C", oracle, "all"])
    if os.path.isdir("/root/reference"):
        _run(["make", "-C", oracle, "ref"])
    if verbose:
        print("built oracle")


def build(verbose=True):
    build_library(verbose=verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    build()
    sys.exit(0)
