"""Builds libeig_amd.so (hand-written gfx950 HIP kernels + C-ABI) and the smartpca
CLI in-tree.  Run: python -m eig_amd.build  (or via __graft_entry__.build()).

Every .hip translation unit is compiled on its own for gfx950 only; files whose
FP64 arithmetic must round exactly as the reference's C code (column statistics)
are built without FMA contraction.  All device code is built with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libeig_amd.so")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# translation units and whether they may contract a*b+c into FMA
UNITS = {
    "ctx.hip": False,
    "syrk.hip": True,
    "eigh.hip": True,
    "project.hip": False,
    "serve.hip": False,
    "fpca.hip": False,
    "gemm.hip": True,
    "normal.hip": False,
}

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature: drop that notice
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for unit, contract in UNITS.items():
        src = os.path.join(CSRC, unit)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
        flags = COMMON + ["-ffp-contract=" + ("fast" if contract else "off")]
        jobs.append([HIPCC] + flags + ["-c", src, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    build_cli()
    return LIB


def build_cli() -> None:
    """The smartpca drop-in (C host code linked against libeig_amd.so), the twstats
    tool, and the host statistics as a library of their own for the CPU tests."""
    if not os.path.isdir(HOST):
        return
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                  if f.endswith(".c") and f != "twstats_main.c")
    os.makedirs(BIN, exist_ok=True)
    cflags = ["gcc", "-O2", "-std=c11", "-Wall", "-Wno-unused-result", "-D_GNU_SOURCE",
              "-I" + os.path.join(ROOT, "include")]
    exe = os.path.join(BIN, "smartpca")
    _run(cflags + ["-o", exe + ".tmp"] + srcs +
         ["-L" + PKG, "-leig_amd", "-Wl,-rpath,$ORIGIN/..", "-lm", "-lpthread"])
    os.replace(exe + ".tmp", exe)
    tw = os.path.join(BIN, "twstats")
    _run(cflags + ["-o", tw + ".tmp", os.path.join(HOST, "twstats_main.c"),
                   os.path.join(HOST, "stats.c"), os.path.join(HOST, "parfile.c"), "-lm"])
    os.replace(tw + ".tmp", tw)
    so = os.path.join(PKG, "libeig_hoststats.so")
    _run(cflags + ["-fPIC", "-shared", "-o", so + ".tmp", os.path.join(HOST, "stats.c"), "-lm"])
    os.replace(so + ".tmp", so)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
