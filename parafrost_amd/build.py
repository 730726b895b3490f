"""Build the native libraries in-tree.

libpfgpu.so : the MI355X simplifier (HIP for gfx950 + host driver + C ABI)
libpfgen.so : seeded CNF generators (host C)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PFG_ARCH", "gfx950")

HIP_SOURCES = ["prims.hip", "cnf.hip", "lcve.hip", "sub.hip", "ve.hip", "ere.hip", "proof.hip", "noalloc.hip",
               "sigma.cpp"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter",
             "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    tmp = os.path.join(HERE, "libpfgpu.so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, os.path.join(HERE, "libpfgpu.so"))
    gtmp = os.path.join(HERE, "libpfgen.so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", gtmp, os.path.join(CSRC, "gen.c"), "-lm"])
    os.replace(gtmp, os.path.join(HERE, "libpfgen.so"))


if __name__ == "__main__":
    build(verbose=True)
