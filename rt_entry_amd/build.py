"""Builds librte.so (HIP kernels + C ABI, gfx950) in-tree.

    python -m rt_entry_amd.build [--jobs N]

HIP sources are compiled with hipcc for gfx950 only; host C++ (the kd-tree)
with g++ and OpenMP.  Packed-FP32 instructions are disabled for device code
and floating-point contraction is off so that the only fused multiply-adds
are the explicit ones.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librte.so")
DRIVER = os.path.join(HERE, "rte_search")
HOST = os.path.join(HERE, "host")

HIP_SOURCES = ["api.hip", "search.hip", "search_nr1.hip", "search_nr2.hip", "search_nr4.hip",
               "search_nr8.hip", "search_nr1_mw.hip", "search_nr2_mw.hip", "search_nr4_mw.hip",
               "search_nr8_mw.hip", "search_nr1_lq.hip", "search_nr2_lq.hip",
               "search_nr4_lq.hip", "search_nr8_lq.hip", "noalloc.hip", "links.hip", "entry.hip", "project.hip", "mlp.hip", "rerank.hip", "knn.hip", "pca.hip", "kdtree_gpu.hip"]
CXX_SOURCES = ["kdtree.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
             f"-I{INCLUDE}", f"-I{CSRC}"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        cmds.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    jobs = jobs or min(len(cmds), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    # host C++ API (Graph mirror) and the reference-style driver over the ABI
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}",
          f"-I{rocm}/include", os.path.join(HOST, "graph.cpp"), os.path.join(HOST, "main.cpp"),
          "-o", tmp, f"-L{HERE}", "-lrte", f"-L{rocm}/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{rocm}/lib"])
    os.replace(tmp, DRIVER)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
