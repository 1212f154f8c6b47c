"""Builds libknn_mi355x.so in-tree with hipcc for gfx950.

Each .hip translation unit is compiled to an object in parallel, then linked
into one shared library next to this file.  Flags:
  -ffp-contract=off      no FMA contraction: distances are bit-exact to the
                         reference's sequential FP32 arithmetic;
  -packed-fp32-ops off   no packed FP32 instructions are generated.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libknn_mi355x.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KNN_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]

SOURCES = [
    "knn_prims.hip", "knn_build.hip", "knn_bf.hip", "knn_bf16.hip", "knn_kd_query.hip",
    "knn_ball_query.hip", "knn_kd_packet.hip", "knn_ball_packet.hip", "knn_kd_packet_pad.hip",
    "knn_ball_packet_pad.hip", "knn_tree_host.hip", "knn_learned.hip", "knn_capi.hip",
    "knn_batch.hip", "knn_serialize.hip", "knn_noalloc.hip", "knn_tuned.hip",
]

TUNED_JSON = os.path.join(HERE, "tuned_options.json")
TUNED_INC = os.path.join(CSRC, "knn_tuned_table.inc")
_CLASSES = {"brute": 0, "kd": 1, "ball": 2, "bf16": 3}


def gen_tuned_table(src: str = TUNED_JSON, dst: str = TUNED_INC) -> int:
    """Writes the C++ form of the autotuner's table (one TunedEntry per key
    "kind:D<d>:K<k>:N2^<b>"); returns the entry count."""
    import json
    table = {}
    if os.path.exists(src):
        with open(src) as f:
            table = json.load(f)
    rows = []
    for key in sorted(table):
        kind, d, k, nb = key.split(":")
        best = table[key]["best"]
        block = int(best.get("BLOCK_SIZE", 0) or 0)
        delta = int(best.get("leaf_span_delta", 0) or 0)
        tile = int(best.get("tile", 0) or 0)
        rows.append(f"    {{{_CLASSES[kind]}, {int(nb[3:])}, {int(d[1:])}, {int(k[1:])}, {block}, {delta}, {tile}}},"
                    f"  // {key}: {table[key].get('ms', 0):.4g} ms")
    body = "\n".join(rows) if rows else "    {0, 0, 0, 0, 0, 0, 0},  // no tuned entries"
    text = ("// Generated from tuned_options.json by cudaexactknn_amd/build.py; do not edit.\n"
            "// {cls, log2n, dims, k, block, span_delta, tile}\n"
            f"const TunedEntry kTunedTable[] = {{\n{body}\n}};\n"
            f"const size_t kTunedCount = {len(rows)};\n")
    old = open(dst).read() if os.path.exists(dst) else None
    if old != text:
        with open(dst, "w") as f:
            f.write(text)
    return len(rows)


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


ROOT = os.path.dirname(HERE)
CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "functionality_tests.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "functionality_tests")


def build_cpp_test() -> str:
    """Compiles the C++ drop-in test against include/KnnTrees.hpp and the library."""
    cmd = [HIPCC, *COMMON, "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC, "-L", HERE, "-lknn_mi355x",
           "-Wl,-rpath,$ORIGIN/../../cudaexactknn_amd", "-o", CPP_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{r.stdout}\n{r.stderr}")
    return CPP_TEST_BIN


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    gen_tuned_table()
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    build_cpp_test()
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
