"""Build recipe for libgooseberry.so (HIP kernels for gfx950 + host C) and
the `gooseberry` command-line program, in-tree under gooseberry_amd/lib and
gooseberry_amd/bin.

Usage: python -m gooseberry_amd.build   (or python gooseberry_amd/build.py)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "lib", "obj")
LIB = os.path.join(PKG, "lib", "libgooseberry.so")
BIN = os.path.join(PKG, "bin", "gooseberry")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_SOURCES = ["gb_runtime.hip", "gb_index.hip", "gb_spmv.hip", "gb_cg.hip",
               "gb_batch.hip", "gb_gen.hip", "gb_pagerank.hip", "gb_spgemm.hip",
               "gb_ingest.hip", "gb_cm.hip"]
C_SOURCES = ["gb_api.c", "gb_io.c", "gb_gen.c", "gb_batch_api.c", "gb_tune.c",
             "gb_ingest_api.c"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # every product and sum rounds separately, as in the reference's
    # x86-64 build (no FMA contraction)
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
C_FLAGS = ["-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-ffp-contract=off"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd),
                                                          p.stdout))
    return p.stdout


def _compile(src):
    base = os.path.basename(src)
    out = os.path.join(OBJ, base + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-I", INC, "-c", src, "-o", out]
    else:
        cmd = [CC] + C_FLAGS + ["-I", INC, "-c", src, "-o", out]
    _run(cmd)
    return out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = [os.path.join(CSRC, "hip", s) for s in HIP_SOURCES]
    srcs += [os.path.join(CSRC, "host", s) for s in C_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
         + ["-lpthread", "-lm"])
    os.replace(tmp, LIB)
    cli = os.path.join(CSRC, "host", "gb_cli.c")
    _run([CC] + C_FLAGS + ["-I", INC, cli, "-o", BIN,
                           "-L", os.path.dirname(LIB), "-lgooseberry",
                           "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    if verbose:
        print("built", LIB)
        print("built", BIN)
    return LIB


def build_oracle(verbose=True):
    """Build the oracle (test infrastructure) with its own Makefile."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built", os.path.join(ROOT, "oracle", "_build", "liboracle.so"))


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
