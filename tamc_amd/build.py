"""Build libtamc_gpu.so for gfx950 in-tree (tamc_amd/lib/), and the drop-in `tamc` CLI.

Steps:
  1. compile and run expf_table_gen (host) -> csrc/generated/expf_exceptions.inc
  2. hipcc every csrc/*.hip for gfx950 (packed-FP32 instructions disabled, no FP contraction)
  3. link the shared library and the CLI

Run ``python -m tamc_amd.build`` or call ``build()``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "generated")
OBJ = os.path.join(HERE, "_obj")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libtamc_gpu.so")
CLI = os.path.join(LIB_DIR, "tamc")
CLI_PT_OPT = os.path.join(LIB_DIR, "tamc-pt-opt")
CLI_PERCOLATE = os.path.join(LIB_DIR, "percolate")
ROOT = os.path.dirname(HERE)

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
ARCH = os.environ.get("TAMC_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = ["pt_kernels.hip", "pt_kernels_any.hip", "instance.hip", "pt_runtime.hip", "lockstep.hip", "elementary.hip", "sa.hip",
           "writers.hip", "server.hip", "gla.hip", "pt_opt.hip", "percolation.hip"]


def _run(cmd, quiet=False):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    if not quiet and r.stderr.strip():
        sys.stderr.write(r.stderr)
    return r


def gen_expf_table():
    os.makedirs(GEN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    exe = os.path.join(OBJ, "expf_table_gen")
    _run([CLANGXX, "-O2", "-std=c++17", "-ffp-contract=off", "-march=x86-64-v3", "-o", exe,
          os.path.join(CSRC, "expf_table_gen.cpp"), "-lpthread"])
    out = os.path.join(GEN, "expf_exceptions.inc")
    _run([exe, out], quiet=True)
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj], quiet=True)
    return obj


def build(verbose=True):
    gen_expf_table()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli_main.cpp")
    if os.path.exists(cli_src):
        # the command-line front end is plain host C++ over the C ABI
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", CLI, cli_src,
              "-L", LIB_DIR, "-ltamc_gpu", "-Wl,-rpath,$ORIGIN"])
        # tamc-pt-opt and percolate are the same program under other names (it dispatches on argv[0])
        shutil.copy2(CLI, CLI_PT_OPT)
        shutil.copy2(CLI, CLI_PERCOLATE)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
