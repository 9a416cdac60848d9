"""Builds the gfx950 HIP library libbnn_flip.so in-tree (hipcc, no JIT cache)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "flip_solver.hip")
OUT = os.path.join(HERE, "libbnn_flip.so")
DEPS = sorted(os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc")) if f.endswith((".hip", ".hpp"))) + [
    os.path.join(os.path.dirname(HERE), "include", "bnn_flip.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
    # exact-order float arithmetic (bit-exact cell indices and fluid SDF)
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


CLI_SRC = os.path.join(HERE, "csrc", "bnn_flip_cli.cpp")
CLI_OUT = os.path.join(HERE, "bnn_flip_cli")


def build(force=True, verbose=False):
    if not force and os.path.exists(OUT) and all(os.path.getmtime(OUT) >= os.path.getmtime(d) for d in DEPS):
        return OUT
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    # the command-line driver links the library (rpath = its own directory)
    cli = ["g++", "-std=c++17", "-O2", "-o", CLI_OUT + ".tmp", CLI_SRC, "-L" + HERE, "-lbnn_flip",
           "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cli), flush=True)
    subprocess.check_call(cli)
    os.replace(CLI_OUT + ".tmp", CLI_OUT)
    return OUT


if __name__ == "__main__":
    build(force=True, verbose=True)
    print(OUT)
    sys.exit(0)
