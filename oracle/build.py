"""Build recipe for the CPU oracle (test infrastructure, see falm_oracle.cpp).

The reference itself (CUDA + MPI) cannot be compiled in this image, so there
is no oracle/_ref build; the oracle is the FP64 restatement only.

    python oracle/build.py        # -> oracle/liboracle.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "falm_oracle.cpp")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    # x86-64-v3 (AVX2/FMA) keeps the library loadable on any recent host CPU.
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
           "-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
