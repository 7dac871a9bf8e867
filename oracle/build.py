"""Build recipe for the CPU oracle (test infrastructure, see falm_oracle.cpp).

The reference's time step (CUDA + MPI) cannot be compiled in this image, so
the oracle of the step is the FP64 restatement only.  Two header-only parts
of the reference do compile on their own: its Mesher
(src/mesher/mesher.hpp) and its vendored JSON library
(src/nlohmann/json.hpp).  When the reference tree is present they are built
from their sources where they lie, with the driver ref_tools.cpp, into
oracle/_ref/ref_tools (kept out of git; the checks that use it run on the
CPU only).

    python oracle/build.py        # -> oracle/liboracle.so [+ oracle/_ref/ref_tools]
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


REFERENCE_SRC = "/root/reference/src"
REF_DIR = os.path.join(HERE, "_ref")
REF_TOOLS = os.path.join(REF_DIR, "ref_tools")


def build_ref(verbose: bool = False):
    """oracle/_ref/ref_tools from the reference's Mesher and JSON library;
    None when the reference tree is absent."""
    if not os.path.exists(os.path.join(REFERENCE_SRC, "mesher", "mesher.hpp")):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", REFERENCE_SRC,
           os.path.join(HERE, "ref_tools.cpp"), "-o", REF_TOOLS + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(REF_TOOLS + ".tmp", REF_TOOLS)
    return REF_TOOLS


if __name__ == "__main__":
    build(verbose=True)
    build_ref(verbose=True)
    sys.exit(0)
