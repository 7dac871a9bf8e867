"""Build falm_amd/lib/libfalm_amd.so from falm_amd/csrc/*.hip for gfx950.

    python -m falm_amd.build          # or: python falm_amd/build.py

hipcc cross-compiles without a GPU.  Packed-FP32 instruction selection is
disabled for every device compile (-target-feature -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libfalm_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FALM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["runtime.hip", "cfd.hip", "linalg.hip", "solver.hip", "bc.hip", "alm.hip", "case.hip", "adjoint.hip",
           "mesher.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-result"] + DEVICE_FLAGS


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):  # host-only code
        cmd = ["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the device-only feature flag
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(lines))
    if verbose and lines:
        print("\n".join(lines), flush=True)
    return obj


def build(verbose: bool = False, jobs: int = 0) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_examples(verbose)
    return LIB


EXAMPLES = {"falm_windtunnel": "windtunnel.cpp", "falm_nedo": "nedo.cpp"}


def build_examples(verbose: bool = False) -> None:
    """C++ applications written against include/falm/*.h, linked to the library."""
    root = os.path.dirname(HERE)
    for exe, src in EXAMPLES.items():
        out = os.path.join(LIBDIR, exe)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include"),
               os.path.join(HERE, "examples", src), "-L", LIBDIR, "-lfalm_amd", "-Wl,-rpath,$ORIGIN",
               "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
