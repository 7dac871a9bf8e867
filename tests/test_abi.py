"""The C ABI without a GPU: the library loads, exports every entry point the
header declares, and the ctypes mirror of the header's structs has the C
layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from falm_amd import _abi as abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "falm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(falm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from falm_amd import runtime
    L = runtime.lib()
    names = declared_functions()
    assert len(names) >= 49
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python bindings cover the same set
    assert set(runtime.EXPORTS) == set(names)


def test_struct_layouts_match_header():
    probe = r"""
#include <stdio.h>
#include <stddef.h>
#include "falm_amd.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(falm_domain) S(falm_face_bc) S(falm_inlet_noise) S(falm_solver_config) S(falm_solve_info)
  S(falm_turbine) S(falm_alm_config) S(falm_case_config) S(falm_step_info)
  printf("off_case_solver %zu\n", offsetof(falm_case_config, solver));
  printf("off_case_alm %zu\n", offsetof(falm_case_config, alm));
  printf("off_alm_cutoff %zu\n", offsetof(falm_alm_config, cutoff));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write(probe)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    for cname, py in abi.ABI_STRUCTS.items():
        assert int(sizes[cname]) == C.sizeof(py), cname
    assert int(sizes["off_case_solver"]) == abi.CaseConfig.solver.offset
    assert int(sizes["off_case_alm"]) == abi.CaseConfig.alm.offset
    assert int(sizes["off_alm_cutoff"]) == abi.AlmConfig.cutoff.offset


def test_status_strings_without_device():
    from falm_amd import runtime
    L = runtime.lib()
    assert L.falm_status_string(abi.FALM_ERR_ALIAS) == b"output aliases an input"
    assert L.falm_status_string(abi.FALM_ERR_HIP + 2) == b"HIP error"


def test_host_only_build_ap_matches_oracle():
    """falm_build_ap is host code: it runs without a GPU."""
    import numpy as np

    from falm_amd import cases
    from oracle import pyoracle as O
    rr, ch, tw, attack, cl, cd = cases.blade_table(0.5, 0.08, 0.02, 20.0, 0.0, n_station=7)
    ref = O.build_ap(rr, ch, tw, cl, cd, 2, 3, 8, 0.5, 0.02)
    from falm_amd import runtime
    nap, nat = 2 * 3 * 8, attack.size
    out = [np.zeros(nap), np.zeros(nap), np.zeros(nap), np.zeros(nap * nat), np.zeros(nap * nat)]
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (rr, ch, tw, cl, cd)]
    runtime.call("falm_build_ap", rr.size, nat, *[v.ctypes.data_as(C.c_void_p) for v in args], 2, 3, 8, 0.5, 0.02,
                 *[v.ctypes.data_as(C.c_void_p) for v in out])
    for a, b in zip(out, ref):  # host compilers may contract a*b+c differently
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))


def test_missing_device_fails_loudly():
    """Without a GPU the library reports an error instead of computing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from falm_amd import runtime
    with pytest.raises(runtime.FalmError):
        runtime.init(0)


def test_cpp_class_api_compiles_and_links():
    """include/falm/*.h (the FalmCFD / FalmEq / FalmMV / Matrix / CPM / Alm
    class API) compiles warning-free and the wind-tunnel application links
    against the library."""
    from falm_amd import runtime
    lib_dir = os.path.dirname(runtime.LIB_PATH)
    for src in (os.path.join(ROOT, "falm_amd", "examples", "windtunnel.cpp"),
                os.path.join(ROOT, "tests", "native", "eq_config.cpp")):
        with tempfile.TemporaryDirectory() as td:
            exe = os.path.join(td, "wt")
            r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I",
                                os.path.join(ROOT, "include"), src, "-L", lib_dir, "-lfalm_amd", "-o", exe],
                               capture_output=True, text=True)
            assert r.returncode == 0, r.stderr


def test_oversized_domains_are_refused_before_any_device_call():
    """The k-marching kernels address a field column with 32-bit byte
    offsets: a double column of 2^29 cells or more is refused with
    FALM_ERR_UNSUPPORTED by the solver and by the step driver, before either
    touches the device."""
    from falm_amd import cases, runtime
    L = runtime.lib()
    dom = abi.Domain.make((1028, 1028, 520), gc=2)  # 5.5e8 cells: 4.4 GB of double
    h = C.c_void_p()
    cfg = abi.solver_config()
    assert L.falm_solver_create(C.byref(dom), C.byref(cfg), C.byref(h)) == abi.FALM_ERR_UNSUPPORTED
    assert not h.value
    setup = cases.windtunnel(n=(1024, 1024, 516), with_rotor=False)
    assert L.falm_case_create(C.byref(setup.struct), C.byref(h)) == abi.FALM_ERR_UNSUPPORTED
    assert b"4 GiB" in L.falm_last_error()
