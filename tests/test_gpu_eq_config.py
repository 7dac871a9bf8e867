"""The FalmEq class API with settings changed between solves
(tests/native/eq_config.cpp) against the oracle's solver with the same
settings.  The reference reads type, tol, maxit, relax and the
preconditioner settings on every Solve (src/FalmEq.h:56-64): a first
PBiCGStab solve, then SOR(), then a tighter tol, then init() with another
preconditioner, then a smaller maxit must each run with their own
settings."""
import os
import subprocess

import numpy as np
import pytest

from falm_amd import _abi as abi
from oracle import pyoracle as O
from tests.helpers import domain

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = (20, 12, 9)


def build_program(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "native", "eq_config.cpp"), "-L", os.path.join(ROOT, "falm_amd", "lib"),
           "-lfalm_amd", "-Wl,-rpath," + os.path.join(ROOT, "falm_amd", "lib"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _system():
    dom = domain(N)
    h = 0.1
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    _, _, g, ja = O.metrics(dom, *cen, *[np.full(dom.shape[m], h) for m in range(3)])
    a, _ = O.poisson_matrix(dom, abi.POISSON_OUTLET_XPLUS, g, ja)
    b = np.random.default_rng(17).standard_normal(dom.ncell) * 1e-2
    return dom, a, b


def test_solver_settings_follow_the_members(gpu, tmp_path):
    exe = build_program(tmp_path / "eq_config")
    dom, a, b = _system()
    a.tofile(tmp_path / "a.bin")
    b.tofile(tmp_path / "b.bin")
    r = subprocess.run([str(exe), *map(str, N), str(tmp_path / "a.bin"), str(tmp_path / "b.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = {}
    for line in r.stdout.splitlines():
        f = line.split()
        if f[0] == "solvers":
            nsolvers = int(f[1])
        else:
            rows[f[0]] = (int(f[1]), float(f[2]))

    def oracle(**kw):
        cfg = abi.solver_config(**kw)
        x, res = np.zeros(dom.ncell), np.zeros(dom.ncell)
        info = O.Solver(dom, cfg).solve(a, x, b, res)
        return info.it, info.err

    bic = dict(type=abi.LS_PBICGSTAB, maxit=1000, relax=1.2, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.5)
    expect = {
        "bicgstab_sor_1e-8": oracle(tol=1e-8, **bic),
        "sor_1e-8": oracle(type=abi.LS_SOR, maxit=1000, tol=1e-8, relax=1.2, pc_type=abi.LS_SOR, pc_maxit=2,
                           pc_relax=1.5),
        "bicgstab_sor_1e-8_again": oracle(tol=1e-8, **bic),
        "bicgstab_sor_1e-11": oracle(tol=1e-11, **bic),
        "bicgstab_jacobi_1e-9": oracle(type=abi.LS_PBICGSTAB, maxit=1000, tol=1e-9, relax=1.2,
                                       pc_type=abi.LS_JACOBI, pc_maxit=3, pc_relax=1.0),
        "bicgstab_jacobi_maxit3": oracle(type=abi.LS_PBICGSTAB, maxit=3, tol=1e-9, relax=1.2,
                                         pc_type=abi.LS_JACOBI, pc_maxit=3, pc_relax=1.0),
    }
    for label, (it_o, err_o) in expect.items():
        it_g, err_g = rows[label]
        if label.startswith("sor") or label.endswith("maxit3"):
            # no dot products in the iterate (SOR) or a fixed count: same
            # iterations, residual norms equal to rounding
            assert it_g == it_o, (label, it_g, it_o)
            assert abs(err_g - err_o) <= 1e-8 * err_o, (label, err_g, err_o)
        else:
            assert abs(it_g - it_o) <= 1, (label, it_g, it_o)
            tol = {"bicgstab_sor_1e-8": 1e-8, "bicgstab_sor_1e-8_again": 1e-8, "bicgstab_sor_1e-11": 1e-11,
                   "bicgstab_jacobi_1e-9": 1e-9}[label]
            assert err_g <= tol, (label, err_g)
    # the settings really changed: the SOR solve and the tighter tolerance
    # take more iterations, maxit 3 stops at 3
    assert rows["sor_1e-8"][0] > rows["bicgstab_sor_1e-8"][0]
    assert rows["bicgstab_sor_1e-11"][0] > rows["bicgstab_sor_1e-8"][0]
    assert rows["bicgstab_jacobi_maxit3"][0] == 3
    assert nsolvers == 5  # one library solver per distinct configuration
