"""The C++ wind-tunnel application (falm_amd/examples/windtunnel.cpp, written
against include/falm/*.h the way app/windtunnel/main.cpp is written against
the reference's classes) run as a child process on the GPU, against the
oracle's wind-tunnel case with the same settings: no-slip z- wall
(app/windtunnel/bcdevcall.cu:82-97), steady inflow, no rotor."""
import os
import subprocess

import numpy as np
import pytest

from falm_amd import _abi as abi
from falm_amd import cases
from oracle import pyoracle as O
from tests.helpers import interior_mask, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "falm_amd", "lib", "falm_windtunnel")


def test_windtunnel_app_matches_oracle(gpu, tmp_path):
    if not os.path.exists(APP):
        pytest.fail("falm_windtunnel is not built (python -m falm_amd.build)")
    n, h, re, u_in, steps = (40, 16, 12), 1.0 / 8, 2.5e5, 1.0, 4
    dt = 0.3 * h
    out = tmp_path / "wt.bin"
    r = subprocess.run([APP, str(n[0]), str(n[1]), str(n[2]), repr(h), repr(re), repr(dt), str(steps), repr(u_in),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == steps

    setup = cases.windtunnel(n=n, cells_per_d=8, re=re, inlet_intensity=0.0, with_rotor=False, tol=1e-12,
                             pc_maxit=2, pc_relax=1.5)
    setup.struct.zminus_slip = 0
    ref = O.Case(setup)
    info = ref.step(steps)

    dom = ref.dom
    data = np.fromfile(out, dtype=np.float64)
    assert data.size == 4 * dom.ncell
    u, p = data[:3 * dom.ncell], data[3 * dom.ncell:]
    assert rel_err(u, ref.get(abi.FIELD_U), interior_mask(dom, 3)) < 1e-10
    assert rel_err(p, ref.get(abi.FIELD_P), interior_mask(dom)) < 1e-10
    # the divergence printed for the last step (5 significant digits)
    div_app = float(lines[-1].split(",")[1])
    assert abs(div_app - info.div) <= 1e-4 * abs(info.div) + 1e-300
