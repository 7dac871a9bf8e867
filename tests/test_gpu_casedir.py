"""A reference-format case directory run on the GPU.

tests/golden/casedir holds a small wind-turbine case in the reference's
formats: setup.json, a stretched node list converted by the Mesher, a
turbine block with two rotors (one pitching, one rolled with its base x read
from a file) and the blade table.  Two drivers run it:

  falm_nedo (falm_amd/examples/nedo.cpp), written against include/falm the way
      app/nedo/main.cpp is written against the reference's classes, which
      writes snapshots, time averages, the index file and the power/thrust
      CSV; each output is checked against the oracle stepped through the
      same 28 steps;
  falmio.case_from_setup + runtime.Case (the device-resident step driver),
      in FP64 against the oracle and in mixed precision against FP64.
"""
import csv
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from falm_amd import _abi as abi
from falm_amd import falmio
from oracle import pyoracle as O
from tests.helpers import interior_mask, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASE = os.path.join(ROOT, "tests", "golden", "casedir")
CV_GOLDEN = os.path.join(ROOT, "tests", "golden", "io", "casedir_innerNode_gc2.cv")
APP = os.path.join(ROOT, "falm_amd", "lib", "falm_nedo")


def _read(path):
    with open(path) as f:
        return f.read()


def _oracle_run(setup, nsteps, on_step):
    ref = O.Case(setup)
    dt = setup.struct.dt
    for it in range(1, nsteps + 1):
        for tid, (b, v) in falmio.base_motion(setup, it * dt).items():
            ref.set_base(tid, b, v)
        info = ref.step(1)
        on_step(it, ref, info)
    return ref


@pytest.mark.parametrize("variant", ["nedo", "windtunnel"])
def test_nedo_app_matches_oracle(gpu, tmp_path, variant):
    """nedo: the case directory as committed (two rotors, RB-SOR
    preconditioner).  windtunnel: the same directory turned into an
    app/windtunnel setup: no turbine block (the inlet then takes the
    inflow velocity, here 0.8) and the Jacobi preconditioner of the
    reference's shipped setups, 8 iterations per step."""
    if not os.path.exists(APP):
        pytest.fail("falm_nedo is not built (python -m falm_amd.build)")
    case = tmp_path / "case"
    shutil.copytree(CASE, case)
    turbine = variant == "nedo"
    if not turbine:
        cfg = json.loads(_read(str(case / "setup.json")))
        del cfg["turbine"]
        cfg["inflow"]["velocity"] = 0.8
        ls = cfg["solver"]["linearSolver"]
        ls["preconditioner"] = {"type": "Jacobi", "iteration": 2}
        # a fixed 8 iterations per step on both sides: the weaker
        # preconditioner reaches a tight tolerance only after many iterations,
        # where the two sides may stop a few iterations apart
        ls["tolerance"], ls["iteration"] = 1e-30, 8
        (case / "setup.json").write_text(json.dumps(cfg, indent=4))
    print(f"falm_nedo on tests/golden/casedir ({variant}, 28 steps)", flush=True)
    r = subprocess.run([APP, "setup.json"], cwd=str(case), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    steps = [ln for ln in r.stdout.splitlines() if ln.strip() and ln.split()[0].isdigit() and "," in ln]
    assert len(steps) == 28

    # the control-volume file is the reference Mesher's, byte for byte;
    # the copies of FalmCore::computation_init (src/falm.h:356-360)
    assert _read(str(case / "controlVolume.txt")) == _read(CV_GOLDEN)
    assert _read(str(case / "run.cv")) == _read(CV_GOLDEN)
    assert json.loads(_read(str(case / "run_setup.json"))) == json.loads(_read(str(case / "setup.json")))
    idx = falmio.read_index_file(str(case / "run.json"))
    assert idx["global"] == [40, 20, 20] and idx["guidePoint"] == 2 and idx["dataType"] == "float64"
    assert idx["outputSteps"] == [{"time": 0.01 * 10, "step": 10, "timeAvg": 1},
                                  {"time": 0.01 * 20, "step": 20, "timeAvg": 1}]
    assert _read(str(case / "run_endmsg.txt")) == "Falm/Main ends normally."

    rows = []
    if turbine:
        with open(case / "run_PT.csv") as f:
            rows = list(csv.reader(f))
        assert rows[0] == ["t", "P0", "T0", "P1", "T1"] and len(rows) == 29
    else:
        assert not (case / "run_PT.csv").exists()

    setup = falmio.case_from_setup(str(case / "setup.json"), build=False)
    dt, mask3 = setup.struct.dt, None
    state = {"tu": None, "tp": None}
    tip = [setup.alm.turbines[t].tip_rate for t in range(2)] if turbine else []

    def check(it, ref, info):
        nonlocal mask3
        dom = ref.dom
        if mask3 is None:
            mask3 = interior_mask(dom, 3)
        u, p = ref.get(abi.FIELD_U), ref.get(abi.FIELD_P)
        if it >= 5:  # FalmCore::TAvg (src/falm.h:433-444), averaging from step 5
            b = 1.0 / (it - 5 + 1)
            a = 1.0 - b
            state["tu"] = b * u if state["tu"] is None else a * state["tu"] + b * u
            state["tp"] = b * p if state["tp"] is None else a * state["tp"] + b * p
        if turbine:
            _, _, tq, th = ref.alm_get()
            row = [float(v) for v in rows[it]]
            assert row[0] == pytest.approx(it * dt, rel=1e-5)
            for t in range(2):  # 6 significant digits in the CSV
                assert row[1 + 2 * t] == pytest.approx(tq[t] * tip[t], rel=2e-5, abs=1e-12)
                assert row[2 + 2 * t] == pytest.approx(th[t], rel=2e-5, abs=1e-12)
        if it in (10, 20):
            for tavg, (uu, pp) in ((False, (u, p)), (True, (state["tu"], state["tp"]))):
                hd, data = falmio.read_vector_file(falmio.snapshot_path(str(case), "run", it, tavg=tavg))
                assert (hd["step"], hd["n"], hd["gc"], hd["dsize"]) == (it, 4, 2, 8)
                assert hd["time"] == it * dt
                assert (hd["imax"], hd["jmax"], hd["kmax"]) == tuple(dom.shape)
                flat = data.reshape(4, -1)
                assert rel_err(flat[:3].ravel(), uu, mask3) < 1e-10, (it, tavg)
                assert rel_err(flat[3], pp, interior_mask(dom)) < 1e-10, (it, tavg)
        if it == 28:
            div_app = float(steps[-1].split(",")[1])
            assert abs(div_app - info.div) <= 1e-4 * abs(info.div) + 1e-300

    ref = _oracle_run(setup, 28, check)
    if turbine:  # both rotors load the flow (turbine 1 turns backwards, tipRate -3)
        last = [float(v) for v in rows[-1]]
        assert last[2] > 0 and last[4] != 0
    else:  # the inlet ghost cells hold the inflow velocity
        u = ref.get(abi.FIELD_U)
        nx, ny, nz = ref.dom.shape
        assert np.all(u[:nx * ny * nz].reshape(nz, ny, nx)[2:-2, 2:-2, :2] == 0.8)


def test_case_directory_on_the_step_driver(gpu, tmp_path):
    from falm_amd import runtime
    case = tmp_path / "case"
    shutil.copytree(CASE, case)
    setup = falmio.case_from_setup(str(case / "setup.json"))  # converts the node list first
    assert _read(str(case / "controlVolume.txt")) == _read(CV_GOLDEN)
    dev = runtime.Case(setup)
    outs = []

    def step_dev(it):
        for tid, (b, v) in falmio.base_motion(setup, it * setup.struct.dt).items():
            dev.set_base(tid, b, v)
        return dev.step(1)

    ref_fields = {}

    def on_step(it, ref, info):
        step_dev(it)
        if it in (3, 6):
            ref_fields[it] = (ref.get(abi.FIELD_U), ref.get(abi.FIELD_P), dev.get(abi.FIELD_U), dev.get(abi.FIELD_P))

    print("step driver FP64 on tests/golden/casedir (6 steps)", flush=True)
    ref = _oracle_run(setup, 6, on_step)
    m3, m1 = interior_mask(ref.dom, 3), interior_mask(ref.dom)
    for it, (ur, pr, ud, pd) in ref_fields.items():
        assert rel_err(ud, ur, m3) < 1e-10, it
        assert rel_err(pd, pr, m1) < 1e-9, it
    outs.append(ref_fields[6][2])
    dev.close()

    print("step driver mixed precision on tests/golden/casedir (6 steps)", flush=True)
    mixed = falmio.case_from_setup(str(case / "setup.json"), precision=abi.CASE_MIXED, build=False)
    dm = runtime.Case(mixed)
    for it in range(1, 7):
        for tid, (b, v) in falmio.base_motion(mixed, it * mixed.struct.dt).items():
            dm.set_base(tid, b, v)
        dm.step(1)
    um = dm.get(abi.FIELD_U)
    dm.close()
    # FP32 stencils with FP64 defect correction: velocity within 1e-4 of FP64
    assert rel_err(um, outs[0], m3) < 1e-4
