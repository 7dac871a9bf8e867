"""Floating tandem-rotor farm segment: two rotors in line, the upstream one
on a platform in prescribed surge and pitch, yawed, with a static part of
the platform pitch; 8 % inlet turbulence.

  * FP64 steps against the CPU oracle on a small domain (1e-10), mixed
    precision against FP64 on the same domain;
  * the adjoint gradient of the summed power of both rotors with respect to
    each rotor's yaw and the platform's pitch offset against central
    differences of the GPU forward and of the oracle's forward window
    (h = 1e-4 rad), with the pitch/twist gradient checked the same way on the
    yawed rotor;
  * the 1024 x 256 x 256 segment in mixed precision for a few steps
    (finite, converged, both rotors loaded, divergence at the solver's
    level).

No reference counterpart for yaw, the platform offset or the adjoint: the
reference turns a rotor about one Euler axis (src/alm/almDevCall.cu:43-129,
src/alm/turbineHandler.h:152-190)."""
import numpy as np
import pytest

from falm_amd import _abi as abi
from falm_amd import cases
from oracle import pyoracle as O
from tests.helpers import interior_mask, rel_err

pytestmark = pytest.mark.gpu

YAW, OFFSET = 12.0, 2.5


def _small(precision=abi.CASE_F64, tol=1e-13, maxit=1000, yaw=YAW, offset=OFFSET):
    # D = 6 cells, rotors at 1.5 D and 4.5 D, a fast platform period so that
    # the short windows see the motion
    return cases.farm_segment(n=(40, 12, 12), cells_per_d=6, spacing_d=3.0, hub_x_d=1.5, re=1e3, cfl=0.3,
                              precision=precision, tol=tol, maxit=maxit, pc_maxit=2, pc_relax=1.5,
                              period_revs=2.0, yaw_deg=yaw, angle_offset_deg=offset,
                              rotor_kw=dict(n_ap_per_blade=4))


def test_farm_steps_match_oracle(gpu):
    """FP64 steps of the yawed, surging, pitching upstream rotor and the fixed
    downstream rotor against the oracle: fields, point forces and rotor
    loads at 1e-10 (a fixed iteration count, so both solvers stop alike)."""
    setup = _small(tol=1e-30, maxit=6)
    ref, dev = O.Case(setup), gpu.Case(setup)
    for _ in range(5):
        ir, ig = ref.step(1), dev.step(1)
        assert ir.solver_it == ig.solver_it
    m3 = interior_mask(ref.dom, 3)
    for field, mask in ((abi.FIELD_U, m3), (abi.FIELD_P, interior_mask(ref.dom)), (abi.FIELD_NUT, None),
                        (abi.FIELD_UU, None), (abi.FIELD_FF, None)):
        err = rel_err(dev.get(field), ref.get(field), mask)
        assert err < 1e-10, (field, err)
    xyz, f, tq, th = dev.alm_get()
    xyz_r, f_r, tq_r, th_r = ref.alm_get()
    assert rel_err(xyz, xyz_r) < 1e-14
    assert rel_err(f, f_r) < 1e-10 and rel_err(tq, tq_r) < 1e-10 and rel_err(th, th_r) < 1e-10
    assert np.all(th > 0), th
    # the yaw turns the upstream rotor's points about the vertical axis: the
    # same case without yaw puts them elsewhere
    plain = O.Case(_small(tol=1e-30, maxit=6, yaw=0.0))
    plain.step(5)
    assert rel_err(plain.alm_get()[0][:12], xyz_r[:12]) > 1e-2


def test_farm_mixed_tracks_fp64(gpu):
    s64, smx = _small(tol=1e-10), _small(precision=abi.CASE_MIXED, tol=1e-10)
    a, b = gpu.Case(s64), gpu.Case(smx)
    for _ in range(8):
        ia, ib = a.step(1), b.step(1)
        assert ia.status == abi.FALM_OK and ib.status == abi.FALM_OK and ib.solver_err <= 1e-10
    m3 = interior_mask(a.dom, 3)
    ua, ub = a.get(abi.FIELD_U)[m3], b.get(abi.FIELD_U)[m3]
    ru = np.linalg.norm(ua - ub) / np.linalg.norm(ua)
    print(f"farm mixed vs fp64 after 8 steps: u rel L2 {ru:.3e}")
    assert ru < 1e-5


def _fd(case, W, turbine, which, h):
    """Central difference of the window-mean summed power in yaw or offset."""
    base = {0: (YAW, OFFSET), 1: (0.0, 0.0)}[turbine]

    def J(e):
        yaw, off = base
        if which == "yaw":
            yaw += e
        else:
            off += e
        case.set_attitude(turbine, yaw, off)
        return case.window(W)

    d = (J(h) - J(-h)) / (2 * h)
    J(0.0)
    return d


def test_attitude_gradient_matches_central_differences(gpu):
    setup = _small()
    dev = gpu.Case(setup)
    dev.step(6)
    u0 = dev.get(abi.FIELD_U)
    W = 8
    J, gp, gt, _ = dev.adjoint(W, 3, 1e-13)
    gy, go = dev.adjoint_attitude()
    assert np.array_equal(dev.get(abi.FIELD_U), u0)
    print(f"J {J:.8g}  dJ/dyaw {gy}  dJ/doffset {go}", flush=True)
    assert go[1] == 0.0  # the downstream rotor has no platform angle
    h = np.degrees(1e-4)
    scale = max(np.abs(gy).max(), abs(go[0]))
    assert scale > 0
    for turbine, which, ad in ((0, "yaw", gy[0]), (0, "offset", go[0]), (1, "yaw", gy[1])):
        fd = _fd(dev, W, turbine, which, h)
        print(f"turbine {turbine} {which}: adjoint {ad:.10e}  central difference {fd:.10e}", flush=True)
        assert abs(fd - ad) <= 1e-5 * scale, (turbine, which, fd, ad)
    # the pitch gradient on the yawed rotor, by the same check
    dev.set_attitude(0, YAW, OFFSET)
    v = np.array([0.6, -0.3, 0.74])

    def Jp(e):
        dev.set_pitch(0, e * v)
        return dev.window(W)

    fd = (Jp(h) - Jp(-h)) / (2 * h)
    Jp(0.0)
    ad = gp[0, :3] @ v
    assert abs(fd - ad) <= 1e-5 * np.abs(gp[0, :3]).max(), (fd, ad)


def test_attitude_gradient_matches_oracle_differences(gpu):
    """The GPU adjoint's attitude gradient against central differences of
    the oracle's forward window on the same case."""
    setup = _small()
    dev, ref = gpu.Case(setup), O.Case(setup)
    dev.step(4)
    ref.step(4)
    W = 6
    dev.adjoint(W, 3, 1e-13)
    gy, go = dev.adjoint_attitude()
    h = np.degrees(1e-4)
    scale = max(np.abs(gy).max(), abs(go[0]))
    for turbine, which, ad in ((0, "yaw", gy[0]), (0, "offset", go[0])):
        fd = _fd(ref, W, turbine, which, h)
        assert abs(fd - ad) <= 1e-5 * scale, (turbine, which, fd, ad)


def test_full_size_farm_segment(gpu):
    """1024 x 256 x 256 cells, two rotors 7 D apart, mixed precision, 3 steps."""
    setup = cases.farm_segment()
    c = gpu.Case(setup)
    info = None
    for _ in range(3):
        info = c.step(1)
        assert info.status == abi.FALM_OK, info.status
        assert info.solver_err <= 1e-6
        # div ~ dt maxdiag |r|/N: 3.1e-3 * 6144 * 1e-6 = 1.9e-5 at most at the tolerance
        assert np.isfinite(info.div) and info.div < 5e-5
    _, f, tq, th = c.alm_get()
    assert np.all(np.isfinite(f)) and np.all(th > 0), th
