"""The device adjoint of the window-mean rotor power (no reference
counterpart: the reference has no adjoint) against the oracle's adjoint on
the same case, central differences of the device forward (h = 1e-4 rad),
checkpoint-spacing invariance, state restoration, and the
torch.autograd.Function wrapper checked with torch's gradcheck."""
import numpy as np
import pytest

from falm_amd import _abi as abi
from falm_amd import cases
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _setup(motion=False, sgs=abi.SGS_SMAGORINSKY, precision=abi.CASE_F64):
    kw = dict(n_ap_per_blade=4)
    if motion:
        kw.update(angle_type=2, motion=(np.radians(4.0), 3.0, 0.3), surge=(0.1, 5.0, 0.2))
    setup = cases.windtunnel(n=(24, 12, 12), cells_per_d=6, hub_x_d=1.5, re=1e3, tol=1e-13, rotor_kw=kw,
                             precision=precision, pc_maxit=2, pc_relax=1.5)
    setup.struct.sgs_model = sgs
    return setup


def _gnorm(gp, gt):
    return np.sqrt(gp[0, :3] @ gp[0, :3] + gt @ gt)


@pytest.mark.parametrize("window,ckpt,spin,motion,sgs", [(6, 2, 4, False, abi.SGS_SMAGORINSKY),
                                                         (8, 3, 6, True, abi.SGS_SMAGORINSKY),
                                                         (5, 5, 3, False, abi.SGS_EMPTY)])
def test_adjoint_matches_oracle(gpu, window, ckpt, spin, motion, sgs):
    setup = _setup(motion, sgs)
    print(f"adjoint vs oracle: window={window} ckpt={ckpt} spin={spin} motion={motion} sgs={sgs}", flush=True)
    ref, dev = O.Case(setup), gpu.Case(setup)
    ref.step(spin)
    dev.step(spin)
    Jr, gpr, gtr = ref.adjoint(window, ckpt, 1e-13)
    Jd, gpd, gtd, nbytes = dev.adjoint(window, ckpt, 1e-13)
    assert Jr > 0 and nbytes > 0
    assert abs(Jd - Jr) <= 1e-10 * abs(Jr), (Jd, Jr)
    assert abs(dev.window(window) - Jd) <= 1e-12 * abs(Jd)
    g = _gnorm(gpr, gtr)
    assert g > 0
    assert np.abs(gpd[0, :3] - gpr[0, :3]).max() <= 1e-8 * g, (gpd[0, :3], gpr[0, :3])
    assert np.abs(gtd - gtr).max() <= 1e-8 * g, (gtd, gtr)
    # unused blade slots carry no gradient
    assert np.all(gpd[0, 3:] == 0.0)


def test_adjoint_matches_central_differences_and_restores_state(gpu):
    setup = _setup(True)
    dev = gpu.Case(setup)
    dev.step(6)
    u0 = dev.get(abi.FIELD_U)
    W = 8
    J, gp, gt, _ = dev.adjoint(W, 3, 1e-13)
    assert np.array_equal(dev.get(abi.FIELD_U), u0)
    rng = np.random.default_rng(2024)
    h = np.degrees(1e-4)  # 1e-4 rad in the API's degrees
    g = _gnorm(gp, gt)
    for _ in range(4):
        vp, vt = rng.standard_normal(3), rng.standard_normal(4)
        nrm = np.sqrt(vp @ vp + vt @ vt)
        vp, vt = vp / nrm, vt / nrm

        def Jat(e):
            dev.set_pitch(0, e * vp)
            dev.set_twist_offset(e * vt)
            return dev.window(W)

        fd = (Jat(h) - Jat(-h)) / (2 * h)
        Jat(0.0)
        ad = gp[0, :3] @ vp + gt @ vt
        assert abs(fd - ad) <= 1e-5 * g, (fd, ad, g)
    # the window left the state and the step counter as they were
    assert np.array_equal(dev.get(abi.FIELD_U), u0)
    assert dev.step(1).step == 7
    # checkpoint spacing does not change the result
    dev2 = gpu.Case(setup)
    dev2.step(6)
    J1, gp1, gt1, b1 = dev2.adjoint(W, 1, 1e-13)
    J8, gp8, gt8, b8 = dev2.adjoint(W, 8, 1e-13)
    assert abs(J1 - J) <= 1e-13 * abs(J) and abs(J8 - J) <= 1e-13 * abs(J)
    for a, b in ((gp1, gp), (gp8, gp), (gt1, gt), (gt8, gt)):
        assert np.abs(a - b).max() <= 1e-10 * g


def test_adjoint_rejects_unsupported_cases(gpu):
    dev = gpu.Case(_setup(precision=abi.CASE_MIXED))
    dev.step(1)
    assert dev.window(2) > 0  # the forward window works in any precision
    with pytest.raises(gpu.FalmError) as e:
        dev.adjoint(2, 1)
    assert e.value.status == abi.FALM_ERR_UNSUPPORTED


def test_autograd_function_gradcheck(gpu):
    import torch
    from falm_amd.adjoint import rotor_power
    dev = gpu.Case(_setup())
    dev.step(4)
    W = 5
    pitch = torch.zeros(1, 3, dtype=torch.float64, requires_grad=True)
    twist = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    J = rotor_power(dev, pitch, twist, W, ckpt=2, adj_rtol=1e-13)
    J.backward()
    Jw, gp, gt, _ = dev.adjoint(W, 2, 1e-13)
    assert abs(J.item() - Jw) <= 1e-13 * abs(Jw)
    g = _gnorm(gp, gt)
    # two adjoint runs differ only in the order of their atomic sums
    assert np.abs(pitch.grad.numpy() - gp[:, :3]).max() <= 1e-10 * g
    assert np.abs(twist.grad.numpy() - gt).max() <= 1e-10 * g
    # torch's own central differences through the same Function
    p0 = torch.tensor([[0.5, -0.3, 0.2]], dtype=torch.float64, requires_grad=True)
    t0 = torch.tensor([0.1, 0.0, -0.2, 0.3], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda p, t: rotor_power(dev, p, t, W, 2, 1e-13), (p0, t0),
                                    eps=float(np.degrees(1e-4)), atol=1e-5 * g, rtol=1e-4)
    # without a gradient request only the forward window runs
    with torch.no_grad():
        Jn = rotor_power(dev, pitch.detach(), twist.detach(), W)
    assert abs(Jn.item() - Jw) <= 1e-12 * abs(Jw)
