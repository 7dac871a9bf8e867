"""Rotor power and its gradient with respect to blade pitch and section twist
as a ``torch.autograd.Function`` over the device adjoint
(``falm_case_adjoint`` in the C ABI).

The objective is the mean rotor power P = sum_t Omega_t * torque_t over the
next ``nsteps`` steps of a device-resident wind-tunnel case, the power
``AlmHandler::writePowerThrust`` reports (src/alm/alm.h:126-138).  The
reference has no adjoint; this module has no counterpart there.

Pitch (per turbine and blade) and the per-section twist offsets are in
degrees, as in the rotor configuration; gradients are per degree.  The case
is left in the state it had: every call integrates the same window.
"""
import numpy as np
import torch

from . import runtime as rt


def _apply(case, pitch, twist):
    p = pitch.detach().to("cpu", torch.float64).numpy()
    for t in range(p.shape[0]):
        case.set_pitch(t, p[t])
    case.set_twist_offset(twist.detach().to("cpu", torch.float64).numpy())


class RotorPower(torch.autograd.Function):
    """J(pitch [n_turbine, n_blade], twist [n_ap_per_blade]) -> scalar.

    With want_grad false the forward pass runs the window forward only;
    otherwise it runs the checkpointed adjoint, which returns J and its
    gradient together, and the backward pass scales that gradient.
    """

    @staticmethod
    def forward(ctx, pitch, twist, case, nsteps, ckpt, adj_rtol, want_grad):
        _apply(case, pitch, twist)
        if not want_grad:
            return torch.tensor(case.window(nsteps), dtype=torch.float64)
        J, gp, gt, _ = case.adjoint(nsteps, ckpt, adj_rtol)
        nb = pitch.shape[1]
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(gp[:, :nb])).to(pitch.device),
                              torch.from_numpy(gt).to(twist.device))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        if not ctx.saved_tensors:
            raise RuntimeError("rotor_power: the forward pass ran without a gradient request")
        gp, gt = ctx.saved_tensors
        return gJ * gp, gJ * gt, None, None, None, None, None


def rotor_power(case, pitch, twist, nsteps, ckpt=20, adj_rtol=1e-10):
    """Mean rotor power of ``case`` (an ``rt.Case`` with a rotor) over the next
    ``nsteps`` steps at the given pitch [n_turbine, n_blade] and twist
    offsets [n_ap_per_blade] (degrees), differentiable in both."""
    if not isinstance(case, rt.Case):
        raise TypeError("rotor_power needs a falm_amd.runtime.Case")
    want = torch.is_grad_enabled() and (pitch.requires_grad or twist.requires_grad)
    return RotorPower.apply(pitch, twist, case, int(nsteps), int(ckpt), float(adj_rtol), want)
