"""falm_dot_noalloc (the allocation-free variant of falm_dot) against falm_dot
and the oracle, on two calls in a row, the second on a smaller domain: the
fixed per-GPU workspace keeps the first call's partial sums, which the second
must not read."""
import ctypes as C

import numpy as np
import pytest

from falm_amd import _abi as abi
from oracle import pyoracle as O
from tests.helpers import domain, to_dev

pytestmark = pytest.mark.gpu


def _pair(gpu, dom, prec, seed):
    import torch
    rng = np.random.default_rng(seed)
    a, b = rng.standard_normal(dom.ncell), rng.standard_normal(dom.ncell)
    dt = torch.float64 if prec == abi.F64 else torch.float32
    da, db = to_dev(a, dt), to_dev(b, dt)
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    gpu.call("falm_dot_noalloc", C.byref(dom), prec, gpu.ptr(da), gpu.ptr(db), gpu.ptr(out), None)
    torch.cuda.synchronize()
    got = float(out.cpu()[0])
    host = C.c_double()
    gpu.call("falm_dot", C.byref(dom), prec, gpu.ptr(da), gpu.ptr(db), C.byref(host), None)
    # the oracle sees the operands at the precision the kernel reads them
    a_r, b_r = da.double().cpu().numpy(), db.double().cpu().numpy()
    return got, host.value, O.dot(dom, a_r, b_r), float(np.sum(np.abs(a_r * b_r)))


@pytest.mark.parametrize("prec", [abi.F64, abi.F32])
def test_dot_noalloc_matches_dot_on_two_calls(gpu, prec):
    # first call: 3 x 3 x 3 partial column tiles; second: one small tile
    for n, seed in (((150, 20, 37), 1), ((13, 10, 8), 2)):
        dom = domain(n)
        got, host, ref, scale = _pair(gpu, dom, prec, seed)
        assert got == host, (n, got, host)  # same kernels, same summation order
        assert abs(got - ref) <= 1e-13 * scale, (n, got, ref)


def test_dot_noalloc_rejects_bad_arguments(gpu):
    import torch
    dom = domain((8, 8, 8))
    x = torch.zeros(dom.ncell, dtype=torch.float64, device="cuda")
    st = gpu.lib().falm_dot_noalloc(C.byref(dom), abi.F64, gpu.ptr(x), gpu.ptr(x), None, None)
    assert st == abi.FALM_ERR_INVALID
