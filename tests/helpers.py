"""Shared helpers of the test-suite: seeded fields in the reference Matrix
layout and host<->device moves."""
import numpy as np

from falm_amd import _abi as abi


def domain(n, gc=2, offset=(0, 0, 0), gshape=None):
    shape = tuple(int(v) + 2 * gc for v in n)
    return abi.Domain.make(shape, gc=gc, offset=offset, gshape=gshape)


def smooth_field(dom, ncol, rng, base=1.0, amp=0.2):
    """A seeded field: a smooth large-scale part plus small noise, so that
    stencil differences are O(1) but not trivially zero."""
    nx, ny, nz = dom.shape
    x = np.linspace(0, 1, nx)[:, None, None]
    y = np.linspace(0, 1, ny)[None, :, None]
    z = np.linspace(0, 1, nz)[None, None, :]
    cols = []
    for c in range(ncol):
        a, b, d = rng.uniform(0.5, 2.5, 3)
        f = base + amp * np.sin(a * 3.1 * x + c) * np.cos(b * 2.3 * y) * np.sin(d * 1.7 * z + 0.3)
        f = f + 0.05 * amp * rng.standard_normal((nx, ny, nz))
        cols.append(np.ascontiguousarray(f.transpose(2, 1, 0)).ravel())  # x fastest
    return np.concatenate(cols)


def uniform_metrics(dom, h=(0.05, 0.05, 0.05), origin=(0.0, 0.0, 0.0), stretch=0.0):
    """1-D centres/widths (optionally geometrically stretched) and the
    derived metric fields from the oracle."""
    from oracle import pyoracle as O
    out = []
    for m in range(3):
        n = dom.shape[m]
        w = h[m] * (1.0 + stretch) ** np.arange(n)
        c = origin[m] + np.cumsum(w) - 0.5 * w - 2 * h[m]
        out += [c, w]
    x, hx, y, hy, z, hz = out
    xyz, kx, g, ja = O.metrics(dom, x, y, z, hx, hy, hz)
    return (x, y, z, hx, hy, hz), (xyz, kx, g, ja)


def to_dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


class Keep:
    """Holds device tensors for the duration of a test: passing a pointer
    of a temporary tensor lets torch's caching allocator reuse its block
    before the (asynchronous) kernel reads it."""

    def __init__(self):
        self.held = []

    def __call__(self, a, dtype=None):
        import ctypes
        t = a if hasattr(a, "data_ptr") else to_dev(a, dtype)
        self.held.append(t)
        return ctypes.c_void_p(t.data_ptr())


def from_dev(t):
    import torch
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def interior_mask(dom, ncol=1):
    nx, ny, nz = dom.shape
    g = dom.gc
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[g:nz - g, g:ny - g, g:nx - g] = True
    return np.tile(m.ravel(), ncol)


def rel_err(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)
