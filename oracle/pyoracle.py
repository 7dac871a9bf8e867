"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  Arrays are float64
numpy arrays in the reference Matrix layout, flattened (col-major over
[ncell, ncol]).
"""
import contextlib
import ctypes as C
import os

import numpy as np

from falm_amd import _abi as abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b  # noqa: WPS433
            _b.build()
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path):
    """Route every call of this module to another build of the oracle (the
    planted-mistake builds of the pinning tests) for the duration of a
    with-block."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def _load(path):
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    dom = C.POINTER(abi.Domain)
    sig = {
        "orc_pseudo_u": [dom, ip, C.c_double, C.c_double] + [vp] * 9,
        "orc_u_to_uu": [dom, vp, vp, vp, vp],
        "orc_divergence": [dom, vp, vp, vp],
        "orc_scale": [vp, C.c_size_t, C.c_double],
        "orc_poisson_rhs": [dom, vp, vp, vp, C.c_double, C.c_double],
        "orc_project_p": [dom, C.c_double, vp, vp, vp, vp, vp, vp, vp],
        "orc_sgs": [dom, ip, C.c_double, vp, vp, vp, vp],
        "orc_mv": [dom, vp, vp, vp],
        "orc_residual": [dom, vp, vp, vp, vp],
        "orc_reduce": [dom, ip, vp, ip, ip, dp],
        "orc_matrix_add": [vp, vp, C.c_size_t],
        "orc_vecaxby": [C.c_double, vp, C.c_double, vp, vp, C.c_size_t],
        "orc_sor_sweep": [dom, vp, vp, vp, C.c_double, ip],
        "orc_jacobi_sweep": [dom, vp, vp, vp, vp],
        "orc_solver_create": [dom, C.POINTER(abi.SolverConfig), C.POINTER(vp)],
        "orc_solver_destroy": [vp],
        "orc_solver_solve": [vp, vp, vp, vp, vp, C.POINTER(abi.SolveInfo)],
        "orc_inlet_normals": [C.c_uint64, C.c_int64, ip, ip, dp],
        "orc_bc_velocity": [dom, C.POINTER(abi.FaceBC), vp, vp, vp, C.c_double,
                            C.POINTER(abi.InletNoise)],
        "orc_bc_pressure": [dom, C.POINTER(C.c_int), vp],
        "orc_bc_face_flux": [dom, C.POINTER(C.c_int), vp],
        "orc_copy_z5": [dom, vp, ip],
        "orc_metrics": [dom] + [vp] * 10,
        "orc_poisson_matrix": [dom, ip, vp, vp, vp, dp],
        "orc_alm_create": [dom, C.POINTER(abi.AlmConfig), vp, vp, vp, C.POINTER(vp)],
        "orc_alm_destroy": [vp],
        "orc_alm_set_twist_offset": [vp, vp],
        "orc_alm_step": [vp, vp, vp, C.c_double],
        "orc_alm_get": [vp, vp, vp, vp, vp],
        "orc_build_ap": [ip, ip, vp, vp, vp, vp, vp, ip, ip, ip, C.c_double, C.c_double,
                         vp, vp, vp, vp, vp],
        "orc_case_create": [C.POINTER(abi.CaseConfig), C.POINTER(vp)],
        "orc_case_destroy": [vp],
        "orc_case_step": [vp, ip, C.POINTER(abi.StepInfo)],
        "orc_case_get_field": [vp, ip, vp, C.c_size_t],
        "orc_case_set_field": [vp, ip, vp, C.c_size_t],
        "orc_case_alm": [vp, C.POINTER(vp)],
        "orc_case_window": [vp, ip, dp],
        "orc_case_adjoint": [vp, ip, ip, C.c_double, dp, dp, dp, dp],
        "orc_alm_set_pitch": [vp, ip, dp],
        "orc_alm_set_attitude": [vp, ip, C.c_double, C.c_double],
        "orc_alm_set_base": [vp, ip, dp, dp],
        "orc_alm_vjp": [vp, vp, C.c_double, vp, C.c_double, vp, vp, vp],
        "orc_adj_sgs": [dom, C.c_double, vp, vp, vp, vp, vp],
        "orc_adj_pseudo_u": [dom, C.c_double, C.c_double, C.c_double] + [vp] * 11,
        "orc_adj_u_to_uu": [dom, vp, vp, vp, vp],
        "orc_adj_project": [dom, C.c_double, vp, vp, vp, vp, vp, vp, vp],
        "orc_adj_ubc": [dom, C.POINTER(abi.FaceBC), vp, C.c_double, vp, vp, vp],
        "orc_adj_rhs": [dom, vp, C.c_double, vp, vp],
        "orc_adj_pbc": [dom, C.POINTER(C.c_int), vp],
        "orc_num_threads": [],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.orc_dot.argtypes = [dom, vp, vp]
    L.orc_dot.restype = C.c_double
    L.orc_norm2.argtypes = [dom, vp, ip]
    L.orc_norm2.restype = C.c_double
    L.orc_inlet_normals.restype = None
    return L


def ptr(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def check(status):
    if status != 0:
        raise RuntimeError(f"oracle call failed with status {status}")


def zeros(dom, ncol=1):
    return np.zeros(dom.ncell * ncol, dtype=np.float64)


# ------------------------------------------------------------------ stencils
def pseudo_u(dom, scheme, ReI, dt, un, u, uu, nut, kx, g, ja, ff, ua=None):
    ua = zeros(dom, 3) if ua is None else ua
    check(lib().orc_pseudo_u(C.byref(dom), scheme, ReI, dt, ptr(un), ptr(u), ptr(uu), ptr(ua),
                             ptr(nut), ptr(kx), ptr(g), ptr(ja), ptr(ff)))
    return ua


def u_to_uu(dom, u, kx, ja, uu=None):
    uu = zeros(dom, 3) if uu is None else uu
    check(lib().orc_u_to_uu(C.byref(dom), ptr(u), ptr(uu), ptr(kx), ptr(ja)))
    return uu


def divergence(dom, uu, ja, div=None):
    div = zeros(dom) if div is None else div
    check(lib().orc_divergence(C.byref(dom), ptr(uu), ptr(div), ptr(ja)))
    return div


def poisson_rhs(dom, uu, ja, dt, maxdiag, rhs=None):
    rhs = zeros(dom) if rhs is None else rhs
    check(lib().orc_poisson_rhs(C.byref(dom), ptr(uu), ptr(rhs), ptr(ja), dt, maxdiag))
    return rhs


def project_p(dom, dt, u, ua, uu, uua, p, kx, g):
    check(lib().orc_project_p(C.byref(dom), dt, ptr(u), ptr(ua), ptr(uu), ptr(uua), ptr(p),
                              ptr(kx), ptr(g)))


def sgs(dom, model, Cs, u, kx, ja, nut=None):
    nut = zeros(dom) if nut is None else nut
    check(lib().orc_sgs(C.byref(dom), model, Cs, ptr(u), ptr(nut), ptr(kx), ptr(ja)))
    return nut


# -------------------------------------------------------------------- linalg
def mv(dom, a, x, ax=None):
    ax = zeros(dom) if ax is None else ax
    check(lib().orc_mv(C.byref(dom), ptr(a), ptr(x), ptr(ax)))
    return ax


def residual(dom, a, x, b, r=None):
    r = zeros(dom) if r is None else r
    check(lib().orc_residual(C.byref(dom), ptr(a), ptr(x), ptr(b), ptr(r)))
    return r


def dot(dom, a, b):
    return lib().orc_dot(C.byref(dom), ptr(a), ptr(b))


def norm2(dom, a, ncol=1):
    return lib().orc_norm2(C.byref(dom), ptr(a), ncol)


def reduce(dom, op, a, ncol=1, col=0):
    out = C.c_double()
    check(lib().orc_reduce(C.byref(dom), op, ptr(a), ncol, col, C.byref(out)))
    return out.value


def scale(a, s):
    check(lib().orc_scale(ptr(a), a.size, s))


def sor_sweep(dom, a, x, b, omega, color):
    check(lib().orc_sor_sweep(C.byref(dom), ptr(a), ptr(x), ptr(b), omega, color))


def jacobi_sweep(dom, a, x, xp, b):
    check(lib().orc_jacobi_sweep(C.byref(dom), ptr(a), ptr(x), ptr(xp), ptr(b)))


class Solver:
    def __init__(self, dom, cfg):
        self.dom = dom
        h = C.c_void_p()
        check(lib().orc_solver_create(C.byref(dom), C.byref(cfg), C.byref(h)))
        self.h = h

    def solve(self, a, x, b, r):
        info = abi.SolveInfo()
        check(lib().orc_solver_solve(self.h, ptr(a), ptr(x), ptr(b), ptr(r), C.byref(info)))
        return info

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_destroy(self.h)
            self.h = None


# ------------------------------------------------------------------------ BCs
def inlet_normals(seed, step, a, b):
    out = (C.c_double * 3)()
    lib().orc_inlet_normals(seed, step, a, b, out)
    return np.array(list(out))


def bc_velocity(dom, faces, u, uprev, xyz, dt, noise=None):
    arr = (abi.FaceBC * 6)(*faces)
    check(lib().orc_bc_velocity(C.byref(dom), arr, ptr(u), ptr(uprev), ptr(xyz), dt,
                                C.byref(noise) if noise is not None else None))


def bc_pressure(dom, faces, p):
    arr = (C.c_int * 6)(*faces)
    check(lib().orc_bc_pressure(C.byref(dom), arr, ptr(p)))


def bc_face_flux(dom, faces, uu):
    arr = (C.c_int * 6)(*faces)
    check(lib().orc_bc_face_flux(C.byref(dom), arr, ptr(uu)))


def copy_z5(dom, v, ncol):
    check(lib().orc_copy_z5(C.byref(dom), ptr(v), ncol))


# ------------------------------------------------------------ metrics/matrix
def metrics(dom, x, y, z, hx, hy, hz):
    xyz, kx, g, ja = zeros(dom, 3), zeros(dom, 3), zeros(dom, 3), zeros(dom)
    check(lib().orc_metrics(C.byref(dom), ptr(x), ptr(y), ptr(z), ptr(hx), ptr(hy), ptr(hz),
                            ptr(xyz), ptr(kx), ptr(g), ptr(ja)))
    return xyz, kx, g, ja


def poisson_matrix(dom, kind, g, ja):
    a = zeros(dom, 7)
    md = C.c_double()
    check(lib().orc_poisson_matrix(C.byref(dom), kind, ptr(g), ptr(ja), ptr(a), C.byref(md)))
    return a, md.value


# ----------------------------------------------------------------------- ALM
class Alm:
    def __init__(self, dom, cfg, x, y, z):
        self.dom, self.cfg = dom, cfg
        h = C.c_void_p()
        check(lib().orc_alm_create(C.byref(dom), C.byref(cfg.struct), ptr(x), ptr(y), ptr(z),
                                   C.byref(h)))
        self.h = h

    def step(self, u, ff, t):
        check(lib().orc_alm_step(self.h, ptr(u), ptr(ff), t))

    def set_attitude(self, turbine, yaw_deg, angle_offset_deg):
        check(lib().orc_alm_set_attitude(self.h, turbine, float(yaw_deg), float(angle_offset_deg)))

    def set_twist_offset(self, off):
        check(lib().orc_alm_set_twist_offset(self.h, ptr(np.ascontiguousarray(off, dtype=np.float64))))

    def set_base(self, turbine, base, velocity=None):
        b = np.ascontiguousarray(base, dtype=np.float64)
        v = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
        check(lib().orc_alm_set_base(self.h, turbine, b.ctypes.data_as(C.POINTER(C.c_double)),
                                     None if v is None else v.ctypes.data_as(C.POINTER(C.c_double))))

    def vjp(self, u, t, ffbar, wP):
        """(ubar, gpitch[nt*8], gtwist[napb]) of adj_alm."""
        ub = np.zeros(3 * self.dom.ncell)
        gp, gt = np.zeros(8 * self.cfg.n_turbine), np.zeros(self.cfg.n_ap_per_blade)
        check(lib().orc_alm_vjp(self.h, ptr(u), t, ptr(ffbar), wP, ptr(ub), ptr(gp), ptr(gt)))
        return ub, gp, gt

    def get(self):
        nap, nt = self.cfg.n_ap, self.cfg.n_turbine
        xyz, f = np.zeros(3 * nap), np.zeros(3 * nap)
        tq, th = np.zeros(nt), np.zeros(nt)
        check(lib().orc_alm_get(self.h, ptr(xyz), ptr(f), ptr(tq), ptr(th)))
        return xyz.reshape(nap, 3), f.reshape(nap, 3), tq, th

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_alm_destroy(self.h)
            self.h = None


def build_ap(rr, chord, twist, cl, cd, n_turbine, n_blade, n_ap_per_blade, radius, hub_radius):
    nst, nat = cl.shape
    nap = n_turbine * n_blade * n_ap_per_blade
    out = [np.zeros(nap) for _ in range(3)] + [np.zeros(nat * nap), np.zeros(nat * nap)]
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (rr, chord, twist, cl, cd)]
    check(lib().orc_build_ap(nst, nat, *[ptr(v) for v in args], n_turbine, n_blade,
                             n_ap_per_blade, radius, hub_radius, *[ptr(v) for v in out]))
    return out


# ---------------------------------------------------------------------- case
class Case:
    def __init__(self, cfg):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().orc_case_create(C.byref(cfg.struct), C.byref(h)))
        self.h = h
        n = cfg.struct.n
        self.dom = abi.Domain.make([n[0] + 4, n[1] + 4, n[2] + 4])

    def step(self, nsteps=1):
        info = abi.StepInfo()
        check(lib().orc_case_step(self.h, nsteps, C.byref(info)))
        return info

    def get(self, field):
        out = np.zeros(self.dom.ncell * abi.FIELD_NCOL[field])
        check(lib().orc_case_get_field(self.h, field, ptr(out), out.size))
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().orc_case_set_field(self.h, field, ptr(v), v.size))

    def _alm_handle(self):
        h = C.c_void_p()
        check(lib().orc_case_alm(self.h, C.byref(h)))
        return h

    def set_pitch(self, turbine, pitch_deg):
        v = np.zeros(8)
        v[:len(pitch_deg)] = pitch_deg
        check(lib().orc_alm_set_pitch(self._alm_handle(), turbine, v.ctypes.data_as(C.POINTER(C.c_double))))

    def set_attitude(self, turbine, yaw_deg, angle_offset_deg):
        check(lib().orc_alm_set_attitude(self._alm_handle(), turbine, float(yaw_deg), float(angle_offset_deg)))

    def set_twist_offset(self, off):
        check(lib().orc_alm_set_twist_offset(self._alm_handle(), ptr(np.ascontiguousarray(off, dtype=np.float64))))

    def set_base(self, turbine, base, velocity=None):
        b = np.ascontiguousarray(base, dtype=np.float64)
        v = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
        check(lib().orc_alm_set_base(self._alm_handle(), turbine, b.ctypes.data_as(C.POINTER(C.c_double)),
                                     None if v is None else v.ctypes.data_as(C.POINTER(C.c_double))))

    def window(self, nsteps):
        """Mean rotor power over the next nsteps steps (state unchanged)."""
        J = C.c_double()
        check(lib().orc_case_window(self.h, nsteps, C.byref(J)))
        return J.value

    def adjoint(self, nsteps, ckpt, adj_rtol=1e-12, with_state=False):
        """J and dJ/d(pitch [n_turbine x 8], twist offset [napb]) per degree
        (and dJ/du at the window start with with_state)."""
        J = C.c_double()
        gp = np.zeros(8 * self.cfg.alm.n_turbine)
        gt = np.zeros(self.cfg.alm.n_ap_per_blade)
        ub = np.zeros(3 * self.dom.ncell) if with_state else None
        check(lib().orc_case_adjoint(self.h, nsteps, ckpt, adj_rtol, C.byref(J),
                                     gp.ctypes.data_as(C.POINTER(C.c_double)),
                                     gt.ctypes.data_as(C.POINTER(C.c_double)),
                                     ub.ctypes.data_as(C.POINTER(C.c_double)) if with_state else None))
        if with_state:
            return J.value, gp.reshape(-1, 8), gt, ub
        return J.value, gp.reshape(-1, 8), gt

    def alm_get(self):
        h = C.c_void_p()
        check(lib().orc_case_alm(self.h, C.byref(h)))
        nap, nt = self.cfg.alm.n_ap, self.cfg.alm.n_turbine
        xyz, f = np.zeros(3 * nap), np.zeros(3 * nap)
        tq, th = np.zeros(nt), np.zeros(nt)
        check(lib().orc_alm_get(h, ptr(xyz), ptr(f), ptr(tq), ptr(th)))
        return xyz.reshape(nap, 3), f.reshape(nap, 3), tq, th

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_case_destroy(self.h)
            self.h = None


def num_threads():
    return lib().orc_num_threads()


# ---------------------------------------------------------------- adjoints
def adj_sgs(dom, Cs, u, kx, ja, nutbar, ubar=None):
    ubar = np.zeros(3 * dom.ncell) if ubar is None else ubar
    check(lib().orc_adj_sgs(C.byref(dom), Cs, ptr(u), ptr(kx), ptr(ja), ptr(nutbar), ptr(ubar)))
    return ubar


def adj_pseudo_u(dom, alpha, ReI, dt, u, uu, nut, kx, g, ja, uabar):
    ubar, uubar = np.zeros(3 * dom.ncell), np.zeros(3 * dom.ncell)
    nutbar, ffbar = np.zeros(dom.ncell), np.zeros(3 * dom.ncell)
    check(lib().orc_adj_pseudo_u(C.byref(dom), alpha, ReI, dt, ptr(u), ptr(uu), ptr(nut), ptr(kx), ptr(g),
                                 ptr(ja), ptr(uabar), ptr(ubar), ptr(uubar), ptr(nutbar), ptr(ffbar)))
    return ubar, uubar, nutbar, ffbar


def adj_u_to_uu(dom, kx, ja, uubar):
    uabar = np.zeros(3 * dom.ncell)
    check(lib().orc_adj_u_to_uu(C.byref(dom), ptr(kx), ptr(ja), ptr(uubar), ptr(uabar)))
    return uabar


def adj_project(dom, dt, kx, g, ustarbar, uunbar):
    uabar, uustarbar, pbar = np.zeros(3 * dom.ncell), np.zeros(3 * dom.ncell), np.zeros(dom.ncell)
    check(lib().orc_adj_project(C.byref(dom), dt, ptr(kx), ptr(g), ptr(ustarbar), ptr(uunbar), ptr(uabar),
                                ptr(uustarbar), ptr(pbar)))
    return uabar, uustarbar, pbar


def adj_rhs(dom, ja, scale, rhsbar):
    uubar = np.zeros(3 * dom.ncell)
    check(lib().orc_adj_rhs(C.byref(dom), ptr(ja), scale, ptr(rhsbar), ptr(uubar)))
    return uubar


def adj_pbc(dom, faces, pbar):
    out = np.array(pbar, dtype=np.float64, copy=True)
    f = (C.c_int * 6)(*faces)
    check(lib().orc_adj_pbc(C.byref(dom), f, ptr(out)))
    return out


def adj_ubc(dom, faces, xyz, dt, unp1bar):
    fa = (abi.FaceBC * 6)(*faces)
    ustarbar, unbar = np.zeros(3 * dom.ncell), np.zeros(3 * dom.ncell)
    check(lib().orc_adj_ubc(C.byref(dom), fa, ptr(xyz), dt, ptr(unp1bar), ptr(ustarbar), ptr(unbar)))
    return ustarbar, unbar
