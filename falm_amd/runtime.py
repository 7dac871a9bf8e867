"""Loader and thin ctypes bindings of libfalm_amd.so (the C ABI in
include/falm_amd.h).

There is no fallback: if the library is missing or no GPU is present the
calls raise.  Field arguments may be torch tensors on the GPU (their
data_ptr() is passed) or raw integer device pointers.
"""
import ctypes as C
import os

from . import _abi as abi

HERE = os.path.dirname(os.path.abspath(__file__))
# FALM_AMD_LIB selects another build of the library (same ABI)
LIB_PATH = os.environ.get("FALM_AMD_LIB") or os.path.join(HERE, "lib", "libfalm_amd.so")

_lib = None


class FalmError(RuntimeError):
    def __init__(self, status, where, msg=""):
        self.status = status
        super().__init__(f"{where} failed with status {status}: {msg}")


def _sigs(L):
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    dom = C.POINTER(abi.Domain)
    S = {
        "falm_init": [ip],
        "falm_malloc": [C.POINTER(vp), C.c_size_t],
        "falm_free": [vp],
        "falm_memcpy": [vp, vp, C.c_size_t, ip, vp],
        "falm_memset": [vp, ip, C.c_size_t, vp],
        "falm_stream_sync": [vp],
        "falm_pseudo_u": [dom, ip, ip, dp, dp] + [vp] * 10,
        "falm_u_to_uu": [dom, ip, vp, vp, vp, vp, vp],
        "falm_poisson_rhs": [dom, ip, vp, vp, vp, dp, dp, vp],
        "falm_project_p": [dom, ip, dp] + [vp] * 8,
        "falm_sgs": [dom, ip, ip, dp, vp, vp, vp, vp, vp],
        "falm_divergence": [dom, ip, vp, vp, vp, vp],
        "falm_mv": [dom, ip, vp, vp, vp, vp],
        "falm_dot": [dom, ip, vp, vp, C.POINTER(dp), vp],
        "falm_dot_noalloc": [dom, ip, vp, vp, vp, vp],
        "falm_norm2": [dom, ip, vp, ip, C.POINTER(dp), vp],
        "falm_reduce": [dom, ip, ip, vp, ip, ip, C.POINTER(dp), vp],
        "falm_scale": [ip, vp, C.c_size_t, dp, vp],
        "falm_matrix_add": [ip, vp, vp, C.c_size_t, vp],
        "falm_vecaxby": [ip, dp, vp, dp, vp, vp, C.c_size_t, vp],
        "falm_residual": [dom, ip, vp, vp, vp, vp, vp],
        "falm_sor_sweep": [dom, ip, vp, vp, vp, dp, ip, vp],
        "falm_jacobi_sweep": [dom, ip, vp, vp, vp, vp, vp],
        "falm_solver_create": [dom, C.POINTER(abi.SolverConfig), C.POINTER(vp)],
        "falm_solver_solve": [vp, vp, vp, vp, vp, vp, C.POINTER(abi.SolveInfo)],
        "falm_solver_set_matrix": [vp, vp, vp],
        "falm_solver_destroy": [vp],
        "falm_bc_velocity": [dom, ip, C.POINTER(abi.FaceBC), vp, vp, vp, dp, C.POINTER(abi.InletNoise), vp],
        "falm_bc_pressure": [dom, ip, C.POINTER(C.c_int), vp, vp],
        "falm_bc_face_flux": [dom, ip, C.POINTER(C.c_int), vp, vp],
        "falm_copy_z5": [dom, ip, vp, ip, vp],
        "falm_metrics": [dom, ip] + [vp] * 11,
        "falm_poisson_matrix": [dom, ip, ip, vp, vp, vp, C.POINTER(dp), vp],
        "falm_alm_create": [dom, C.POINTER(abi.AlmConfig), vp, vp, vp, C.POINTER(vp)],
        "falm_alm_step": [vp, ip, vp, vp, dp, vp],
        "falm_alm_get": [vp, vp, vp, vp, vp, vp],
        "falm_alm_set_twist_offset": [vp, vp],
        "falm_alm_destroy": [vp],
        "falm_build_ap": [ip, ip, vp, vp, vp, vp, vp, ip, ip, ip, dp, dp, vp, vp, vp, vp, vp],
        "falm_case_create": [C.POINTER(abi.CaseConfig), C.POINTER(vp)],
        "falm_case_step": [vp, ip, C.POINTER(abi.StepInfo)],
        "falm_case_get_field": [vp, ip, vp, C.c_size_t],
        "falm_case_set_field": [vp, ip, vp, C.c_size_t],
        "falm_case_domain": [vp, dom],
        "falm_case_alm": [vp, C.POINTER(vp)],
        "falm_case_destroy": [vp],
        "falm_case_window": [vp, ip, C.POINTER(dp)],
        "falm_case_adjoint": [vp, ip, ip, dp, C.POINTER(dp), vp, vp, C.POINTER(C.c_size_t)],
        "falm_alm_set_pitch": [vp, ip, vp],
        "falm_alm_set_attitude": [vp, ip, dp, dp],
        "falm_case_set_tolerance": [vp, dp, ip],
        "falm_case_adjoint_attitude": [vp, vp, vp],
        "falm_alm_set_base": [vp, ip, vp, vp, vp],
        "falm_mesher_build": [C.c_char_p, C.c_char_p, C.c_char_p, ip],
        "falm_pseudo_u_vjp": [dom, ip, dp, dp] + [vp] * 12,
        "falm_u_to_uu_vjp": [dom, vp, vp, vp, vp, vp],
        "falm_poisson_rhs_vjp": [dom, vp, dp, dp, vp, vp, vp],
        "falm_project_p_vjp": [dom, dp] + [vp] * 8,
        "falm_sgs_vjp": [dom, dp, vp, vp, vp, vp, vp, vp],
        "falm_bc_velocity_vjp": [dom, C.POINTER(abi.FaceBC), vp, dp, vp, vp, vp, vp],
        "falm_bc_pressure_vjp": [dom, C.POINTER(C.c_int), vp, vp],
        "falm_alm_vjp": [vp, vp, dp, vp, dp, vp, vp, vp, vp],
        "falm_timing_enable": [ip],
        "falm_timing_read": [C.POINTER(ip), C.POINTER(dp), C.POINTER(dp)],
    }
    for name, args in S.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.falm_last_error.argtypes = []
    L.falm_last_error.restype = C.c_char_p
    L.falm_status_string.argtypes = [ip]
    L.falm_status_string.restype = C.c_char_p
    return list(S) + ["falm_last_error", "falm_status_string"]


EXPORTS = None


def lib():
    """Load libfalm_amd.so (raises if it has not been built)."""
    global _lib, EXPORTS
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m falm_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        EXPORTS = _sigs(L)
        _lib = L
    return _lib


def check(status, where):
    if status != abi.FALM_OK:
        msg = lib().falm_last_error().decode(errors="replace")
        raise FalmError(status, where, msg)
    return status


def ptr(x):
    """Device pointer of a torch tensor (or an int / None)."""
    if x is None:
        return None
    if isinstance(x, C.c_void_p):
        return x
    if isinstance(x, int):
        return C.c_void_p(x)
    return C.c_void_p(x.data_ptr())


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)


def init(device=0):
    call("falm_init", device)


class Solver:
    def __init__(self, dom, cfg):
        h = C.c_void_p()
        call("falm_solver_create", C.byref(dom), C.byref(cfg), C.byref(h))
        self.h = h

    def set_matrix(self, a, stream=None):
        call("falm_solver_set_matrix", self.h, ptr(a), stream)

    def solve(self, a, x, b, r, stream=None):
        info = abi.SolveInfo()
        call("falm_solver_solve", self.h, ptr(a), ptr(x), ptr(b), ptr(r), stream, C.byref(info))
        return info

    def close(self):
        if getattr(self, "h", None):
            lib().falm_solver_destroy(self.h)
            self.h = None

    __del__ = close


class Alm:
    def __init__(self, dom, setup, x, y, z):
        import numpy as np
        self.setup = setup
        self._xyz = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        h = C.c_void_p()
        call("falm_alm_create", C.byref(dom), C.byref(setup.struct),
             *[v.ctypes.data_as(C.c_void_p) for v in self._xyz], C.byref(h))
        self.h = h

    def step(self, prec, u, ff, t, stream=None):
        call("falm_alm_step", self.h, prec, ptr(u), ptr(ff), t, stream)

    def set_twist_offset(self, off):
        import numpy as np
        v = np.ascontiguousarray(off, dtype=np.float64)
        call("falm_alm_set_twist_offset", self.h, v.ctypes.data_as(C.c_void_p))

    def set_pitch(self, turbine, pitch_deg):
        import numpy as np
        v = np.zeros(8)
        v[:len(pitch_deg)] = pitch_deg
        call("falm_alm_set_pitch", self.h, turbine, v.ctypes.data_as(C.c_void_p))

    def get(self, stream=None):
        import numpy as np
        nap, nt = self.setup.n_ap, self.setup.n_turbine
        xyz, f = np.zeros(3 * nap), np.zeros(3 * nap)
        tq, th = np.zeros(nt), np.zeros(nt)
        call("falm_alm_get", self.h, *[v.ctypes.data_as(C.c_void_p) for v in (xyz, f, tq, th)], stream)
        return xyz.reshape(nap, 3), f.reshape(nap, 3), tq, th

    def close(self):
        if getattr(self, "h", None):
            lib().falm_alm_destroy(self.h)
            self.h = None

    __del__ = close


class Case:
    """A device-resident simulation (falm_case_*)."""

    def __init__(self, setup):
        self.setup = setup
        h = C.c_void_p()
        call("falm_case_create", C.byref(setup.struct), C.byref(h))
        self.h = h
        self.dom = abi.Domain()
        call("falm_case_domain", self.h, C.byref(self.dom))

    def step(self, nsteps=1):
        info = abi.StepInfo()
        call("falm_case_step", self.h, nsteps, C.byref(info))
        return info

    def get(self, field):
        import numpy as np
        out = np.zeros(self.dom.ncell * abi.FIELD_NCOL[field])
        call("falm_case_get_field", self.h, field, out.ctypes.data_as(C.c_void_p), out.size)
        return out

    def set(self, field, values):
        import numpy as np
        v = np.ascontiguousarray(values, dtype=np.float64)
        call("falm_case_set_field", self.h, field, v.ctypes.data_as(C.c_void_p), v.size)

    def _alm_handle(self):
        h = C.c_void_p()
        call("falm_case_alm", self.h, C.byref(h))
        return h

    def set_pitch(self, turbine, pitch_deg):
        """Per-blade pitch of one turbine in degrees (up to 8 values)."""
        import numpy as np
        v = np.zeros(8)
        v[:len(pitch_deg)] = pitch_deg
        call("falm_alm_set_pitch", self._alm_handle(), turbine, v.ctypes.data_as(C.c_void_p))

    def set_tolerance(self, tol, maxit):
        """Poisson tolerance (|r|/N) and iteration limit from the next step on."""
        call("falm_case_set_tolerance", self.h, float(tol), int(maxit))

    def set_attitude(self, turbine, yaw_deg, angle_offset_deg):
        """Static yaw and platform angle offset of one turbine, degrees."""
        call("falm_alm_set_attitude", self._alm_handle(), turbine, float(yaw_deg), float(angle_offset_deg))

    def adjoint_attitude(self):
        """dJ/d(yaw_deg) and dJ/d(angle_offset_deg) per turbine (per degree)
        from the last adjoint() call."""
        import numpy as np
        nt = self.setup.alm.n_turbine
        gy, go = np.zeros(nt), np.zeros(nt)
        call("falm_case_adjoint_attitude", self.h, gy.ctypes.data_as(C.c_void_p), go.ctypes.data_as(C.c_void_p))
        return gy, go

    def set_base(self, turbine, base, velocity=None):
        """Base position and velocity of one turbine from the next step on."""
        import numpy as np
        b = np.ascontiguousarray(base, dtype=np.float64)
        v = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
        call("falm_alm_set_base", self._alm_handle(), turbine, b.ctypes.data_as(C.c_void_p),
             None if v is None else v.ctypes.data_as(C.c_void_p), None)

    def set_twist_offset(self, off):
        """Per-section twist offset in degrees ([n_ap_per_blade])."""
        import numpy as np
        v = np.ascontiguousarray(off, dtype=np.float64)
        call("falm_alm_set_twist_offset", self._alm_handle(), v.ctypes.data_as(C.c_void_p))

    def window(self, nsteps):
        """Mean rotor power over the next nsteps steps (state unchanged)."""
        J = C.c_double()
        call("falm_case_window", self.h, nsteps, C.byref(J))
        return J.value

    def adjoint(self, nsteps, ckpt, adj_rtol=1e-12):
        """J and dJ/d(pitch [n_turbine x 8], twist offset [n_ap_per_blade]),
        per degree, by the checkpointed discrete adjoint; also the device
        bytes it allocated.  The state is unchanged."""
        import numpy as np
        a = self.setup.alm
        gp, gt = np.zeros(a.n_turbine * 8), np.zeros(a.n_ap_per_blade)
        J, nb = C.c_double(), C.c_size_t()
        call("falm_case_adjoint", self.h, nsteps, ckpt, adj_rtol, C.byref(J), gp.ctypes.data_as(C.c_void_p),
             gt.ctypes.data_as(C.c_void_p), C.byref(nb))
        return J.value, gp.reshape(a.n_turbine, 8), gt, nb.value

    def alm_get(self):
        import numpy as np
        h = self._alm_handle()
        a = self.setup.alm
        xyz, f = np.zeros(3 * a.n_ap), np.zeros(3 * a.n_ap)
        tq, th = np.zeros(a.n_turbine), np.zeros(a.n_turbine)
        call("falm_alm_get", h, *[v.ctypes.data_as(C.c_void_p) for v in (xyz, f, tq, th)], None)
        return xyz.reshape(a.n_ap, 3), f.reshape(a.n_ap, 3), tq, th

    def close(self):
        if getattr(self, "h", None):
            lib().falm_case_destroy(self.h)
            self.h = None

    __del__ = close


def timing_enable(family):
    call("falm_timing_enable", family)


def timing_read():
    n, ms, b = C.c_int(), C.c_double(), C.c_double()
    call("falm_timing_read", C.byref(n), C.byref(ms), C.byref(b))
    return n.value, ms.value, b.value
