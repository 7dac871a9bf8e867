"""Per-stage adjoints (VJPs) of the HIP library through the C ABI.

Each is checked two ways: against the oracle's transpose of the same stage
on seeded inputs (1e-12 relative; the device gathers or adds with atomics,
so only the order of additions differs), and, independently of the oracle,
against the library's own forward stage on the GPU: the transposition identity
<J v, w> = <v, J^T w> for the linear stages and central differences for the
nonlinear ones (pseudo-velocity, Smagorinsky, actuator line)."""
import ctypes as C

import numpy as np
import pytest

from falm_amd import _abi as abi
from oracle import pyoracle as O
from tests.helpers import domain, from_dev, interior_mask, rel_err, smooth_field, to_dev, uniform_metrics

pytestmark = pytest.mark.gpu

TOL = 1e-12
N = (14, 10, 8)


def _fields(dom, seed):
    rng = np.random.default_rng(seed)
    _, (xyz, kx, g, ja) = uniform_metrics(dom, stretch=0.03)
    return rng, dict(u=smooth_field(dom, 3, rng, base=0.8, amp=0.5),
                     uu=smooth_field(dom, 3, rng, base=0.1, amp=0.1),
                     nut=np.abs(smooth_field(dom, 1, rng, base=1e-3, amp=1e-3)),
                     ff=smooth_field(dom, 3, rng, base=0.0, amp=0.3), xyz=xyz, kx=kx, g=g, ja=ja)


def _face_mask(dom, ncol=1):
    """The face region (interior shifted down by one layer, one larger)."""
    nx, ny, nz = dom.shape
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[1:nz - 2, 1:ny - 2, 1:nx - 2] = True
    return np.tile(m.ravel(), ncol)


def _bc_mask(dom, ncol=3):
    """Interior cells and the face ghost cells the boundary fill writes (one
    index in a ghost layer, the other two interior)."""
    nx, ny, nz = dom.shape
    g = dom.gc
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = [(i < g) | (i >= nx - g), (j < g) | (j >= ny - g), (k < g) | (k >= nz - g)]
    nghost = out[0].astype(int) + out[1] + out[2]
    return np.tile((nghost <= 1).ravel(), ncol)


def _ptrs(gpu, *arrs):
    return [gpu.ptr(a) for a in arrs]


def test_u_to_uu_vjp(gpu):
    dom = domain(N)
    rng, F = _fields(dom, 1)
    fm = _face_mask(dom, 3)
    w = rng.standard_normal(3 * dom.ncell) * fm
    ref = O.adj_u_to_uu(dom, F["kx"], F["ja"], w)
    kx, ja, wd = to_dev(F["kx"]), to_dev(F["ja"]), to_dev(w)
    ub = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_u_to_uu_vjp", C.byref(dom), *_ptrs(gpu, kx, ja, wd, ub), None)
    got = from_dev(ub)
    assert rel_err(got, ref) < TOL
    # <UtoUU(v), w> = <v, UtoUU^T w> with the forward on the GPU
    v = rng.standard_normal(3 * dom.ncell)
    vd, uu = to_dev(v), to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_u_to_uu", C.byref(dom), abi.F64, *_ptrs(gpu, vd, uu, kx, ja), None)
    lhs, rhs = from_dev(uu) @ w, v @ got
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_poisson_rhs_vjp(gpu):
    dom = domain(N)
    rng, F = _fields(dom, 2)
    dt, md = 2e-3, 7.5
    w = rng.standard_normal(dom.ncell) * interior_mask(dom)
    ref = O.adj_rhs(dom, F["ja"], 1.0 / (dt * md), w)
    ja, wd = to_dev(F["ja"]), to_dev(w)
    ub = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_poisson_rhs_vjp", C.byref(dom), gpu.ptr(ja), dt, md, gpu.ptr(wd), gpu.ptr(ub), None)
    got = from_dev(ub)
    assert rel_err(got, ref) < TOL
    v = rng.standard_normal(3 * dom.ncell)
    vd, rhs = to_dev(v), to_dev(np.zeros(dom.ncell))
    gpu.call("falm_poisson_rhs", C.byref(dom), abi.F64, gpu.ptr(vd), gpu.ptr(rhs), gpu.ptr(ja), dt, md, None)
    lhs, rh = from_dev(rhs) @ w, v @ got
    assert abs(lhs - rh) <= 1e-12 * abs(lhs)


def test_project_p_vjp(gpu):
    dom = domain(N)
    rng, F = _fields(dom, 3)
    dt = 3e-3
    ub = rng.standard_normal(3 * dom.ncell) * interior_mask(dom, 3)
    uub = rng.standard_normal(3 * dom.ncell) * _face_mask(dom, 3)
    ra, rua, rp = O.adj_project(dom, dt, F["kx"], F["g"], ub, uub)
    kx, g = to_dev(F["kx"]), to_dev(F["g"])
    ubd, uubd = to_dev(ub), to_dev(uub)
    a, ua, p = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(dom.ncell))
    gpu.call("falm_project_p_vjp", C.byref(dom), dt, *_ptrs(gpu, kx, g, ubd, uubd, a, ua, p), None)
    ga, gua, gp = from_dev(a), from_dev(ua), from_dev(p)
    assert rel_err(ga, ra) < TOL and rel_err(gua, rua) < TOL and rel_err(gp, rp) < TOL
    # the forward (u, uu) = (va - dt kx grad vp, vua - dt g grad vp) on the GPU
    va, vua, vp = (rng.standard_normal(3 * dom.ncell), rng.standard_normal(3 * dom.ncell),
                   rng.standard_normal(dom.ncell))
    vad, vuad, vpd = to_dev(va), to_dev(vua), to_dev(vp)
    uo, uuo = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_project_p", C.byref(dom), abi.F64, dt, *_ptrs(gpu, uo, vad, uuo, vuad, vpd, kx, g), None)
    lhs = from_dev(uo) @ ub + from_dev(uuo) @ uub
    rhs = va @ ga + vua @ gua + vp @ gp
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_sgs_vjp(gpu):
    dom = domain(N)
    rng, F = _fields(dom, 4)
    Cs = 0.12
    w = rng.standard_normal(dom.ncell) * interior_mask(dom)
    ref = O.adj_sgs(dom, Cs, F["u"], F["kx"], F["ja"], w)
    u, kx, ja, wd = to_dev(F["u"]), to_dev(F["kx"]), to_dev(F["ja"]), to_dev(w)
    ub = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_sgs_vjp", C.byref(dom), Cs, *_ptrs(gpu, u, kx, ja, wd, ub), None)
    got = from_dev(ub)
    assert rel_err(got, ref) < TOL
    # central differences of the GPU's own Smagorinsky
    v = rng.standard_normal(3 * dom.ncell)
    eps = 1e-6

    def nut_at(e):
        ud, nut = to_dev(F["u"] + e * v), to_dev(np.zeros(dom.ncell))
        gpu.call("falm_sgs", C.byref(dom), abi.F64, abi.SGS_SMAGORINSKY, Cs, *_ptrs(gpu, ud, nut, kx, ja), None)
        return from_dev(nut)

    fd = (nut_at(eps) - nut_at(-eps)) @ w / (2 * eps)
    assert abs(fd - v @ got) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("scheme", [abi.ADV_UPWIND3, abi.ADV_UTOPIA, abi.ADV_KK])
def test_pseudo_u_vjp(gpu, scheme):
    dom = domain(N)
    rng, F = _fields(dom, 10 + scheme)
    ReI, dt = 1.0 / 2500.0, 2e-3
    alpha = {abi.ADV_UPWIND3: 1 / 24, abi.ADV_UTOPIA: 1 / 12, abi.ADV_KK: 1 / 4}[scheme]
    w = rng.standard_normal(3 * dom.ncell) * interior_mask(dom, 3)
    r_u, r_uu, r_nut, r_ff = O.adj_pseudo_u(dom, alpha, ReI, dt, F["u"], F["uu"], F["nut"], F["kx"], F["g"],
                                            F["ja"], w)
    d = {k: to_dev(v) for k, v in F.items()}
    wd = to_dev(w)
    ub, uub = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(3 * dom.ncell))
    nb, fb = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_pseudo_u_vjp", C.byref(dom), scheme, ReI, dt,
             *_ptrs(gpu, d["u"], d["uu"], d["nut"], d["kx"], d["g"], d["ja"], wd, ub, uub, nb, fb), None)
    g_u, g_uu, g_nut, g_ff = from_dev(ub), from_dev(uub), from_dev(nb), from_dev(fb)
    for a, b in ((g_u, r_u), (g_uu, r_uu), (g_nut, r_nut), (g_ff, r_ff)):
        assert rel_err(a, b) < TOL
    # central differences of the GPU's own pseudo-velocity, every input
    # perturbed at once (u in both of its roles)
    vu, vuu = rng.standard_normal(3 * dom.ncell), rng.standard_normal(3 * dom.ncell)
    vnut, vff = rng.standard_normal(dom.ncell) * 1e-3, rng.standard_normal(3 * dom.ncell)
    eps = 1e-6

    def ua_at(e):
        u = to_dev(F["u"] + e * vu)
        uu, nut, ff = to_dev(F["uu"] + e * vuu), to_dev(F["nut"] + e * vnut), to_dev(F["ff"] + e * vff)
        ua = to_dev(np.zeros(3 * dom.ncell))
        gpu.call("falm_pseudo_u", C.byref(dom), abi.F64, scheme, ReI, dt, gpu.ptr(u), gpu.ptr(u), gpu.ptr(uu),
                 gpu.ptr(ua), gpu.ptr(nut), gpu.ptr(d["kx"]), gpu.ptr(d["g"]), gpu.ptr(d["ja"]), gpu.ptr(ff), None)
        return from_dev(ua)

    fd = (ua_at(eps) - ua_at(-eps)) @ w / (2 * eps)
    ad = vu @ g_u + vuu @ g_uu + vnut @ g_nut + vff @ g_ff
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_pseudo_u_vjp_rejects_other_schemes(gpu):
    dom = domain((6, 6, 6))
    z3, z1 = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(dom.ncell))
    with pytest.raises(gpu.FalmError) as e:
        gpu.call("falm_pseudo_u_vjp", C.byref(dom), abi.ADV_QUICK, 1e-3, 1e-3,
                 *_ptrs(gpu, z3, z3, z1, z3, z3, z1, z3, z3, z3, z1, z3), None)
    assert e.value.status == abi.FALM_ERR_UNSUPPORTED


def _windtunnel_faces(inlet=0.0):
    f = (abi.FaceBC * 6)()
    f[0].type, f[0].value[0] = abi.BC_INLET, inlet
    f[1].type = abi.BC_CONVECTIVE
    f[2].type = f[3].type = f[5].type = abi.BC_SLIP
    f[4].type = abi.BC_WALL0
    return f


def test_bc_velocity_vjp(gpu):
    dom = domain(N)
    rng, F = _fields(dom, 5)
    dt = 4e-3
    faces = _windtunnel_faces()
    w = rng.standard_normal(3 * dom.ncell) * _bc_mask(dom)
    r_s, r_n = O.adj_ubc(dom, list(faces), F["xyz"], dt, w)
    xyz, wd = to_dev(F["xyz"]), to_dev(w)
    sb, nb = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_bc_velocity_vjp", C.byref(dom), faces, *_ptrs(gpu, xyz), dt, *_ptrs(gpu, wd, sb, nb), None)
    g_s, g_n = from_dev(sb), from_dev(nb)
    assert rel_err(g_s, r_s) < TOL and rel_err(g_n, r_n) < TOL
    # the forward fill is linear in (u, uprev) with a zero inlet value
    vs, vn = rng.standard_normal(3 * dom.ncell), rng.standard_normal(3 * dom.ncell)
    u, up = to_dev(vs), to_dev(vn)
    gpu.call("falm_bc_velocity", C.byref(dom), abi.F64, faces, gpu.ptr(u), gpu.ptr(up), gpu.ptr(xyz), dt, None,
             None)
    lhs = from_dev(u) @ w
    rhs = (vs * interior_mask(dom, 3)) @ g_s + vn @ g_n
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_bc_pressure_vjp(gpu):
    dom = domain(N)
    rng = np.random.default_rng(6)
    faces = [abi.PBC_NEUMANN, abi.PBC_ZERO, abi.PBC_NEUMANN, abi.PBC_NEUMANN, abi.PBC_NEUMANN, abi.PBC_NEUMANN]
    w = rng.standard_normal(dom.ncell)
    ref = O.adj_pbc(dom, faces, w)
    fa = (C.c_int * 6)(*faces)
    pb = to_dev(w)
    gpu.call("falm_bc_pressure_vjp", C.byref(dom), fa, gpu.ptr(pb), None)
    got = from_dev(pb)
    assert rel_err(got, ref) < TOL
    v = rng.standard_normal(dom.ncell)
    p = to_dev(v)
    gpu.call("falm_bc_pressure", C.byref(dom), abi.F64, fa, gpu.ptr(p), None)
    lhs, rhs = from_dev(p) @ w, v @ got
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_alm_vjp(gpu):
    """The actuator-line step at time t transposed: u, blade pitch and twist
    offsets from the adjoint of the projected force plus a weight on the
    rotor power; against the oracle, and against central differences of the
    GPU's own ALM step."""
    from falm_amd import cases
    dom = domain((32, 20, 20))
    h = 1.0 / 8.0
    setup = cases.rotor(h, base=(1.6, 1.25, 1.25), radius=0.5, tsr=7.0, n_ap_per_blade=8, angle_type=2,
                        motion=(np.radians(3.0), 0.8, 0.3), cutoff=0.0)
    x, y, z = ((np.arange(dom.shape[m]) - 1.5) * h for m in range(3))
    rng = np.random.default_rng(41)
    u = smooth_field(dom, 3, rng, base=1.0, amp=0.2)
    ffbar = rng.standard_normal(3 * dom.ncell) * interior_mask(dom, 3)
    t, wP = 0.37, 0.25
    nt, napb = setup.n_turbine, setup.n_ap_per_blade
    ref = O.Alm(dom, setup, x, y, z)
    r_u, r_gp, r_gt = ref.vjp(u, t, ffbar, wP)
    A = gpu.Alm(dom, setup, x, y, z)
    ud, fb = to_dev(u), to_dev(ffbar)
    ub, gp, gt = to_dev(np.zeros(3 * dom.ncell)), to_dev(np.zeros(8 * nt)), to_dev(np.zeros(napb))
    gpu.call("falm_alm_vjp", A.h, gpu.ptr(ud), t, gpu.ptr(fb), wP, *_ptrs(gpu, ub, gp, gt), None)
    g_u, g_p, g_t = from_dev(ub), from_dev(gp), from_dev(gt)
    assert rel_err(g_u, r_u) < 1e-10 and rel_err(g_p, r_gp) < 1e-10 and rel_err(g_t, r_gt) < 1e-10
    assert np.any(g_p[:3] != 0) and np.all(g_p[3:8] == 0)  # three blades
    tip = setup.turbines[0].tip_rate

    def L(uh, pitch=(0.0, 0.0, 0.0), twist=None):
        A.set_pitch(0, list(pitch))
        A.set_twist_offset(np.zeros(napb) if twist is None else twist)
        ff, uh_d = to_dev(np.zeros(3 * dom.ncell)), to_dev(uh)
        A.step(abi.F64, uh_d, ff, t)
        _, _, tq, _ = A.get()
        return from_dev(ff) @ ffbar + wP * tip * tq[0]

    v = rng.standard_normal(3 * dom.ncell) * interior_mask(dom, 3)
    eps = 1e-6
    fd = (L(u + eps * v) - L(u - eps * v)) / (2 * eps)
    assert abs(fd - v @ g_u) <= 1e-6 * abs(fd)
    e = 1e-5  # degrees
    fd = (L(u, pitch=(0.0, e, 0.0)) - L(u, pitch=(0.0, -e, 0.0))) / (2 * e)
    assert abs(fd - g_p[1]) <= 1e-5 * abs(fd) + 1e-12
    dt_ = np.zeros(napb)
    dt_[3] = e
    fd = (L(u, twist=dt_) - L(u, twist=-dt_)) / (2 * e)
    assert abs(fd - g_t[3]) <= 1e-5 * abs(fd) + 1e-12


def test_vjp_outputs_must_be_device_memory(gpu):
    """The VJP kernels accumulate with hardware FP64 atomics: an output in
    host-pinned memory is refused with FALM_ERR_INVALID (and nothing is
    written); the same call with device memory succeeds."""
    import torch
    dom = domain(N)
    rng, F = _fields(dom, 9)
    kx, ja = to_dev(F["kx"]), to_dev(F["ja"])
    wd = to_dev(rng.standard_normal(3 * dom.ncell))
    pinned = torch.zeros(3 * dom.ncell, dtype=torch.float64, pin_memory=True)
    st = gpu.lib().falm_u_to_uu_vjp(C.byref(dom), *_ptrs(gpu, kx, ja, wd), C.c_void_p(pinned.data_ptr()), None)
    assert st == abi.FALM_ERR_INVALID
    assert b"device memory" in gpu.lib().falm_last_error()
    assert float(pinned.abs().max()) == 0.0
    ub = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_u_to_uu_vjp", C.byref(dom), *_ptrs(gpu, kx, ja, wd, ub), None)
