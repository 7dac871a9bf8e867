"""Per-operation parity of the HIP library (through the C ABI) with the CPU
oracle on seeded inputs.  FP64 bar: 1e-12 relative (max-norm) unless the
operation reorders a long reduction; FP32 bar: 1e-5 relative."""
import ctypes as C

import numpy as np
import pytest

from falm_amd import _abi as abi
from oracle import pyoracle as O
from tests.helpers import Keep, domain, from_dev, interior_mask, rel_err, smooth_field, to_dev, uniform_metrics

pytestmark = pytest.mark.gpu

TOL64 = 1e-12

# device buffers passed by pointer stay referenced for the whole module
K = Keep()


def _fields(dom, seed=0, stretch=0.0):
    rng = np.random.default_rng(seed)
    _, (xyz, kx, g, ja) = uniform_metrics(dom, stretch=stretch)
    u = smooth_field(dom, 3, rng, base=0.8, amp=0.5)
    uu = smooth_field(dom, 3, rng, base=0.1, amp=0.1)
    nut = np.abs(smooth_field(dom, 1, rng, base=1e-3, amp=1e-3))
    ff = smooth_field(dom, 3, rng, base=0.0, amp=0.3)
    un = u + 0.01 * smooth_field(dom, 3, rng, base=0.0, amp=1.0)
    return dict(u=u, un=un, uu=uu, nut=nut, ff=ff, xyz=xyz, kx=kx, g=g, ja=ja)


@pytest.mark.parametrize("scheme", [abi.ADV_UPWIND1, abi.ADV_UPWIND3, abi.ADV_QUICK, abi.ADV_UTOPIA, abi.ADV_KK])
@pytest.mark.parametrize("stretch", [0.0, 0.03])
def test_pseudo_u(gpu, scheme, stretch):
    dom = domain((19, 13, 11))
    F = _fields(dom, seed=scheme, stretch=stretch)
    ReI, dt = 1.0 / 3200.0, 1e-3
    ref = O.pseudo_u(dom, scheme, ReI, dt, F["un"], F["u"], F["uu"], F["nut"], F["kx"], F["g"], F["ja"], F["ff"])
    d = {k: to_dev(v) for k, v in F.items()}
    ua = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_pseudo_u", C.byref(dom), abi.F64, scheme, ReI, dt, *[gpu.ptr(d[k]) for k in ("un", "u", "uu")],
             gpu.ptr(ua), *[gpu.ptr(d[k]) for k in ("nut", "kx", "g", "ja", "ff")], None)
    m = interior_mask(dom, 3)
    assert rel_err(from_dev(ua), ref, m) < TOL64
    # float variant
    d32 = {k: to_dev(v.astype(np.float32)) for k, v in F.items()}
    ua32 = to_dev(np.zeros(3 * dom.ncell, dtype=np.float32))
    gpu.call("falm_pseudo_u", C.byref(dom), abi.F32, scheme, ReI, dt, *[gpu.ptr(d32[k]) for k in ("un", "u", "uu")],
             gpu.ptr(ua32), *[gpu.ptr(d32[k]) for k in ("nut", "kx", "g", "ja", "ff")], None)
    assert rel_err(from_dev(ua32), ref, m) < 1e-5


def test_pseudo_u_rejects_alias(gpu):
    dom = domain((8, 8, 8))
    F = _fields(dom)
    d = {k: to_dev(v) for k, v in F.items()}
    with pytest.raises(gpu.FalmError) as e:
        gpu.call("falm_pseudo_u", C.byref(dom), abi.F64, abi.ADV_UPWIND3, 1e-3, 1e-3, gpu.ptr(d["u"]), gpu.ptr(d["u"]),
                 gpu.ptr(d["uu"]), gpu.ptr(d["u"]), *[gpu.ptr(d[k]) for k in ("nut", "kx", "g", "ja", "ff")], None)
    assert e.value.status == abi.FALM_ERR_ALIAS


def test_u_to_uu_rhs_divergence(gpu):
    dom = domain((17, 12, 9))
    F = _fields(dom, seed=5, stretch=0.02)
    uu_ref = O.u_to_uu(dom, F["u"], F["kx"], F["ja"])
    d = {k: to_dev(v) for k, v in F.items()}
    uu = to_dev(np.zeros(3 * dom.ncell))
    gpu.call("falm_u_to_uu", C.byref(dom), abi.F64, gpu.ptr(d["u"]), gpu.ptr(uu), gpu.ptr(d["kx"]), gpu.ptr(d["ja"]),
             None)
    got = from_dev(uu)
    # the face region: interior shifted down by one layer, one larger
    nx, ny, nz = dom.shape
    fm = np.zeros((nz, ny, nx), dtype=bool)
    fm[1:nz - 2, 1:ny - 2, 1:nx - 2] = True
    assert rel_err(got, uu_ref, np.tile(fm.ravel(), 3)) < TOL64
    # rhs = div / (dt maxdiag) incl. the ghost scaling
    rhs0 = np.random.default_rng(1).standard_normal(dom.ncell)
    rhs_ref = rhs0.copy()
    O.poisson_rhs(dom, uu_ref, F["ja"], 2e-3, 7.5, rhs_ref)
    rhs = to_dev(rhs0)
    gpu.call("falm_poisson_rhs", C.byref(dom), abi.F64, gpu.ptr(uu), gpu.ptr(rhs), gpu.ptr(d["ja"]), 2e-3, 7.5, None)
    assert rel_err(from_dev(rhs), rhs_ref) < TOL64
    div = to_dev(np.zeros(dom.ncell))
    gpu.call("falm_divergence", C.byref(dom), abi.F64, gpu.ptr(uu), gpu.ptr(div), gpu.ptr(d["ja"]), None)
    div_ref = O.divergence(dom, uu_ref, F["ja"])
    assert rel_err(from_dev(div), div_ref, interior_mask(dom)) < TOL64


@pytest.mark.parametrize("inplace", [True, False])
def test_project_p(gpu, inplace):
    dom = domain((15, 14, 10))
    F = _fields(dom, seed=7, stretch=0.01)
    p = smooth_field(dom, 1, np.random.default_rng(3), base=0.0, amp=1.0)
    u_ref, uu_ref = F["u"].copy(), F["uu"].copy()
    ua, uua = (u_ref, uu_ref) if inplace else (F["un"], F["uu"] * 1.1)
    O.project_p(dom, 3e-3, u_ref, ua.copy() if not inplace else u_ref, uu_ref, uua.copy() if not inplace else uu_ref,
                p, F["kx"], F["g"])
    du, duu = to_dev(F["u"]), to_dev(F["uu"])
    dua, duua = (du, duu) if inplace else (to_dev(ua), to_dev(uua))
    gpu.call("falm_project_p", C.byref(dom), abi.F64, 3e-3, gpu.ptr(du), gpu.ptr(dua), gpu.ptr(duu), gpu.ptr(duua),
             K(to_dev(p)), K(to_dev(F["kx"])), K(to_dev(F["g"])), None)
    assert rel_err(from_dev(du), u_ref) < TOL64
    assert rel_err(from_dev(duu), uu_ref) < TOL64


def test_smagorinsky(gpu):
    dom = domain((16, 11, 13))
    F = _fields(dom, seed=9, stretch=0.02)
    ref = O.sgs(dom, abi.SGS_SMAGORINSKY, 0.1, F["u"], F["kx"], F["ja"])
    nut = to_dev(np.zeros(dom.ncell))
    gpu.call("falm_sgs", C.byref(dom), abi.F64, abi.SGS_SMAGORINSKY, 0.1, K(to_dev(F["u"])), gpu.ptr(nut),
             K(to_dev(F["kx"])), K(to_dev(F["ja"])), None)
    assert rel_err(from_dev(nut), ref, interior_mask(dom)) < 1e-11
    with pytest.raises(gpu.FalmError):
        gpu.call("falm_sgs", C.byref(dom), abi.F64, abi.SGS_CSM, 0.1, K(to_dev(F["u"])), gpu.ptr(nut),
                 K(to_dev(F["kx"])), K(to_dev(F["ja"])), None)


def _matrix(dom, kind=abi.POISSON_NEUMANN, stretch=0.02):
    _, (xyz, kx, g, ja) = uniform_metrics(dom, stretch=stretch)
    a, md = O.poisson_matrix(dom, kind, g, ja)
    return a, md, g, ja


@pytest.mark.parametrize("kind", [abi.POISSON_NEUMANN, abi.POISSON_OUTLET_XPLUS, abi.POISSON_2D])
def test_poisson_matrix(gpu, kind):
    dom = domain((12, 9, 1 if kind == abi.POISSON_2D else 7))
    a_ref, md_ref, g, ja = _matrix(dom, kind)
    a = to_dev(np.zeros(7 * dom.ncell))
    md = C.c_double()
    gpu.call("falm_poisson_matrix", C.byref(dom), abi.F64, kind, K(to_dev(g)), K(to_dev(ja)), gpu.ptr(a),
             C.byref(md), None)
    assert abs(md.value - md_ref) <= 1e-14 * md_ref
    assert rel_err(from_dev(a), a_ref) < TOL64


def test_mv_residual_dot_norm(gpu):
    dom = domain((20, 11, 9))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS)
    rng = np.random.default_rng(11)
    x, b = rng.standard_normal(dom.ncell), rng.standard_normal(dom.ncell)
    da, dx, db = to_dev(a), to_dev(x), to_dev(b)
    ax, r = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
    gpu.call("falm_mv", C.byref(dom), abi.F64, gpu.ptr(da), gpu.ptr(dx), gpu.ptr(ax), None)
    gpu.call("falm_residual", C.byref(dom), abi.F64, gpu.ptr(da), gpu.ptr(dx), gpu.ptr(db), gpu.ptr(r), None)
    m = interior_mask(dom)
    assert rel_err(from_dev(ax), O.mv(dom, a, x), m) < TOL64
    assert rel_err(from_dev(r), O.residual(dom, a, x, b), m) < TOL64
    out = C.c_double()
    gpu.call("falm_dot", C.byref(dom), abi.F64, gpu.ptr(dx), gpu.ptr(db), C.byref(out), None)
    assert abs(out.value - O.dot(dom, x, b)) < 1e-12 * np.sum(np.abs(x * b)[m])
    gpu.call("falm_norm2", C.byref(dom), abi.F64, gpu.ptr(dx), 1, C.byref(out), None)
    assert abs(out.value - O.norm2(dom, x)) < 1e-13 * O.norm2(dom, x)
    for op in range(6):
        gpu.call("falm_reduce", C.byref(dom), abi.F64, op, gpu.ptr(da), 7, 2, C.byref(out), None)
        assert out.value == O.reduce(dom, op, a, 7, 2)


def test_dot_known_answer(gpu):
    """test/dot.cpp + test/dot.py of the reference: 12^3 box, 1 guide cell."""
    from tests.golden_data import dot_case
    dom, a, b, expect = dot_case()
    out = C.c_double()
    # gc = 1 is below the library's stencil radius; use the same interior
    # embedded in a gc = 2 domain
    dom2 = abi.Domain.make((14, 14, 14), gc=2)
    a2, b2 = np.zeros(14 ** 3), np.zeros(14 ** 3)
    a2.reshape(14, 14, 14)[1:13, 1:13, 1:13] = a.reshape(12, 12, 12)
    b2.reshape(14, 14, 14)[1:13, 1:13, 1:13] = b.reshape(12, 12, 12)
    gpu.call("falm_dot", C.byref(dom2), abi.F64, K(to_dev(a2)), K(to_dev(b2)), C.byref(out), None)
    assert out.value == expect["sumab"]
    gpu.call("falm_norm2", C.byref(dom2), abi.F64, K(to_dev(a2)), 1, C.byref(out), None)
    assert out.value == expect["sumaa"]
    gpu.call("falm_reduce", C.byref(dom2), abi.F64, abi.RED_COL_ABSMAX, K(to_dev(a2)), 1, 0, C.byref(out), None)
    assert out.value == expect["maxa"]


@pytest.mark.parametrize("offset", [(0, 0, 0), (3, 0, 0)])
def test_sor_jacobi_sweeps(gpu, offset):
    dom = domain((13, 10, 8), offset=offset, gshape=(17 + 3, 14, 12))
    a, _, _, _ = _matrix(domain((13, 10, 8)), abi.POISSON_NEUMANN)
    rng = np.random.default_rng(13)
    x, b = rng.standard_normal(dom.ncell), rng.standard_normal(dom.ncell)
    for color in (abi.BLACK, abi.RED):
        xr = x.copy()
        O.sor_sweep(dom, a, xr, b, 1.5, color)
        dx = to_dev(x)
        gpu.call("falm_sor_sweep", C.byref(dom), abi.F64, K(to_dev(a)), gpu.ptr(dx), K(to_dev(b)), 1.5,
                 color, None)
        assert rel_err(from_dev(dx), xr) < TOL64
    xj = np.zeros(dom.ncell)
    O.jacobi_sweep(dom, a, xj, x, b)
    dxj = to_dev(np.zeros(dom.ncell))
    gpu.call("falm_jacobi_sweep", C.byref(dom), abi.F64, K(to_dev(a)), gpu.ptr(dxj), K(to_dev(x)),
             K(to_dev(b)), None)
    assert rel_err(from_dev(dxj), xj) < TOL64


def _faces(types, values=None):
    out = []
    for f, t in enumerate(types):
        fb = abi.FaceBC()
        fb.type = t
        if values and values[f]:
            for m in range(3):
                fb.value[m] = values[f][m]
        out.append(fb)
    return out


@pytest.mark.parametrize("setup", ["windtunnel", "lid"])
def test_bc_fills(gpu, setup):
    dom = domain((14, 9, 8))
    F = _fields(dom, seed=21)
    if setup == "windtunnel":
        faces = _faces([abi.BC_INLET, abi.BC_CONVECTIVE, abi.BC_SLIP, abi.BC_SLIP, abi.BC_WALL0, abi.BC_SLIP],
                       [(1.0, 0, 0)] + [None] * 5)
        noise = abi.InletNoise(0.05, 1234, 17)
        pf = [abi.PBC_NEUMANN, abi.PBC_ZERO] + [abi.PBC_NEUMANN] * 4
    else:
        faces = _faces([abi.BC_REFLECT] * 6, [None, None, None, (1.0, 0, 0), None, None])
        noise = None
        pf = [abi.PBC_NEUMANN] * 6
    u_ref = F["u"].copy()
    O.bc_velocity(dom, faces, u_ref, F["un"], F["xyz"], 2e-3, noise)
    du = to_dev(F["u"])
    farr = (abi.FaceBC * 6)(*faces)
    gpu.call("falm_bc_velocity", C.byref(dom), abi.F64, farr, gpu.ptr(du), K(to_dev(F["un"])),
             K(to_dev(F["xyz"])), 2e-3, C.byref(noise) if noise else None, None)
    assert rel_err(from_dev(du), u_ref) < TOL64
    p = smooth_field(dom, 1, np.random.default_rng(2))
    p_ref = p.copy()
    O.bc_pressure(dom, pf, p_ref)
    dp = to_dev(p)
    gpu.call("falm_bc_pressure", C.byref(dom), abi.F64, (C.c_int * 6)(*pf), gpu.ptr(dp), None)
    assert np.array_equal(from_dev(dp), p_ref)
    uu_ref = F["uu"].copy()
    O.bc_face_flux(dom, [1] * 6, uu_ref)
    duu = to_dev(F["uu"])
    gpu.call("falm_bc_face_flux", C.byref(dom), abi.F64, (C.c_int * 6)(*([1] * 6)), gpu.ptr(duu), None)
    assert np.array_equal(from_dev(duu), uu_ref)


def test_copy_z5(gpu):
    dom = domain((9, 7, 1))
    v = smooth_field(dom, 3, np.random.default_rng(4))
    ref = v.copy()
    O.copy_z5(dom, ref, 3)
    dv = to_dev(v)
    gpu.call("falm_copy_z5", C.byref(dom), abi.F64, gpu.ptr(dv), 3, None)
    assert np.array_equal(from_dev(dv), ref)


@pytest.mark.parametrize("pc", [abi.LS_SOR, abi.LS_JACOBI])
def test_bicgstab_solve(gpu, pc):
    """Same iteration trajectory as the oracle for the first iterations
    (x after a fixed small maxit, where rounding has not been amplified),
    then a tight solve whose answer is fixed by the system, not by where
    the iteration stops."""
    dom = domain((16, 12, 10))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS)
    rng = np.random.default_rng(17)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom)
    m = interior_mask(dom)
    for maxit, tol, bar in ((3, 1e-30, 1e-11), (2000, 1e-13, 1e-9)):
        cfg = abi.solver_config(type=abi.LS_PBICGSTAB, maxit=maxit, tol=tol, pc_type=pc, pc_maxit=2, pc_relax=1.2)
        x_ref, r_ref = np.zeros(dom.ncell), np.zeros(dom.ncell)
        info_ref = O.Solver(dom, cfg).solve(a, x_ref, b, r_ref)
        S = gpu.Solver(dom, cfg)
        dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
        info = S.solve(to_dev(a), dx, to_dev(b), dr)
        if maxit == 3:
            assert info.it == info_ref.it == 3
            assert abs(info.err - info_ref.err) <= 1e-10 * info_ref.err
        else:
            assert info.status == abi.FALM_OK and info_ref.status == abi.FALM_OK
            assert np.sqrt(np.sum(O.residual(dom, a, from_dev(dx), b)[m] ** 2)) / m.sum() < 1e-12
        assert rel_err(from_dev(dx), x_ref) < bar


@pytest.mark.parametrize("stretch", [0.0, 0.02])
@pytest.mark.parametrize("precision", [abi.SOLVE_F64, abi.SOLVE_MIXED])
def test_declared_matrix_paths(gpu, stretch, precision):
    """falm_solver_set_matrix: on a uniform mesh the off-diagonal columns are
    one-index arrays and the solver switches to the compressed operator; on a
    stretched mesh it keeps the seven columns.  Both reproduce the oracle's
    first iterations and solve the system."""
    dom = domain((18, 12, 10))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS, stretch=stretch)
    rng = np.random.default_rng(31)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom)
    m = interior_mask(dom)
    da, db = to_dev(a), to_dev(b)
    if precision == abi.SOLVE_F64:
        cfg = abi.solver_config(maxit=3, tol=1e-30, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2)
        x_ref, r_ref = np.zeros(dom.ncell), np.zeros(dom.ncell)
        info_ref = O.Solver(dom, cfg).solve(a, x_ref, b, r_ref)
        S = gpu.Solver(dom, cfg)
        S.set_matrix(da)
        dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
        info = S.solve(da, dx, db, dr)
        assert info.it == info_ref.it == 3
        assert rel_err(from_dev(dx), x_ref) < 1e-11
    cfg = abi.solver_config(maxit=1000, tol=1e-12, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2, precision=precision)
    S = gpu.Solver(dom, cfg)
    S.set_matrix(da)
    dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
    info = S.solve(da, dx, db, dr)
    assert info.status == abi.FALM_OK
    assert np.sqrt(np.sum(O.residual(dom, a, from_dev(dx), b)[m] ** 2)) / m.sum() <= 1.01e-12


def _four_iterations_both_views(gpu, dom, a, b):
    """Four BiCGStab + one-pass SOR iterations with the caller's seven
    columns and with the declared operator, each against the oracle's."""
    cfg = abi.solver_config(maxit=4, tol=1e-30, pc_type=abi.LS_SOR, pc_maxit=1, pc_relax=1.2)
    x_ref, r_ref = np.zeros(dom.ncell), np.zeros(dom.ncell)
    info_ref = O.Solver(dom, cfg).solve(a, x_ref, b, r_ref)
    da, db = to_dev(a), to_dev(b)
    xs = []
    for declared in (False, True):
        S = gpu.Solver(dom, cfg)
        if declared:
            S.set_matrix(da)
        dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
        info = S.solve(da, dx, db, dr)
        assert info.it == info_ref.it == 4
        assert abs(info.err - info_ref.err) <= 1e-10 * info_ref.err, declared
        assert rel_err(from_dev(dx), x_ref) < 1e-11, declared
        xs.append(from_dev(dx))
    return xs


@pytest.mark.parametrize("stretch", [0.0, 0.02])
@pytest.mark.parametrize("precision", [abi.SOLVE_F64, abi.SOLVE_MIXED])
def test_one_pass_sor_preconditioner(gpu, stretch, precision):
    """pc_maxit = 1 runs each Krylov vector and both SOR half-sweeps as one
    k-marching kernel, which recomputes the black values at the neighbours of
    a tile and alternates p between two buffers.  On a domain of several
    column tiles and k chunks (70 x 10 x 34 interior: 2 x 2 x 3 workgroups)
    the first four iterations follow the oracle's with the caller's seven
    columns and with the declared (compressed, on the uniform mesh) operator,
    and a tight solve reaches its tolerance in both precisions."""
    dom = domain((74, 14, 38))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS, stretch=stretch)
    rng = np.random.default_rng(41)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom)
    m = interior_mask(dom)
    da, db = to_dev(a), to_dev(b)
    if precision == abi.SOLVE_F64:
        xs = _four_iterations_both_views(gpu, dom, a, b)
        if stretch == 0.0:
            # the declared operator (compressed off-diagonals, diagonal from
            # its per-axis classes on this uniform mesh) computes what the
            # seven columns do, bit for bit
            assert np.array_equal(xs[0], xs[1])
    cfg = abi.solver_config(maxit=3000, tol=1e-11, pc_type=abi.LS_SOR, pc_maxit=1, pc_relax=1.0, precision=precision)
    S = gpu.Solver(dom, cfg)
    S.set_matrix(da)
    dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
    info = S.solve(da, dx, db, dr)
    assert info.status == abi.FALM_OK
    assert np.sqrt(np.sum(O.residual(dom, a, from_dev(dx), b)[m] ** 2)) / m.sum() <= 1.01e-11


def test_diagonal_classes_fall_back(gpu):
    """A uniform-mesh matrix whose diagonal is changed at one interior cell
    keeps its one-index off-diagonals but loses the per-class diagonal: the
    declared solver then reads the stored diagonal and still follows the
    oracle, and matches the seven-column solve bit for bit."""
    dom = domain((74, 14, 38))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS, stretch=0.0)
    nx, ny, _ = tuple(dom.shape)
    a[30 + nx * (7 + ny * 20)] *= 1.25
    b = np.random.default_rng(43).standard_normal(dom.ncell) * interior_mask(dom)
    xs = _four_iterations_both_views(gpu, dom, a, b)
    assert np.array_equal(xs[0], xs[1])


@pytest.mark.parametrize("precision", [abi.SOLVE_F64, abi.SOLVE_MIXED])
def test_solve_without_residual_output(gpu, precision):
    """r = NULL: the solve is the same (same x, iterations and err) and only
    the residual field is not returned."""
    dom = domain((18, 12, 10))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS)
    rng = np.random.default_rng(37)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom)
    da, db = to_dev(a), to_dev(b)
    cfg = abi.solver_config(maxit=1000, tol=1e-10, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2, precision=precision)
    out = []
    for with_r in (True, False):
        S = gpu.Solver(dom, cfg)
        S.set_matrix(da)
        dx = to_dev(np.zeros(dom.ncell))
        dr = to_dev(np.zeros(dom.ncell)) if with_r else None
        info = S.solve(da, dx, db, dr)
        assert info.status == abi.FALM_OK
        out.append((from_dev(dx), info.it, info.err, info.outer_it))
        if with_r:
            m = interior_mask(dom)
            r_true = O.residual(dom, a, from_dev(dx), b)
            if precision == abi.SOLVE_MIXED:  # b - A x of the final x
                assert rel_err(from_dev(dr), r_true, m) < 1e-6
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_sor_solver(gpu):
    dom = domain((24, 20, 1))
    a, _, _, _ = _matrix(dom, abi.POISSON_2D, stretch=0.0)
    rng = np.random.default_rng(19)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom)
    b[interior_mask(dom)] -= b[interior_mask(dom)].mean()
    cfg = abi.solver_config(type=abi.LS_SOR, maxit=20000, tol=1e-9, relax=1.2, pc_type=abi.LS_EMPTY)
    x_ref, r_ref = np.zeros(dom.ncell), np.zeros(dom.ncell)
    info_ref = O.Solver(dom, cfg).solve(a, x_ref, b, r_ref)
    dx, dr = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
    info = gpu.Solver(dom, cfg).solve(to_dev(a), dx, to_dev(b), dr)
    assert info.it == info_ref.it
    assert rel_err(from_dev(dx), x_ref) < 1e-10


def test_mixed_solver_meets_tolerance(gpu):
    dom = domain((24, 16, 16))
    a, _, _, _ = _matrix(dom, abi.POISSON_OUTLET_XPLUS)
    rng = np.random.default_rng(23)
    b = rng.standard_normal(dom.ncell) * interior_mask(dom) * 1e-2
    m = interior_mask(dom)
    # defect correction reaches tolerances far below float round-off
    for tol, bar in ((1e-8, None), (1e-12, 1e-7)):
        cfg64 = abi.solver_config(tol=tol, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2)
        cfgmx = abi.solver_config(tol=tol, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2, precision=abi.SOLVE_MIXED)
        x64, r64 = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
        i64 = gpu.Solver(dom, cfg64).solve(to_dev(a), x64, to_dev(b), r64)
        xmx, rmx = to_dev(np.zeros(dom.ncell)), to_dev(np.zeros(dom.ncell))
        S = gpu.Solver(dom, cfgmx)
        imx = S.solve(to_dev(a), xmx, to_dev(b), rmx)
        assert imx.status == abi.FALM_OK and imx.err <= tol
        true = np.sqrt(np.sum(O.residual(dom, a, from_dev(xmx), b)[m] ** 2)) / m.sum()
        assert true <= tol * 1.01
        assert i64.status == abi.FALM_OK
        if bar:
            assert rel_err(from_dev(xmx), from_dev(x64), m) < bar


@pytest.mark.parametrize("cutoff", [0.0, abi.ALM_CUTOFF_ROUNDING])
def test_alm_matches_oracle(gpu, cutoff):
    # the oracle always sums every point at every cell (the reference's
    # all-pairs projection); the rounding cutoff drops terms below 2^-53 of a
    # point's peak, so both settings agree with it to rounding
    from falm_amd import cases
    dom = domain((40, 24, 24))
    h = 1.0 / 8.0
    setup = cases.rotor(h, base=(2.0, 1.5, 1.5), radius=0.5, tsr=7.0, n_ap_per_blade=12,
                        angle_type=2, motion=(np.radians(3.0), 0.8, 0.3), cutoff=cutoff)
    x = 0.0 + (np.arange(dom.shape[0]) - 1.5) * h
    y = (np.arange(dom.shape[1]) - 1.5) * h
    z = (np.arange(dom.shape[2]) - 1.5) * h
    rng = np.random.default_rng(29)
    u = smooth_field(dom, 3, rng, base=1.0, amp=0.2)
    ref = O.Alm(dom, setup, x, y, z)
    ff_ref = np.zeros(3 * dom.ncell)
    ref.step(u, ff_ref, 0.37)
    A = gpu.Alm(dom, setup, x, y, z)
    ff = to_dev(np.zeros(3 * dom.ncell))
    A.step(abi.F64, K(to_dev(u)), ff, 0.37)
    xyz, f, tq, th = A.get()
    xyz_r, f_r, tq_r, th_r = ref.get()
    assert rel_err(xyz, xyz_r) < 1e-14
    assert rel_err(f, f_r) < 1e-12
    assert rel_err(tq, tq_r) < 1e-12 and rel_err(th, th_r) < 1e-12
    assert rel_err(from_dev(ff), ff_ref) < 1e-12
