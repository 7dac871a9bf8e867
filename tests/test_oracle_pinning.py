"""Pins the CPU oracle (the checker of the HIP library) against what the
reference's own tests hold and what the mathematics fixes.  Every check is
chosen so that a plausible restatement mistake — a dropped term, a wrong
sign, index or coefficient order, a transposed operand — fails it;
tests/test_oracle_mutants.py plants such mistakes in a copy of the oracle
and shows that these checks catch every one of them."""
import math

import numpy as np
import pytest

from falm_amd import _abi as abi
from falm_amd import cases
from oracle import pyoracle as O
from tests.golden_data import dot_case, heat1d
from tests.helpers import domain, interior_mask


# ---------------------------------------------------------------- reductions
def test_dot_norm_max_known_answer():
    """test/dot.cpp + test/dot.py: sum(ab), sum(a^2) over the interior."""
    dom, a, b, e = dot_case()
    assert O.dot(dom, a, b) == e["sumab"] == 14677400
    assert O.norm2(dom, a) == e["sumaa"] == 68505400
    assert O.reduce(dom, abi.RED_COL_ABSMAX, a) == e["maxa"]
    assert O.reduce(dom, abi.RED_COL_ABSMAX, b) == e["maxb"]


# -------------------------------------------------------------- linear solve
def _heat1d_system():
    """test/heat1d.cpp:18-49: ghost-Dirichlet 1-D Poisson, T_W = 0, T_E = 100,
    stored in the current [c,w,e,s,n,b,t] column order, scaled by max|a_c|
    (matrix and right-hand side, :51-53)."""
    h = heat1d()
    n, tw, te = h["n"], h["TW"], h["TE"]
    dom = domain((n, 1, 1))
    dx = h["L"] / n
    a = np.zeros(7 * dom.ncell)
    b = np.zeros(dom.ncell)
    nc = dom.ncell
    for i in range(n):
        c = (i + 2) + dom.shape[0] * (2 + dom.shape[1] * 2)
        if i == 0:
            ae, aw, ac, bc = 1 / dx ** 2, 0.0, -(1 / dx ** 2 + 2 / dx ** 2), -2 * tw / dx ** 2
        elif i == n - 1:
            ae, aw, ac, bc = 0.0, 1 / dx ** 2, -(1 / dx ** 2 + 2 / dx ** 2), -2 * te / dx ** 2
        else:
            ae, aw, ac, bc = 1 / dx ** 2, 1 / dx ** 2, -2 / dx ** 2, 0.0
        a[c], a[c + nc], a[c + 2 * nc], b[c] = ac, aw, ae, bc
    md = O.reduce(dom, abi.RED_COL_ABSMAX, a, 7, 0)
    O.scale(a, 1.0 / md)
    O.scale(b, 1.0 / md)
    return dom, a, b, np.array(h["T"])


def test_heat1d_exact_solution_has_zero_residual():
    """The linear profile T_i = 100(i+1/2)/N solves the discrete system; a
    swapped w/e column or a wrong sign in the residual breaks this."""
    dom, a, b, T = _heat1d_system()
    x = np.zeros(dom.ncell)
    m = interior_mask(dom)
    x[m] = T
    r = O.residual(dom, a, x, b)
    assert np.max(np.abs(r[m])) < 1e-12


@pytest.mark.parametrize("pc", [abi.LS_SOR, abi.LS_JACOBI])
def test_heat1d_bicgstab(pc):
    """heat1d.cpp's solver set-up: PBiCGStab, tol 1e-9, SOR PC (5, 1.5)."""
    dom, a, b, T = _heat1d_system()
    x, r = np.zeros(dom.ncell), np.zeros(dom.ncell)
    cfg = abi.solver_config(type=abi.LS_PBICGSTAB, maxit=10000, tol=1e-12, pc_type=pc, pc_maxit=5, pc_relax=1.5)
    info = O.Solver(dom, cfg).solve(a, x, b, r)
    assert info.status == abi.FALM_OK
    assert np.max(np.abs(x[interior_mask(dom)] - T)) < 1e-7


def _textbook_laplacian(n, outlet, zcouple=True):
    """7-point Laplacian of a uniform grid (spacing h): Neumann walls by
    dropping the coupling, an x+ Dirichlet wall (ghost p = 0) keeping it,
    divided by max|a_c|.  Built with scipy, independent of the oracle."""
    import scipy.sparse as sp
    nx, ny, nz = n
    N = nx * ny * nz
    idx = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros(N)
    dirs = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)) + (((0, 0, 1), (0, 0, -1)) if zcouple else ())
    for (di, dj, dk) in dirs:
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    ii, jj, kk = i + di, j + dj, k + dk
                    me = idx[k, j, i]
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        rows.append(me)
                        cols.append(idx[kk, jj, ii])
                        vals.append(1.0)
                        diag[me] -= 1.0
                    elif outlet and ii == nx:
                        diag[me] -= 1.0  # coupling to the p = 0 ghost
    A = sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr() + sp.diags(diag)
    return A / np.max(np.abs(diag))


@pytest.mark.parametrize("kind", [abi.POISSON_NEUMANN, abi.POISSON_OUTLET_XPLUS])
def test_poisson_matrix_is_textbook_laplacian(kind):
    n = (7, 5, 4)
    dom = domain(n)
    h = 0.1
    x = (np.arange(dom.shape[0]) - 1.5) * h
    y = (np.arange(dom.shape[1]) - 1.5) * h
    z = (np.arange(dom.shape[2]) - 1.5) * h
    hh = [np.full(dom.shape[m], h) for m in range(3)]
    _, _, g, ja = O.metrics(dom, x, y, z, *hh)
    a, md = O.poisson_matrix(dom, kind, g, ja)
    assert abs(md - 6.0 / h ** 2) < 1e-9 * md
    A = _textbook_laplacian(n, kind == abi.POISSON_OUTLET_XPLUS)
    # apply both to random interior vectors
    rng = np.random.default_rng(0)
    m = interior_mask(dom)
    for _ in range(3):
        v = rng.standard_normal(m.sum())
        xv = np.zeros(dom.ncell)
        xv[m] = v
        assert np.max(np.abs(O.mv(dom, a, xv)[m] - A @ v)) < 1e-13
    if kind == abi.POISSON_NEUMANN:
        rows = a.reshape(7, -1)[:, m].sum(axis=0)
        assert np.max(np.abs(rows)) < 1e-15  # zero row sums


def test_bicgstab_and_sor_match_direct_solve():
    import scipy.sparse.linalg as spla
    n = (6, 5, 4)
    dom = domain(n)
    h = 1.0
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    _, _, g, ja = O.metrics(dom, *cen, *[np.full(dom.shape[m], h) for m in range(3)])
    a, _ = O.poisson_matrix(dom, abi.POISSON_OUTLET_XPLUS, g, ja)
    A = _textbook_laplacian(n, True)
    rng = np.random.default_rng(1)
    m = interior_mask(dom)
    bv = rng.standard_normal(m.sum())
    exact = spla.spsolve(A.tocsc(), bv)
    b = np.zeros(dom.ncell)
    b[m] = bv
    for cfg in (abi.solver_config(type=abi.LS_PBICGSTAB, tol=1e-14, maxit=500, pc_type=abi.LS_SOR),
                abi.solver_config(type=abi.LS_PBICGSTAB, tol=1e-14, maxit=500, pc_type=abi.LS_JACOBI),
                abi.solver_config(type=abi.LS_SOR, tol=1e-14, maxit=20000, relax=1.2, pc_type=abi.LS_EMPTY),
                abi.solver_config(type=abi.LS_JACOBI, tol=1e-13, maxit=200000, pc_type=abi.LS_EMPTY)):
        x, r = np.zeros(dom.ncell), np.zeros(dom.ncell)
        info = O.Solver(dom, cfg).solve(a, x, b, r)
        assert info.status == abi.FALM_OK, cfg.type
        assert np.max(np.abs(x[m] - exact)) < 1e-9 * np.max(np.abs(exact)), cfg.type


def test_red_black_sweep_is_colour_gauss_seidel():
    """One red-black half sweep equals a Jacobi update restricted to the
    cells of that colour (parity of i+j+k+sum(offset)); the other colour is
    untouched.  Checked against a numpy shift implementation."""
    n = (6, 5, 4)
    for offset in ((0, 0, 0), (1, 0, 0), (1, 1, 1)):
        dom = domain(n, offset=offset, gshape=tuple(v + 4 + 2 for v in n))
        rng = np.random.default_rng(2)
        a = rng.uniform(0.1, 1.0, 7 * dom.ncell)
        a[:dom.ncell] = -6.0
        x = rng.standard_normal(dom.ncell)
        b = rng.standard_normal(dom.ncell)
        nx, ny, nz = dom.shape
        X = x.reshape(nz, ny, nx)
        A = a.reshape(7, nz, ny, nx)
        B = b.reshape(nz, ny, nx)
        for color in (0, 1):
            ref = X.copy()
            s = slice(2, -2)
            ax = (A[0] * X)[s, s, s]
            nb = {1: X[s, s, 1:-3], 2: X[s, s, 3:-1], 3: X[s, 1:-3, s], 4: X[s, 3:-1, s], 5: X[1:-3, s, s],
                  6: X[3:-1, s, s]}
            for q, v in nb.items():
                ax = ax + A[q][s, s, s] * v
            kk, jj, ii = np.meshgrid(np.arange(2, nz - 2), np.arange(2, ny - 2), np.arange(2, nx - 2), indexing="ij")
            par = (ii + jj + kk + sum(offset)) % 2 == color
            upd = X[s, s, s] + 1.3 * (B[s, s, s] - ax) / A[0][s, s, s]
            ref[s, s, s] = np.where(par, upd, X[s, s, s])
            got = x.copy()
            O.sor_sweep(dom, a, got, b, 1.3, color)
            assert np.max(np.abs(got - ref.ravel())) < 1e-13


# --------------------------------------------------------------- advection
def _uniform(dom, h):
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    wid = [np.full(dom.shape[m], h) for m in range(3)]
    xyz, kx, g, ja = O.metrics(dom, *cen, *wid)
    return cen, xyz, kx, g, ja


def _grid(dom, cen):
    nx, ny, nz = dom.shape
    X = np.broadcast_to(cen[0][None, None, :], (nz, ny, nx)).ravel()
    Y = np.broadcast_to(cen[1][None, :, None], (nz, ny, nx)).ravel()
    Z = np.broadcast_to(cen[2][:, None, None], (nz, ny, nx)).ravel()
    return X, Y, Z


@pytest.mark.parametrize("scheme", [abi.ADV_UPWIND3, abi.ADV_UTOPIA, abi.ADV_KK, abi.ADV_QUICK])
def test_upwind3_family_exact_for_quadratics(scheme):
    """With uniform face flux U (uu = U*J*kx) every Upwind3-family scheme
    gives adv = U . grad(phi) exactly for quadratic phi (the fourth difference
    vanishes), so ua = un - dt*U.grad(phi) + dt*ff.  QUICK too: its face
    value (3 phi_d + 6 phi_u - phi_uu)/8 from the upwind side
    (src/FalmCFDScheme.h:33-66) is the quadratic through three cells, exact
    at the face, and the face difference of exact face values of a quadratic
    is its exact derivative.  U has both signs across the axes."""
    dom = domain((9, 8, 7))
    h = 0.125
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    U = np.array([0.7, -0.4, 0.25])
    # three quadratic components
    phis = [0.3 + X * X - 0.5 * X * Y + 0.2 * Z * Z + X,
            1.0 - Y * Y + 0.3 * X * Z + 0.1 * Y,
            0.5 * Z * Z - X * Y + 2.0 * Z]
    grads = [(2 * X - 0.5 * Y + 1, -0.5 * X, 0.4 * Z),
             (0.3 * Z, -2 * Y + 0.1, 0.3 * X),
             (-Y, -X, Z + 2.0)]
    u = np.concatenate(phis)
    uu = np.concatenate([np.full(dom.ncell, U[d] * h * h) for d in range(3)])  # J*kx = h^2
    nut = np.zeros(dom.ncell)
    ff = np.concatenate([np.full(dom.ncell, 0.1 * d) for d in range(3)])
    un = u + 1.0
    dt = 0.01
    ua = O.pseudo_u(dom, scheme, 0.0, dt, un, u, uu, nut, kx, g, ja, ff)
    m = interior_mask(dom)
    for d in range(3):
        adv = sum(U[e] * grads[d][e] for e in range(3))
        expect = un[d * dom.ncell:(d + 1) * dom.ncell] + dt * (-adv + 0.1 * d)
        assert np.max(np.abs(ua[d * dom.ncell:(d + 1) * dom.ncell][m] - expect[m])) < 1e-12


def test_upwind3_dissipation_term_is_fourth_difference():
    """For phi = x^4 (unit spacing, U = 1 through the cell and its faces) the
    Upwind3 result departs from U dphi/dx by |U| * (4th difference)/24 * ...:
    the central part of x^4 at x = 0 vanishes and the fourth difference of
    x^4 is 24, so adv = |U|*24/24/h (per unit spacing) = 1/h * 1."""
    dom = domain((9, 5, 5))
    h = 1.0
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    x0 = cen[0][6]
    phi = (X - x0) ** 4
    u = np.concatenate([phi, np.zeros(dom.ncell), np.zeros(dom.ncell)])
    # the velocity used for |U| is the component itself: make it 1 at x0 by
    # adding 1 (a constant does not change differences)
    u[:dom.ncell] += 1.0
    uu = np.concatenate([np.full(dom.ncell, 1.0), np.zeros(dom.ncell), np.zeros(dom.ncell)])
    zero = np.zeros(3 * dom.ncell)
    ua = O.pseudo_u(dom, abi.ADV_UPWIND3, 0.0, 1.0, zero, u, uu, np.zeros(dom.ncell), kx, g, ja, zero)
    c = 6 + dom.shape[0] * (2 + dom.shape[1] * 2)
    # central part: 0.5*(UDL+UDR) with UDL = (-16+27*1-0+1)/24 = 0.5, UDR = (-1+0-27+16)/24 = -0.5
    assert abs(-ua[c] - 1.0 * 24 / 24) < 1e-12


def test_diffusion_exact_for_cubics():
    """Central diffusion with constant nu_t: vis = (1/Re + nu_t) lap(phi),
    exact for cubic phi on a uniform grid."""
    dom = domain((8, 7, 9))
    h = 0.2
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    phi = X ** 3 - 2 * X * Y * Y + Z ** 3 / 3 + X * Y * Z
    lap = 6 * X - 4 * X + 2 * Z
    u = np.concatenate([phi, phi, phi])
    uu = np.zeros(3 * dom.ncell)  # no advection
    nut = np.full(dom.ncell, 0.01)
    ReI = 1e-3
    zero = np.zeros(3 * dom.ncell)
    ua = O.pseudo_u(dom, abi.ADV_UPWIND3, ReI, 1.0, zero, u, uu, nut, kx, g, ja, zero)
    m = interior_mask(dom)
    # Upwind3 |U| term: |u*kx| * 4th difference of a cubic = 0
    for d in range(3):
        got = ua[d * dom.ncell:(d + 1) * dom.ncell][m]
        assert np.max(np.abs(got - (ReI + 0.01) * lap[m])) < 1e-11


def test_upwind1_closed_form_for_quadratics():
    """Upwind1 (src/FalmCFDScheme.h:8-31) with uniform U: the flux through a
    face is U times the upwind cell value, so for quadratic phi
    adv = U . grad(phi) - (h/2) sum_e |U_e| d2phi/dx_e2 exactly: linear fields
    are advected exactly and the error of a quadratic has the upwind sign
    (a downwind flux would flip it)."""
    dom = domain((9, 8, 7))
    h = 0.125
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    U = np.array([0.7, -0.4, 0.25])
    phis = [0.3 + X * X - 0.5 * X * Y + 0.2 * Z * Z + X,
            1.0 - Y * Y + 0.3 * X * Z + 0.1 * Y,
            0.5 * Z * Z - X * Y + 2.0 * Z + 0.4 * X]
    grads = [(2 * X - 0.5 * Y + 1, -0.5 * X, 0.4 * Z),
             (0.3 * Z, -2 * Y + 0.1, 0.3 * X),
             (-Y + 0.4, -X, Z + 2.0)]
    curv = [(2.0, 0.0, 0.4), (0.0, -2.0, 0.0), (0.0, 0.0, 1.0)]  # d2phi/dx_e2
    u = np.concatenate(phis)
    uu = np.concatenate([np.full(dom.ncell, U[d] * h * h) for d in range(3)])
    ff = np.concatenate([np.full(dom.ncell, 0.1 * d) for d in range(3)])
    un = u - 0.5
    dt = 0.01
    ua = O.pseudo_u(dom, abi.ADV_UPWIND1, 0.0, dt, un, u, uu, np.zeros(dom.ncell), kx, g, ja, ff)
    m = interior_mask(dom)
    for d in range(3):
        adv = sum(U[e] * grads[d][e] - 0.5 * h * abs(U[e]) * curv[d][e] for e in range(3))
        expect = un[d * dom.ncell:(d + 1) * dom.ncell] + dt * (-adv + 0.1 * d)
        assert np.max(np.abs(ua[d * dom.ncell:(d + 1) * dom.ncell][m] - expect[m])) < 1e-12


def test_quick_closed_form_for_cubics():
    """QUICK's upwind correction is a third difference, so quadratics cannot
    see it.  For a cubic the upwind-side quadratic interpolant misses the
    face value by sign(U) (d3phi/dx3) h^3/16 on every face; with d3phi/dx3
    constant these cancel between the two faces of a cell and
    adv = sum_e U_e (dphi/dx_e + (h^2/24) d3phi/dx_e3) exactly.  A wrong sign
    or weight in the |U| terms leaves a multiple of the third difference."""
    dom = domain((9, 8, 7))
    h = 0.125
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    U = np.array([0.7, -0.4, 0.25])
    phis = [X ** 3 - 0.5 * X * Y * Y + 0.3 * Z ** 3 + X * Y * Z, 1 - Y ** 3 + 0.2 * X * X * Z,
            0.5 * Z ** 3 - X * Y + 0.4 * X ** 3]
    grads = [(3 * X * X - 0.5 * Y * Y + Y * Z, -X * Y + X * Z, 0.9 * Z * Z + X * Y),
             (0.4 * X * Z, -3 * Y * Y, 0.2 * X * X),
             (-Y + 1.2 * X * X, -X, 1.5 * Z * Z)]
    d3 = [(6.0, 0.0, 1.8), (0.0, -6.0, 0.0), (2.4, 0.0, 3.0)]
    u = np.concatenate(phis)
    uu = np.concatenate([np.full(dom.ncell, U[d] * h * h) for d in range(3)])
    zero = np.zeros(3 * dom.ncell)
    ua = O.pseudo_u(dom, abi.ADV_QUICK, 0.0, 1.0, zero, u, uu, np.zeros(dom.ncell), kx, g, ja, zero)
    m = interior_mask(dom)
    for d in range(3):
        adv = sum(U[e] * (grads[d][e] + h * h / 24 * d3[d][e]) for e in range(3))
        assert np.max(np.abs(-ua[d * dom.ncell:(d + 1) * dom.ncell][m] - adv[m])) < 1e-12


@pytest.mark.parametrize("cfl", [0.1, 0.3])
def test_upwind3_explicit_euler_amplification(cfl):
    """One pseudo-velocity step (explicit Euler, src/FalmCFD.cpp:5-67) of a
    transverse velocity mode v = cos(j theta) carried by a uniform stream U
    along x multiplies the mode by the von Neumann factor of Upwind3
    (src/FalmCFDScheme.h:134-165):
        G = 1 - nu [ i (26 sin theta - sin 2 theta)/24 + (2 - 2 cos theta)^2/24 ],
    nu = U dt / h.  |G| > 1 for long waves at any CFL (the fourth-difference
    damping is O(theta^4), the Euler error O(nu^2 theta^2)): the worst mode
    grows by 2.4 % per step at CFL 0.3 (e-folding in 42 steps) and by 0.12 %
    at CFL 0.1 (808 steps), which is why long inviscid runs at CFL 0.3
    overflow and runs at CFL 0.1 do not within 1000 steps."""
    n = 48
    dom = domain((n, 3, 3))
    h = 1.0
    cen, xyz, kx, g, ja = _uniform(dom, h)
    nx, ny, nz = dom.shape
    I = np.broadcast_to(np.arange(nx)[None, None, :], (nz, ny, nx)).ravel().astype(float)
    U = 1.0
    dt = cfl * h / U
    m = interior_mask(dom)
    for theta in (2 * np.pi / 24, 2 * np.pi / 12, 1.08, 2 * np.pi / 5, 2 * np.pi / 3):
        v = np.cos(I * theta)
        u = np.concatenate([np.full(dom.ncell, U), v, np.zeros(dom.ncell)])
        uu = np.concatenate([np.full(dom.ncell, U * h * h), np.zeros(dom.ncell), np.zeros(dom.ncell)])
        zero = np.zeros(3 * dom.ncell)
        ua = O.pseudo_u(dom, abi.ADV_UPWIND3, 0.0, dt, u, u, uu, np.zeros(dom.ncell), kx, g, ja, zero)
        G = 1.0 - cfl * (1j * (26 * np.sin(theta) - np.sin(2 * theta)) / 24 + (2 - 2 * np.cos(theta)) ** 2 / 24)
        expect = G.real * np.cos(I * theta) - G.imag * np.sin(I * theta)
        got = ua[dom.ncell:2 * dom.ncell]
        assert np.max(np.abs(got[m] - expect[m])) < 1e-13, theta
        assert np.max(np.abs(ua[:dom.ncell][m] - U)) < 1e-14  # the carrier is untouched
    th = np.linspace(1e-3, np.pi, 20001)
    gmax = np.max(np.abs(1.0 - cfl * (1j * (26 * np.sin(th) - np.sin(2 * th)) / 24 + (2 - 2 * np.cos(th)) ** 2 / 24)))
    efold = 1.0 / np.log(gmax)
    if cfl == 0.3:
        assert 1.02 < gmax < 1.03 and efold < 50
    else:
        assert 1.0 < gmax < 1.002 and efold > 700


def test_quick_and_upwind1_constant_field():
    """Any consistent scheme gives adv = 0 for a constant field under a
    constant (divergence-free) face flux."""
    dom = domain((7, 6, 6))
    cen, xyz, kx, g, ja = _uniform(dom, 0.1)
    u = np.full(3 * dom.ncell, 2.5)
    uu = np.concatenate([np.full(dom.ncell, v) for v in (0.3, -0.2, 0.1)])
    zero = np.zeros(3 * dom.ncell)
    m = interior_mask(dom, 3)
    for s in (abi.ADV_QUICK, abi.ADV_UPWIND1):
        ua = O.pseudo_u(dom, s, 1e-3, 1.0, zero, u, uu, np.zeros(dom.ncell), kx, g, ja, zero)
        assert np.max(np.abs(ua[m])) < 1e-12


# ---------------------------------------------------------- face flux, SGS
def test_u_to_uu_linear_exact_and_divergence():
    """For linear u the face flux is J*kx*u at the face centre exactly, and
    the divergence equals J*div(u)/J = du/dx+dv/dy+dw/dz."""
    dom = domain((6, 7, 5))
    h = 0.25
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    G = np.array([[0.3, -1.0, 0.2], [0.7, 0.1, -0.4], [0.05, 0.6, -0.4]])
    u = np.concatenate([0.5 + G[d, 0] * X + G[d, 1] * Y + G[d, 2] * Z for d in range(3)])
    uu = O.u_to_uu(dom, u, kx, ja)
    nx, ny, nz = dom.shape
    fm = np.zeros((nz, ny, nx), dtype=bool)
    fm[1:nz - 2, 1:ny - 2, 1:nx - 2] = True
    fm = fm.ravel()
    for d in range(3):
        Xf = [X, Y, Z]
        Xf = [Xf[e] + (0.5 * h if e == d else 0.0) for e in range(3)]
        face_u = 0.5 + G[d, 0] * Xf[0] + G[d, 1] * Xf[1] + G[d, 2] * Xf[2]
        got = uu[d * dom.ncell:(d + 1) * dom.ncell]
        assert np.max(np.abs(got[fm] - h * h * face_u[fm])) < 1e-13
    div = O.divergence(dom, uu, ja)
    m = interior_mask(dom)
    assert np.max(np.abs(div[m] - np.trace(G))) < 1e-12


def test_smagorinsky_linear_field():
    """nu_t = (Cs*Delta)^2 * sqrt(2 S_ij S_ij) with Delta = J^(1/3) and the
    strain rate S = (G + G^T)/2 of a linear field (central differences are
    exact)."""
    dom = domain((6, 6, 6))
    h = 0.1
    cen, xyz, kx, g, ja = _uniform(dom, h)
    X, Y, Z = _grid(dom, cen)
    G = np.array([[0.3, -1.2, 0.2], [0.7, 0.1, -0.45], [0.05, 0.6, -0.4]])
    u = np.concatenate([G[d, 0] * X + G[d, 1] * Y + G[d, 2] * Z for d in range(3)])
    nut = O.sgs(dom, abi.SGS_SMAGORINSKY, 0.1, u, kx, ja)
    S = 0.5 * (G + G.T)
    expect = (0.1 * h) ** 2 * math.sqrt(2.0 * np.sum(S * S))
    m = interior_mask(dom)
    assert np.max(np.abs(nut[m] - expect)) < 1e-14


def test_projection_makes_divergence_free():
    """Fractional-step identity on the lid cavity matrix: rhs from u*, solve
    A p = rhs tightly, project -> div(uu) ~ 0 (cell and face projection use the
    same gradient as the Poisson operator)."""
    dom = domain((10, 9, 8))
    h = 0.1
    cen, xyz, kx, g, ja = _uniform(dom, h)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(3 * dom.ncell)
    uu = O.u_to_uu(dom, u, kx, ja)
    O.bc_face_flux(dom, [1] * 6, uu)  # zero wall flux (app/lid3d)
    a, md = O.poisson_matrix(dom, abi.POISSON_NEUMANN, g, ja)
    dt = 0.01
    rhs = O.poisson_rhs(dom, uu, ja, dt, md)
    p, r = np.zeros(dom.ncell), np.zeros(dom.ncell)
    cfg = abi.solver_config(tol=1e-15, maxit=3000, pc_type=abi.LS_SOR, pc_maxit=2, pc_relax=1.2)
    O.Solver(dom, cfg).solve(a, p, rhs, r)
    O.bc_pressure(dom, [abi.PBC_NEUMANN] * 6, p)
    O.project_p(dom, dt, u, u.copy(), uu, uu.copy(), p, kx, g)
    div = O.divergence(dom, uu, ja)
    m = interior_mask(dom)
    div0 = O.divergence(dom, O.u_to_uu(dom, rng.standard_normal(3 * dom.ncell), kx, ja), ja)
    assert np.max(np.abs(div[m])) < 1e-9 * np.max(np.abs(div0[m]))


# ----------------------------------------------------------------- boundaries
def test_bc_windtunnel_closed_forms():
    dom = domain((6, 5, 4))
    rng = np.random.default_rng(6)
    u = rng.standard_normal(3 * dom.ncell)
    uprev = rng.standard_normal(3 * dom.ncell)
    cen, xyz, kx, g, ja = _uniform(dom, 0.5)
    faces = []
    for t, v in ((abi.BC_INLET, (1.0, 0, 0)), (abi.BC_CONVECTIVE, None), (abi.BC_SLIP, None), (abi.BC_SLIP, None),
                 (abi.BC_WALL0, None), (abi.BC_SLIP, None)):
        f = abi.FaceBC()
        f.type = t
        if v:
            for q in range(3):
                f.value[q] = v[q]
        faces.append(f)
    out = u.copy()
    O.bc_velocity(dom, faces, out, uprev, xyz, 0.1, None)
    nx, ny, nz = dom.shape
    nc = dom.ncell
    U = out.reshape(3, nz, ny, nx)
    P = uprev.reshape(3, nz, ny, nx)
    U0 = u.reshape(3, nz, ny, nx)
    s = slice(2, -2)
    # inlet: both ghost layers (1,0,0)
    assert np.all(U[0, s, s, 0:2] == 1.0) and np.all(U[1:, s, s, 0:2] == 0.0)
    # convective outlet: u_o = up_o - dt*(3up_o - 4up_i1 + up_i2)/(2dx); o2 = 2 u_o - u_i1
    o1 = P[:, s, s, nx - 2] - 0.1 * (3 * P[:, s, s, nx - 2] - 4 * P[:, s, s, nx - 3] + P[:, s, s, nx - 4]) / (2 * 0.5)
    assert np.max(np.abs(U[:, s, s, nx - 2] - o1)) < 1e-14
    assert np.max(np.abs(U[:, s, s, nx - 1] - (2 * o1 - U0[:, s, s, nx - 3]))) < 1e-14
    # slip y-: tangential copied, normal zero; second layer mirrors
    assert np.array_equal(U[0, s, 1, s], U0[0, s, 2, s]) and np.all(U[1, s, 1, s] == 0.0)
    assert np.array_equal(U[1, s, 0, s], -U0[1, s, 2, s])
    # wall z-: zero and odd reflection of the first interior layer
    assert np.all(U[:, 1, s, s] == 0.0) and np.array_equal(U[:, 0, s, s], -U0[:, 2, s, s])
    # slip z+: w = 0
    assert np.all(U[2, nz - 2, s, s] == 0.0)


def test_bc_lid_reflection_averages_to_wall_velocity():
    """No-slip reflection: the mean of the first ghost and first interior
    cell is the wall velocity (lid (1,0,0) on y+, 0 elsewhere)."""
    dom = domain((5, 6, 4))
    u = np.random.default_rng(7).standard_normal(3 * dom.ncell)
    faces = []
    for f in range(6):
        fb = abi.FaceBC()
        fb.type = abi.BC_REFLECT
        if f == 3:
            fb.value[0] = 1.0
        faces.append(fb)
    out = u.copy()
    O.bc_velocity(dom, faces, out, u, u, 0.0, None)
    nx, ny, nz = dom.shape
    U = out.reshape(3, nz, ny, nx)
    s = slice(2, -2)
    assert np.allclose(0.5 * (U[:, s, ny - 2, s] + U[:, s, ny - 3, s])[0], 1.0, atol=1e-15)
    assert np.allclose(0.5 * (U[:, s, 1, s] + U[:, s, 2, s]), 0.0, atol=1e-15)
    assert np.allclose(0.5 * (U[:, s, 0, s] + U[:, s, 3, s]), 0.0, atol=1e-15)


# ------------------------------------------------------------------ inlet RNG
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors (Random123 kat_vectors)."""
    import ctypes as C
    L = O.lib()
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_philox4x32_10.restype = None
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, expect in kat:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        L.orc_philox4x32_10(c, k)
        assert tuple(c) == expect


def test_inlet_noise_is_standard_normal():
    z = np.array([O.inlet_normals(99, s, a, b) for s in range(20) for a in range(10) for b in range(10)]).ravel()
    assert abs(z.mean()) < 0.06 and abs(z.std() - 1.0) < 0.05


# ------------------------------------------------------------------------ ALM
def test_gaussian_projection_mass_and_centroid():
    """sum_cells eta*exp(-r^2/eps^2) dV = 1 for eps >= 2h (the Gaussian
    quadrature is spectrally accurate), and the projected force has the
    point's position as its centroid."""
    import ctypes as C
    h = 0.1
    dom = domain((40, 40, 40))
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    P = np.array([2.013, 1.987, 2.041])
    F = np.array([1.0, -2.0, 0.5])
    ff = np.zeros(3 * dom.ncell)
    L = O.lib()
    L.orc_alm_project.restype = None
    L.orc_alm_project.argtypes = [C.POINTER(abi.Domain)] + [C.c_void_p] * 3 + [C.c_int, C.c_void_p, C.c_void_p,
                                                                               C.c_double, C.c_void_p]
    for eps in (2 * h, 3 * h):
        L.orc_alm_project(C.byref(dom), *[O.ptr(c) for c in cen], 1, O.ptr(P.copy()), O.ptr(F.copy()), eps, O.ptr(ff))
        X, Y, Z = _grid(dom, cen)
        for d in range(3):
            col = ff[d * dom.ncell:(d + 1) * dom.ncell]
            assert abs(col.sum() * h ** 3 - F[d]) < 1e-9 * abs(F[d])
        w = ff[:dom.ncell]
        assert abs((w * X).sum() / w.sum() - P[0]) < 1e-9
        assert abs((w * Z).sum() / w.sum() - P[2]) < 1e-9


def test_polar_lookup_is_piecewise_linear_with_clamps():
    import ctypes as C
    attack = np.arange(-10.0, 11.0, 1.0)
    nap = 3
    rng = np.random.default_rng(8)
    cl = rng.standard_normal(attack.size * nap)
    cd = rng.uniform(0, 1, attack.size * nap)
    L = O.lib()
    L.orc_polar_lookup.restype = None
    L.orc_polar_lookup.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
    for ap in range(nap):
        col_l = cl.reshape(attack.size, nap)[:, ap]
        col_d = cd.reshape(attack.size, nap)[:, ap]
        for att in np.concatenate([rng.uniform(-14, 14, 50), attack]):
            ol, od = C.c_double(), C.c_double()
            L.orc_polar_lookup(O.ptr(attack), attack.size, O.ptr(cl), O.ptr(cd), nap, ap, att, C.byref(ol),
                               C.byref(od))
            assert abs(ol.value - np.interp(att, attack, col_l)) < 1e-13
            assert abs(od.value - np.interp(att, attack, col_d)) < 1e-13


def test_build_ap_matches_interpolation():
    rr, ch, tw, attack, cl, cd = cases.blade_table(0.5, 0.08, 0.02, 20.0, 0.0, n_station=9)
    cl = cl * np.linspace(1.0, 1.3, cl.shape[0])[:, None]
    out = O.build_ap(rr, ch, tw, cl, cd, 1, 3, 16, 0.5, 0.05)
    host = cases.build_ap_host(rr, ch, tw, cl, cd, 1, 3, 16, 0.5, 0.05)
    for a, b in zip(out, host):
        assert np.max(np.abs(a - b)) < 1e-14
    r = out[0]
    assert np.allclose(r[:16], (np.arange(16) + 0.5) * (0.45 / 16) + 0.05)
    assert np.allclose(out[1], np.interp(r, rr, ch)) and np.allclose(out[2], np.interp(r, rr, tw))


def _one_point_rotor(h, tsr, twist, pitch_type=0, motion=(0, 0, 0), yaw_deg=0.0, angle_offset_deg=0.0):
    setup = cases.rotor(h, base=(2.0, 2.0, 2.0), radius=0.5, tsr=tsr, n_blade=1, n_ap_per_blade=1,
                        twist=(twist, twist), chord=(0.1, 0.1), angle_type=pitch_type, motion=motion,
                        yaw_deg=yaw_deg, angle_offset_deg=angle_offset_deg)
    return setup


def test_alm_blade_element_closed_form():
    """One blade, one point at r = R/2, uniform inflow U along x: the point
    sits at hub + r(cos th, sin th) in the (y, z) plane, the inflow angle is
    atan(U/(Omega r)), lift from the synthetic polar (2 pi alpha) and the
    force on the fluid points upstream (thrust > 0, ff_x < 0), with
    torque = |f_t| r."""
    h = 1.0 / 16
    dom = domain((64, 64, 64))
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    setup = _one_point_rotor(h, tsr=4.0, twist=20.0)
    A = O.Alm(dom, setup, *cen)
    u = np.concatenate([np.ones(dom.ncell), np.zeros(dom.ncell), np.zeros(dom.ncell)])
    ff = np.zeros(3 * dom.ncell)
    t = 0.3
    A.step(u, ff, t)
    xyz, f, tq, th = A.get()
    Om = 4.0 * 1.0 / 0.5
    r = 0.25
    theta = Om * t
    assert np.allclose(xyz[0], [2.0, 2.0 + r * math.cos(theta), 2.0 + r * math.sin(theta)], atol=1e-14)
    phi = math.atan(1.0 / (Om * r))
    alpha = math.degrees(phi) - 20.0          # 6.57 deg: inside the linear range
    table = np.arange(-45.0, 46.0, 1.0)       # the tabulated synthetic polar
    cl = 2 * math.pi * math.radians(alpha)    # linear in alpha: interpolation is exact
    cd = np.interp(alpha, table, 0.01 + 0.05 * np.radians(table) ** 2)
    urel2 = 1.0 + (Om * r) ** 2
    dr = 0.5
    fl, fd = 0.5 * cl * urel2 * 0.1 * dr, 0.5 * cd * urel2 * 0.1 * dr
    fx = fl * math.cos(phi) + fd * math.sin(phi)
    ft = fl * math.sin(phi) - fd * math.sin(phi)  # reference's sin/sin (src/alm/almDevCall.cu:239)
    assert abs(th[0] - fx) < 1e-12 and abs(tq[0] - abs(ft) * r) < 1e-12
    assert np.allclose(f[0], [-fx, ft * math.sin(theta), -ft * math.cos(theta)], atol=1e-12)
    assert th[0] > 0 and ff[:dom.ncell].sum() < 0
    # tangential force opposes... the blade rotates +theta: the fluid is pushed along +theta
    tang = np.array([0.0, -math.sin(theta), math.cos(theta)])
    assert np.dot(f[0], tang) * 1.0 < 0 or ft < 0


def test_alm_pitch_rotation_keeps_radius():
    """A platform pitch of angle psi rotates the rotor about y: the point
    keeps its distance from the base and its y offset."""
    h = 1.0 / 16
    dom = domain((64, 64, 64))
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    psi = math.radians(6.0)
    setup = _one_point_rotor(h, tsr=4.0, twist=5.0, pitch_type=2, motion=(psi, 0.0, 0.5 * math.pi))
    A = O.Alm(dom, setup, *cen)
    u = np.concatenate([np.ones(dom.ncell), np.zeros(dom.ncell), np.zeros(dom.ncell)])
    ff = np.zeros(3 * dom.ncell)
    A.step(u, ff, 0.2)
    xyz, _, _, _ = A.get()
    d = xyz[0] - np.array([2.0, 2.0, 2.0])
    assert abs(np.linalg.norm(d) - 0.25) < 1e-14
    theta = 8.0 * 0.2
    assert abs(d[1] - 0.25 * math.cos(theta)) < 1e-14
    # rotation about y by -psi of (0, r cos, r sin): x = sin(psi)*r sin(theta)
    assert abs(d[0] - math.sin(psi) * 0.25 * math.sin(theta)) < 1e-14


def test_alm_yaw_turns_rotor_about_vertical_axis():
    """A yawed rotor (no reference counterpart) in a uniform stream U along
    x: its point lies on the yawed rotor plane, (x, y) turned by the yaw
    angle about the base, z unchanged; in the rotor's frame the stream is
    (U cos(yaw), -U sin(yaw), 0), so the point's thrust and torque equal
    those of an unyawed rotor in that stream, and its force on the fluid is
    that rotor's force turned by the yaw.  With a static platform pitch
    offset a and yaw 0, the point is the unpitched one turned about y by a."""
    h = 1.0 / 16
    dom = domain((64, 64, 64))
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    t = 0.2
    psi = 25.0
    c, s_ = math.cos(math.radians(psi)), math.sin(math.radians(psi))
    ones = np.ones(dom.ncell)
    zero = np.zeros(dom.ncell)

    def run(setup, U, V=0.0):
        A = O.Alm(dom, setup, *cen)
        ff = np.zeros(3 * dom.ncell)
        A.step(np.concatenate([U * ones, V * ones, zero]), ff, t)
        return A.get()

    xy, fy, tqy, thy = run(_one_point_rotor(h, tsr=4.0, twist=5.0, yaw_deg=psi), 1.0)
    x0, f0, tq0, th0 = run(_one_point_rotor(h, tsr=4.0, twist=5.0), c, -s_)
    d, d0 = xy[0] - 2.0, x0[0] - 2.0
    # rotor frame -> inertial: turned by -yaw about z (x' = c x - s y, y' = s x + c y)
    assert np.allclose(d, [c * d0[0] - s_ * d0[1], s_ * d0[0] + c * d0[1], d0[2]], atol=1e-14, rtol=0)
    assert abs(tqy[0] - tq0[0]) < 1e-12 * abs(tq0[0]) and abs(thy[0] - th0[0]) < 1e-12 * abs(th0[0])
    F0 = f0[0]
    assert np.allclose(fy[0], [c * F0[0] - s_ * F0[1], s_ * F0[0] + c * F0[1], F0[2]], atol=1e-13 * abs(F0).max())
    # a static platform pitch offset a (angle_type 2, no oscillation)
    a = 7.0
    xp, _, _, _ = run(_one_point_rotor(h, tsr=4.0, twist=5.0, pitch_type=2, angle_offset_deg=a), 1.0)
    x1, _, _, _ = run(_one_point_rotor(h, tsr=4.0, twist=5.0), 1.0)
    ca, sa = math.cos(math.radians(a)), math.sin(math.radians(a))
    d, d0 = xp[0] - 2.0, x1[0] - 2.0
    # rot() about y by -a: x' = c x + s z, z' = -s x + c z
    assert np.allclose(d, [ca * d0[0] + sa * d0[2], d0[1], -sa * d0[0] + ca * d0[2]], atol=1e-14, rtol=0)


def test_alm_velocity_sampling_exact_for_linear_field():
    """The actuator point samples u by trilinear interpolation between the
    eight surrounding cell centres (src/alm/almDevCall.cu:28-41,131-253),
    which is exact for a linear field: the point's forces then follow the
    blade-element closed form with u(P) in place of the uniform inflow
    (axial u_x(P), tangential Omega r + u_y(P) sin(theta) - u_z(P) cos(theta)).
    A wrong corner weight, a swapped corner or an index off by one cell
    changes u(P) for a field with gradients along all three axes."""
    h = 1.0 / 16
    dom = domain((64, 64, 64))
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    base = (2.013, 1.987, 2.041)  # off the cell centres and the midpoints along every axis
    setup = cases.rotor(h, base=base, radius=0.5, tsr=4.0, n_blade=1, n_ap_per_blade=1, twist=(20.0, 20.0),
                        chord=(0.1, 0.1))
    A = O.Alm(dom, setup, *cen)
    X, Y, Z = _grid(dom, cen)
    G = np.array([[0.15, -0.2, 0.1], [0.3, 0.05, -0.25], [-0.1, 0.2, 0.12]])
    c0 = np.array([1.0, 0.1, -0.05])
    u = np.concatenate([c0[d] + G[d, 0] * (X - 2) + G[d, 1] * (Y - 2) + G[d, 2] * (Z - 2) for d in range(3)])
    ff = np.zeros(3 * dom.ncell)
    t = 0.3
    A.step(u, ff, t)
    xyz, f, tq, th = A.get()
    P = xyz[0]
    up = c0 + G @ (P - 2.0)
    Om, r = 8.0, 0.25
    theta = Om * t
    ux = up[0]
    ut = Om * r + up[1] * math.sin(theta) - up[2] * math.cos(theta)
    phi = math.atan(ux / ut)
    alpha = math.degrees(phi) - 20.0
    assert -12.0 < alpha < 12.0  # inside the linear part of the synthetic polar
    table = np.arange(-45.0, 46.0, 1.0)
    cl = 2 * math.pi * math.radians(alpha)
    cd = np.interp(alpha, table, 0.01 + 0.05 * np.radians(table) ** 2)
    urel2 = ux * ux + ut * ut
    fl, fd = 0.5 * cl * urel2 * 0.1 * 0.5, 0.5 * cd * urel2 * 0.1 * 0.5
    fx = fl * math.cos(phi) + fd * math.sin(phi)
    ft = fl * math.sin(phi) - fd * math.sin(phi)
    assert abs(th[0] - fx) < 1e-12 and abs(tq[0] - abs(ft) * r) < 1e-12
    assert np.allclose(f[0], [-fx, ft * math.sin(theta), -ft * math.cos(theta)], atol=1e-12, rtol=0)


def test_poisson_matrix_2d_is_textbook_five_point():
    """app/ld2d/main.cpp:14-68: the 2-D cavity's matrix is the 5-point
    Laplacian with Neumann walls (couplings across walls dropped), its z
    columns zero, divided by max|a_c| = 4/h^2.  On a domain of several
    planes each plane is its own 5-point problem (a z-thin domain alone
    cannot tell z couplings from Neumann walls)."""
    n = (7, 5, 3)
    dom = domain(n)
    h = 0.1
    cen = [(np.arange(dom.shape[m]) - 1.5) * h for m in range(3)]
    _, _, g, ja = O.metrics(dom, *cen, *[np.full(dom.shape[m], h) for m in range(3)])
    a, md = O.poisson_matrix(dom, abi.POISSON_2D, g, ja)
    assert abs(md - 4.0 / h ** 2) < 1e-9 * md
    m = interior_mask(dom)
    cols = a.reshape(7, -1)
    assert np.all(cols[5] == 0.0) and np.all(cols[6] == 0.0)
    assert np.max(np.abs(cols[:, m].sum(axis=0))) < 1e-15  # zero row sums
    A = _textbook_laplacian(n, False, zcouple=False)
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal(m.sum())
        x = np.zeros(dom.ncell)
        x[m] = v
        got = O.mv(dom, a, x)[m]
        assert np.max(np.abs(got - A @ v)) < 1e-13


def test_copy_z5_closed_form():
    """app/ld2d/bcdevcall.cu:240-256: planes k = gc-2 .. gc+2 of every column
    equal plane k = gc over the whole padded x-y extent; other planes keep
    their values."""
    for n in ((5, 4, 1), (4, 3, 3)):
        dom = domain(n)
        nx, ny, nz = dom.shape
        v = np.random.default_rng(4).standard_normal(2 * dom.ncell)
        out = v.copy()
        O.copy_z5(dom, out, 2)
        V, W = v.reshape(2, nz, ny, nx), out.reshape(2, nz, ny, nx)
        for k in range(nz):
            if k <= 4:
                assert np.array_equal(W[:, k], V[:, 2])
            else:
                assert np.array_equal(W[:, k], V[:, k])


def test_pressure_fill_closed_form():
    """pbc (app/windtunnel/bcdevcall.cu:153-211): a Neumann face copies the
    first interior layer into the first ghost layer, the Dirichlet outlet
    sets that ghost layer to 0, over the face's interior extent only;
    second ghost layers, edges and corners keep their values."""
    dom = domain((6, 5, 4))
    nx, ny, nz = dom.shape
    p = np.random.default_rng(5).standard_normal(dom.ncell)
    faces = [abi.PBC_NEUMANN, abi.PBC_ZERO, abi.PBC_NEUMANN, abi.PBC_NEUMANN, abi.PBC_NONE, abi.PBC_NEUMANN]
    out = p.copy()
    O.bc_pressure(dom, faces, out)
    P, Q = p.reshape(nz, ny, nx), out.reshape(nz, ny, nx)
    E = P.copy()
    s = slice(2, -2)
    E[s, s, 1] = P[s, s, 2]
    E[s, s, nx - 2] = 0.0
    E[s, 1, s] = P[s, 2, s]
    E[s, ny - 2, s] = P[s, ny - 3, s]
    E[nz - 2, s, s] = P[nz - 3, s, s]
    assert np.array_equal(Q, E)


def test_min_max_reductions_match_numpy():
    """MatColMax/Min/AbsMax/AbsMin of one column and VecMax/Min of the
    per-cell Euclidean norm over all columns (src/MV.h:67-136,
    src/dev/MVDevCall.cu:216-598), interior cells only.  Values are chosen
    so that each extremum sits in the interior and the ghosts hold larger
    and smaller values that a wrong region would pick up."""
    dom = domain((5, 6, 7))
    rng = np.random.default_rng(9)
    a = rng.uniform(0.5, 2.0, 3 * dom.ncell) * rng.choice([-1.0, 1.0], 3 * dom.ncell)
    m = interior_mask(dom)
    ghost = np.nonzero(~m)[0]
    for col in range(3):
        a[col * dom.ncell + ghost] = 50.0 * rng.choice([-1.0, 1.0], ghost.size)
        a[col * dom.ncell + ghost[::7]] = 1e-3
    cols = a.reshape(3, -1)[:, m]
    for col in range(3):
        c = cols[col]
        assert O.reduce(dom, abi.RED_COL_MAX, a, 3, col) == c.max()
        assert O.reduce(dom, abi.RED_COL_MIN, a, 3, col) == c.min()
        assert O.reduce(dom, abi.RED_COL_ABSMAX, a, 3, col) == np.abs(c).max()
        assert O.reduce(dom, abi.RED_COL_ABSMIN, a, 3, col) == np.abs(c).min()
    norm = np.sqrt(cols[0] * cols[0] + cols[1] * cols[1] + cols[2] * cols[2])
    assert O.reduce(dom, abi.RED_VEC_MAX, a, 3, 0) == norm.max()
    assert O.reduce(dom, abi.RED_VEC_MIN, a, 3, 0) == norm.min()


def test_window_power_smooth_in_yaw():
    """The window-mean power of the two-rotor farm case is a smooth function
    of the upstream rotor's yaw when each Poisson solve runs a fixed number
    of BiCGStab iterations below the point where its residual stops falling:
    first differences at a 1e-6 rad spacing agree to 1e-3 of each other
    (no jump from the yawed kinematics, the sampling cell, the polar segment
    or the force box)."""
    from falm_amd import cases
    s = cases.farm_segment(n=(64, 24, 24), cells_per_d=8, spacing_d=3.0, hub_x_d=2.0, cfl=0.3,
                           precision=abi.CASE_F64, tol=1e-30, maxit=10, pc_maxit=1, pc_relax=1.0,
                           period_revs=2.0, rotor_kw=dict(n_ap_per_blade=32))
    c = O.Case(s)
    c.step(30)
    J = []
    for e in np.degrees(1e-6) * np.arange(-2, 3):
        c.set_attitude(0, e, 0.0)
        J.append(c.window(6))
    d = np.diff(np.array(J))
    assert np.all(d != 0.0) and np.all(np.sign(d) == np.sign(d[0]))
    assert np.max(np.abs(d - d.mean())) <= 1e-3 * abs(d.mean()), d
