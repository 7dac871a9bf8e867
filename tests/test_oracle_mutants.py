"""Planted mistakes in the CPU oracle: each row below edits one expression
of oracle/falm_oracle.cpp the way a restatement could plausibly go wrong (a
coefficient, a sign, an index, a dropped term, a swapped operand), builds
that copy into a temporary library and runs the pinning checks named in the
row against it.  At least one of them must fail: that is what shows the
pinning tests would catch such a mistake in the oracle the GPU is compared
against.  Runs on the CPU only; the temporary builds never leave the test's
temporary directory."""
import concurrent.futures as cf
import gc
import os
import subprocess

import pytest

from falm_amd import _abi as abi
from oracle import pyoracle as O
from tests import test_oracle_pinning as P

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "falm_oracle.cpp")

# (name, original text, planted text, checks that must catch it)
MUTANTS = [
    ("upwind3 central coefficient",
     "UDL = JUE * (-ue2 + 27 * ue1 - 27 * ucc + uw1) / 24.;",
     "UDL = JUE * (-ue2 + 28 * ue1 - 27 * ucc + uw1) / 24.;",
     [lambda: P.test_upwind3_family_exact_for_quadratics(abi.ADV_UPWIND3)]),
    ("upwind3 fourth difference sign",
     "UD4 = Uabs * (ue2 - 4 * ue1 + 6 * ucc - 4 * uw1 + uw2) * ALPHA;",
     "UD4 = Uabs * (ue2 - 4 * ue1 + 6 * ucc - 4 * uw1 - uw2) * ALPHA;",
     [P.test_upwind3_dissipation_term_is_fourth_difference]),
    ("QUICK downstream weight",
     "adv += JUE * (-ue2 + 9 * ue1 + 9 * ucc - uw1);",
     "adv += JUE * (-ue2 + 8 * ue1 + 9 * ucc - uw1);",
     [lambda: P.test_upwind3_family_exact_for_quadratics(abi.ADV_QUICK)]),
    ("QUICK upwind term sign",
     "adv -= fabs(JVS) * (un1 - 3 * ucc + 3 * us1 - us2);",
     "adv += fabs(JVS) * (un1 - 3 * ucc + 3 * us1 - us2);",
     [P.test_quick_closed_form_for_cubics]),
    ("upwind1 takes the downwind value",
     "return 0.5 * (UF * (ur + ul) - fabs(UF) * (ur - ul));",
     "return 0.5 * (UF * (ur + ul) + fabs(UF) * (ur - ul));",
     [P.test_upwind1_closed_form_for_quadratics]),
    ("diffusion without the 1/2",
     "vis /= (2 * jac);",
     "vis /= jac;",
     [P.test_diffusion_exact_for_cubics]),
    ("diffusion bottom face sign",
     "vis -= (gzcc + gzb1) * (ReI + 0.5 * (nutcc + nutb1)) * (ucc - ub1);",
     "vis += (gzcc + gzb1) * (ReI + 0.5 * (nutcc + nutb1)) * (ucc - ub1);",
     [P.test_diffusion_exact_for_cubics]),
    ("explicit Euler step sign",
     "ua(cc, d) = un(cc, d) + dt * (-adv + vis + ff(cc, d));",
     "ua(cc, d) = un(cc, d) + dt * (adv + vis + ff(cc, d));",
     [lambda: P.test_upwind3_explicit_euler_amplification(0.3)]),
    ("face flux weight",
     "uu(c, 0) = 0.5 * (uc(c, 0) + uc(D.id(i + 1, j, k), 0));",
     "uu(c, 0) = 0.5 * (uc(c, 0) + uc(D.id(i - 1, j, k), 0));",
     [P.test_u_to_uu_linear_exact_and_divergence]),
    ("Smagorinsky length not squared",
     "nut_[c] = SQ(lc) * Du;",
     "nut_[c] = lc * Du;",
     [P.test_smagorinsky_linear_field]),
    ("Smagorinsky shear sign",
     "d4 = SQ(duy + dvx)",
     "d4 = SQ(duy - dvx)",
     [P.test_smagorinsky_linear_field]),
    ("face projection sign",
     "uu(c, 0) = a0 - dt * dpx;",
     "uu(c, 0) = a0 + dt * dpx;",
     [P.test_projection_makes_divergence_free]),
    ("operator columns w/e swapped",
     "double ac = a(c, 0), aw = a(c, 1), ae = a(c, 2)",
     "double ac = a(c, 0), aw = a(c, 2), ae = a(c, 1)",
     [P.test_heat1d_exact_solution_has_zero_residual]),
    ("red-black colour ignores the offset",
     "if ((i + j + k + osum) % 2 != color) continue;",
     "if ((i + j + k) % 2 != color) continue;",
     [P.test_red_black_sweep_is_colour_gauss_seidel]),
    ("BiCGStab s update sign",
     "return r[cc] - alpha * q[cc];",
     "return r[cc] + alpha * q[cc];",
     [P.test_bicgstab_and_sor_match_direct_solve]),
    ("dot product drops the last plane",
     "for (int k = gc; k < D.nz - gc; k++) s += part[k];",
     "for (int k = gc; k < D.nz - gc - 1; k++) s += part[k];",
     [P.test_dot_norm_max_known_answer]),
    ("2-D matrix keeps z couplings",
     "if (kind != FALM_POISSON_2D) {",
     "if (true) {",
     [P.test_poisson_matrix_2d_is_textbook_five_point]),
    ("outlet coupling dropped",
     "const int xe_lim = (kind == FALM_POISSON_OUTLET_XPLUS) ? D.gnx - gc : D.gnx - gc - 1;",
     "const int xe_lim = D.gnx - gc - 1;",
     [lambda: P.test_poisson_matrix_is_textbook_laplacian(abi.POISSON_OUTLET_XPLUS)]),
    ("convective outlet stencil",
     "(3 * uprev(co1, m) - 4 * uprev(ci1, m) + uprev(ci2, m));",
     "(3 * uprev(co1, m) - 4 * uprev(ci1, m) - uprev(ci2, m));",
     [P.test_bc_windtunnel_closed_forms]),
    ("slip second ghost layer",
     "u(co2, m) = 2 * ub[m] - u(ci1, m);",
     "u(co2, m) = 2 * ub[m] - u(ci2, m);",
     [P.test_bc_windtunnel_closed_forms]),
    ("reflection second ghost layer",
     "u(co2, m) = 2 * bc.value[m] - u(ci2, m);",
     "u(co2, m) = 2 * bc.value[m] - u(ci1, m);",
     [P.test_bc_lid_reflection_averages_to_wall_velocity]),
    ("Dirichlet pressure as an odd reflection",
     "p[co] = (faces[f] == FALM_PBC_ZERO) ? 0.0 : p[ci];",
     "p[co] = (faces[f] == FALM_PBC_ZERO) ? -p[ci] : p[ci];",
     [P.test_pressure_fill_closed_form]),
    ("copy_z5 source plane",
     "double val = v[D.id(i, j, gc) + (size_t)n * D.ncell];",
     "double val = v[D.id(i, j, gc + 1) + (size_t)n * D.ncell];",
     [P.test_copy_z5_closed_form]),
    ("absolute minimum without fabs",
     "if (op == FALM_RED_COL_ABSMAX || op == FALM_RED_COL_ABSMIN) v = fabs(v);",
     "if (op == FALM_RED_COL_ABSMAX) v = fabs(v);",
     [P.test_min_max_reductions_match_numpy]),
    ("trilinear weights swapped",
     "double c01 = c0 * (1. - xd) + c1 * xd;",
     "double c01 = c0 * xd + c1 * (1. - xd);",
     [P.test_alm_velocity_sampling_exact_for_linear_field]),
    ("trilinear z fraction",
     "double zd = fabs(P[2] - z0) / fabs(z1 - z0);",
     "double zd = fabs(P[2] - z1) / fabs(z1 - z0);",
     [P.test_alm_velocity_sampling_exact_for_linear_field]),
    ("tangential force drag term",
     "                                                           : fl * sin(phi) - fd * sin(phi);",
     "                                                           : fl * sin(phi) - fd * cos(phi);",
     [P.test_alm_blade_element_closed_form]),
    ("Gaussian normalisation",
     "ColW ff{ff_, D.ncell};\n    const double eta = 1. / ((eps * sqrt(PI)) * (eps * sqrt(PI)) * (eps * sqrt(PI)));",
     "ColW ff{ff_, D.ncell};\n    const double eta = 1. / ((eps * sqrt(PI)) * (eps * sqrt(PI)) * (eps * PI));",
     [P.test_gaussian_projection_mass_and_centroid]),
    ("polar interpolation fraction",
     "double p = (att - attack[ak]) / (attack[ak + 1] - attack[ak]);",
     "double p = (att - attack[ak]) / (attack[ak + 1] - attack[0]);",
     [P.test_polar_lookup_is_piecewise_linear_with_clamps]),
    ("Philox round word",
     "uint32_t n0 = hi1 ^ ctr[1] ^ k0;",
     "uint32_t n0 = hi1 ^ ctr[3] ^ k0;",
     [P.test_philox_known_answers]),
    ("platform pitch rotation sign",
     "out[0] = C * X - S * Z; out[1] = Y; out[2] = S * X + C * Z;",
     "out[0] = C * X + S * Z; out[1] = Y; out[2] = S * X + C * Z;",
     [P.test_alm_pitch_rotation_keeps_radius]),
]


def _build(src_text, out):
    cpp = out[:-3] + ".cpp"
    with open(cpp, "w") as f:
        f.write(src_text)
    cmd = ["g++", "-O1", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-I", os.path.dirname(ORACLE_SRC),
           "-o", out, cpp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return out


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    with open(ORACLE_SRC) as f:
        src = f.read()
    jobs = {}
    for q, (name, orig, planted, _) in enumerate(MUTANTS):
        assert src.count(orig) == 1, f"{name}: the original text must occur exactly once"
        jobs[name] = (src.replace(orig, planted), str(d / f"m{q:02d}.so"))
    jobs["unchanged"] = (src, str(d / "base.so"))
    with cf.ThreadPoolExecutor(min(8, os.cpu_count() or 2)) as ex:
        futs = {name: ex.submit(_build, *a) for name, a in jobs.items()}
        return {name: f.result() for name, f in futs.items()}


def _fails(path, checks):
    failed = False
    with O.using(path):
        for check in checks:
            try:
                check()
            except AssertionError:
                failed = True
            gc.collect()  # release oracle handles made by this build while it is loaded
            if failed:
                break
    return failed


def test_unchanged_copy_passes(mutant_libs):
    """The copy built the same way with no edit passes every check used
    below (so a failure of a mutant is the planted mistake's doing)."""
    checks = [c for *_, cs in MUTANTS for c in cs]
    assert not _fails(mutant_libs["unchanged"], checks)


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(mutant_libs, name):
    checks = next(m[3] for m in MUTANTS if m[0] == name)
    assert _fails(mutant_libs[name], checks), f"planted mistake not caught: {name}"
