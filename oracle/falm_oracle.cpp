/*
 * falm_oracle.cpp — CPU FP64 restatement of the reference (ki2098/falm)
 * fractional-step LES step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it and never falls
 * back to it.
 *
 * Every function cites the reference file:line it restates and keeps the
 * reference's arithmetic order where that is cheap.  Deliberate semantic
 * choices (each documented where it is made):
 *   - pseudo-velocity is out of place (the reference apps alias ua == u,
 *     a data race in kernel_Cartesian_CalcPseudoU);
 *   - reductions sum in a fixed order (per k-plane partials, then planes in
 *     order), not the reference's block-partial + host order;
 *   - BiCGStab with no preconditioner uses the identity (the reference's
 *     Precondition() leaves pp = 0, which divides by zero);
 *   - torque/thrust are summed sequentially (the reference uses
 *     thrust::reduce, whose order is not specified).
 * Built by oracle/build.py into oracle/liboracle.so (g++ -O3 -fopenmp).
 */
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <functional>
#include "../include/falm_amd.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Dom {
    int nx, ny, nz, gc;
    int ox, oy, oz;
    int gnx, gny, gnz;
    size_t ncell;
    explicit Dom(const falm_domain *d)
        : nx(d->shape[0]), ny(d->shape[1]), nz(d->shape[2]), gc(d->gc),
          ox(d->offset[0]), oy(d->offset[1]), oz(d->offset[2]),
          gnx(d->gshape[0]), gny(d->gshape[1]), gnz(d->gshape[2]),
          ncell((size_t)d->shape[0] * d->shape[1] * d->shape[2]) {}
    /* IDX (src/util.h:29-31) */
    inline size_t id(int i, int j, int k) const { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k); }
    /* gmap.size: global interior cell count (src/FalmEq.cpp:311) */
    double n_interior_global() const { return (double)(gnx - 2 * gc) * (gny - 2 * gc) * (gnz - 2 * gc); }
};

inline double SQ(double x) { return x * x; }

/* ---------------------------------------------------------------- schemes */
/* src/FalmCFDScheme.h:8-10 */
inline double UpwindFlux1st(double ul, double ur, double UF) {
    return 0.5 * (UF * (ur + ul) - fabs(UF) * (ur - ul));
}

/* Upwind1st, src/FalmCFDScheme.h:12-31 */
inline double Upwind1st(double ucc, double ue1, double uw1, double un1, double us1, double ut1, double ub1,
                        double JUE, double JUW, double JVN, double JVS, double JWT, double JWB, double jac) {
    double adv = 0.0;
    adv += UpwindFlux1st(ucc, ue1, JUE);
    adv -= UpwindFlux1st(uw1, ucc, JUW);
    adv += UpwindFlux1st(ucc, un1, JVN);
    adv -= UpwindFlux1st(us1, ucc, JVS);
    adv += UpwindFlux1st(ucc, ut1, JWT);
    adv -= UpwindFlux1st(ub1, ucc, JWB);
    adv /= jac;
    return adv;
}

/* QUICK, src/FalmCFDScheme.h:33-66 */
inline double Quick(double ucc, double ue1, double ue2, double uw1, double uw2,
                    double un1, double un2, double us1, double us2,
                    double ut1, double ut2, double ub1, double ub2,
                    double JUE, double JUW, double JVN, double JVS, double JWT, double JWB, double jac) {
    double adv = 0.;
    adv += JUE * (-ue2 + 9 * ue1 + 9 * ucc - uw1);
    adv += fabs(JUE) * (ue2 - 3 * ue1 + 3 * ucc - uw1);
    adv -= JUW * (-ue1 + 9 * ucc + 9 * uw1 - uw2);
    adv -= fabs(JUW) * (ue1 - 3 * ucc + 3 * uw1 - uw2);
    adv += JVN * (-un2 + 9 * un1 + 9 * ucc - us1);
    adv += fabs(JVN) * (un2 - 3 * un1 + 3 * ucc - us1);
    adv -= JVS * (-un1 + 9 * ucc + 9 * us1 - us2);
    adv -= fabs(JVS) * (un1 - 3 * ucc + 3 * us1 - us2);
    adv += JWT * (-ut2 + 9 * ut1 + 9 * ucc - ub1);
    adv += fabs(JWT) * (ut2 - 3 * ut1 + 3 * ucc - ub1);
    adv -= JWB * (-ut1 + 9 * ucc + 9 * ub1 - ub2);
    adv -= fabs(JWB) * (ut1 - 3 * ucc + 3 * ub1 - ub2);
    adv /= (16 * jac);
    return adv;
}

/* Upwind3rd / UTOPIA / KK share one form with a different 4th-difference
 * weight ALPHA: 1/24 (src/FalmCFDScheme.h:134-165), 1/12 (:101-132),
 * 1/4 (:68-99). */
inline double Upwind3Family(double ALPHA, double ucc, double ue1, double ue2, double uw1, double uw2,
                            double un1, double un2, double us1, double us2,
                            double ut1, double ut2, double ub1, double ub2,
                            double Uabs, double Vabs, double Wabs,
                            double JUE, double JUW, double JVN, double JVS, double JWT, double JWB,
                            double jac) {
    double adv = 0.0, UDL, UDR, UD4;
    UDL = JUE * (-ue2 + 27 * ue1 - 27 * ucc + uw1) / 24.;
    UDR = JUW * (-ue1 + 27 * ucc - 27 * uw1 + uw2) / 24.;
    UD4 = Uabs * (ue2 - 4 * ue1 + 6 * ucc - 4 * uw1 + uw2) * ALPHA;
    adv += .5 * (UDL + UDR) / jac + UD4;
    UDL = JVN * (-un2 + 27 * un1 - 27 * ucc + us1) / 24.;
    UDR = JVS * (-un1 + 27 * ucc - 27 * us1 + us2) / 24.;
    UD4 = Vabs * (un2 - 4 * un1 + 6 * ucc - 4 * us1 + us2) * ALPHA;
    adv += .5 * (UDL + UDR) / jac + UD4;
    UDL = JWT * (-ut2 + 27 * ut1 - 27 * ucc + ub1) / 24.;
    UDR = JWB * (-ut1 + 27 * ucc - 27 * ub1 + ub2) / 24.;
    UD4 = Wabs * (ut2 - 4 * ut1 + 6 * ucc - 4 * ub1 + ub2) * ALPHA;
    adv += .5 * (UDL + UDR) / jac + UD4;
    return adv;
}

/* Diffusion, src/FalmCFDScheme.h:167-191 */
inline double Diffusion(double ReI, double ucc, double ue1, double uw1, double un1, double us1,
                        double ut1, double ub1, double nutcc, double nute1, double nutw1,
                        double nutn1, double nuts1, double nutt1, double nutb1,
                        double gxcc, double gxe1, double gxw1, double gycc, double gyn1,
                        double gys1, double gzcc, double gzt1, double gzb1, double jac) {
    double vis = 0.0;
    vis += (gxcc + gxe1) * (ReI + 0.5 * (nutcc + nute1)) * (ue1 - ucc);
    vis -= (gxcc + gxw1) * (ReI + 0.5 * (nutcc + nutw1)) * (ucc - uw1);
    vis += (gycc + gyn1) * (ReI + 0.5 * (nutcc + nutn1)) * (un1 - ucc);
    vis -= (gycc + gys1) * (ReI + 0.5 * (nutcc + nuts1)) * (ucc - us1);
    vis += (gzcc + gzt1) * (ReI + 0.5 * (nutcc + nutt1)) * (ut1 - ucc);
    vis -= (gzcc + gzb1) * (ReI + 0.5 * (nutcc + nutb1)) * (ucc - ub1);
    vis /= (2 * jac);
    return vis;
}

/* per-k-plane partial sums reduced in plane order: deterministic for any
 * thread count */
template <class F>
double plane_sum(const Dom &D, F f) {
    const int gc = D.gc;
    std::vector<double> part(D.nz, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++) {
        double s = 0.0;
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) s += f(D.id(i, j, k));
        part[k] = s;
    }
    double s = 0.0;
    for (int k = gc; k < D.nz - gc; k++) s += part[k];
    return s;
}

/* Shorthand for column access of the reference's Matrix layout
 * (src/matrix.h:20): element (cell, col) at p[cell + col*ncell]. */
struct Col {
    const double *p;
    size_t n;
    inline double operator()(size_t c, int col = 0) const { return p[c + (size_t)col * n]; }
};
struct ColW {
    double *p;
    size_t n;
    inline double &operator()(size_t c, int col = 0) const { return p[c + (size_t)col * n]; }
};

} // namespace

extern "C" {

/* ============================================================ FalmCFD */

/* kernel_Cartesian_CalcPseudoU, src/dev/FalmCFDDevCall.cu:10-301 (interior
 * region, src/FalmCFD.cpp:33-35).  Out of place: ua must not alias u. */
int orc_pseudo_u(const falm_domain *dom, int scheme, double ReI, double dt,
                 const double *un_, const double *u_, const double *uu_, double *ua_,
                 const double *nut_, const double *kx_, const double *g_, const double *ja_,
                 const double *ff_) {
    Dom D(dom);
    if (ua_ == u_ || ua_ == un_) return FALM_ERR_ALIAS;
    Col un{un_, D.ncell}, u{u_, D.ncell}, uu{uu_, D.ncell}, nut{nut_, D.ncell}, kx{kx_, D.ncell},
        g{g_, D.ncell}, ja{ja_, D.ncell}, ff{ff_, D.ncell};
    ColW ua{ua_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t cc = D.id(i, j, k);
                size_t e1 = D.id(i + 1, j, k), w1 = D.id(i - 1, j, k);
                size_t n1 = D.id(i, j + 1, k), s1 = D.id(i, j - 1, k);
                size_t t1 = D.id(i, j, k + 1), b1 = D.id(i, j, k - 1);
                size_t e2 = D.id(i + 2, j, k), w2 = D.id(i - 2, j, k);
                size_t n2 = D.id(i, j + 2, k), s2 = D.id(i, j - 2, k);
                size_t t2 = D.id(i, j, k + 2), b2 = D.id(i, j, k - 2);
                double uc = u(cc, 0), vc = u(cc, 1), wc = u(cc, 2);
                double Uabs = fabs(uc * kx(cc, 0));
                double Vabs = fabs(vc * kx(cc, 1));
                double Wabs = fabs(wc * kx(cc, 2));
                double UE = uu(cc, 0), UW = uu(w1, 0);
                double VN = uu(cc, 1), VS = uu(s1, 1);
                double WT = uu(cc, 2), WB = uu(b1, 2);
                double nutcc = nut(cc), nute1 = nut(e1), nutw1 = nut(w1), nutn1 = nut(n1),
                       nuts1 = nut(s1), nutt1 = nut(t1), nutb1 = nut(b1);
                double gxcc = g(cc, 0), gxe1 = g(e1, 0), gxw1 = g(w1, 0);
                double gycc = g(cc, 1), gyn1 = g(n1, 1), gys1 = g(s1, 1);
                double gzcc = g(cc, 2), gzt1 = g(t1, 2), gzb1 = g(b1, 2);
                double jac = ja(cc);
                for (int d = 0; d < 3; d++) {
                    double ucc = (d == 0) ? uc : (d == 1) ? vc : wc;
                    double ue1 = u(e1, d), uw1 = u(w1, d), un1 = u(n1, d), us1 = u(s1, d), ut1 = u(t1, d), ub1 = u(b1, d);
                    double ue2 = u(e2, d), uw2 = u(w2, d), un2 = u(n2, d), us2 = u(s2, d), ut2 = u(t2, d), ub2 = u(b2, d);
                    double adv;
                    if (scheme == FALM_ADV_UPWIND3) {
                        adv = Upwind3Family(1. / 24., ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2,
                                            Uabs, Vabs, Wabs, UE, UW, VN, VS, WT, WB, jac);
                    } else if (scheme == FALM_ADV_QUICK) {
                        adv = Quick(ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2,
                                    UE, UW, VN, VS, WT, WB, jac);
                    } else if (scheme == FALM_ADV_UTOPIA) {
                        adv = Upwind3Family(1. / 12., ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2,
                                            Uabs, Vabs, Wabs, UE, UW, VN, VS, WT, WB, jac);
                    } else if (scheme == FALM_ADV_KK) {
                        adv = Upwind3Family(1. / 4., ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2,
                                            Uabs, Vabs, Wabs, UE, UW, VN, VS, WT, WB, jac);
                    } else {
                        adv = Upwind1st(ucc, ue1, uw1, un1, us1, ut1, ub1, UE, UW, VN, VS, WT, WB, jac);
                    }
                    double vis = Diffusion(ReI, ucc, ue1, uw1, un1, us1, ut1, ub1, nutcc, nute1, nutw1,
                                           nutn1, nuts1, nutt1, nutb1, gxcc, gxe1, gxw1, gycc, gyn1, gys1,
                                           gzcc, gzt1, gzb1, jac);
                    ua(cc, d) = un(cc, d) + dt * (-adv + vis + ff(cc, d));
                }
            }
    return FALM_OK;
}

/* FalmCFD::UtoUU, src/FalmCFD.cpp:69-129: UtoCU on Region(shape, gc-1)
 * (src/dev/FalmCFDDevCall.cu:303-325) then InterpolateCU on the face region
 * shape-2gc+1 at offset gc-1 (:327-349). */
int orc_u_to_uu(const falm_domain *dom, const double *u_, double *uu_, const double *kx_,
                const double *ja_) {
    Dom D(dom);
    Col u{u_, D.ncell}, kx{kx_, D.ncell}, ja{ja_, D.ncell};
    ColW uu{uu_, D.ncell};
    std::vector<double> ucv(3 * D.ncell, 0.0);
    ColW uc{ucv.data(), D.ncell};
    const int lo = D.gc - 1;
#pragma omp parallel for schedule(static)
    for (int k = lo; k < D.nz - lo; k++)
        for (int j = lo; j < D.ny - lo; j++)
            for (int i = lo; i < D.nx - lo; i++) {
                size_t c = D.id(i, j, k);
                double jacob = ja(c);
                uc(c, 0) = jacob * kx(c, 0) * u(c, 0);
                uc(c, 1) = jacob * kx(c, 1) * u(c, 1);
                uc(c, 2) = jacob * kx(c, 2) * u(c, 2);
            }
#pragma omp parallel for schedule(static)
    for (int k = lo; k < D.nz - D.gc; k++)
        for (int j = lo; j < D.ny - D.gc; j++)
            for (int i = lo; i < D.nx - D.gc; i++) {
                size_t c = D.id(i, j, k);
                uu(c, 0) = 0.5 * (uc(c, 0) + uc(D.id(i + 1, j, k), 0));
                uu(c, 1) = 0.5 * (uc(c, 1) + uc(D.id(i, j + 1, k), 1));
                uu(c, 2) = 0.5 * (uc(c, 2) + uc(D.id(i, j, k + 1), 2));
            }
    return FALM_OK;
}

/* kernel_Cartesian_Divergence, src/dev/FalmCFDDevCall.cu:557-581 */
int orc_divergence(const falm_domain *dom, const double *uu_, double *div_, const double *ja_) {
    Dom D(dom);
    Col uu{uu_, D.ncell}, ja{ja_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                double UE = uu(c, 0), UW = uu(D.id(i - 1, j, k), 0);
                double VN = uu(c, 1), VS = uu(D.id(i, j - 1, k), 1);
                double WT = uu(c, 2), WB = uu(D.id(i, j, k - 1), 2);
                div_[c] = (UE - UW + VN - VS + WT - WB) / ja(c);
            }
    return FALM_OK;
}

/* kernel_ScaleMatrix, src/dev/MVDevCall.cu:506-522 (all entries) */
int orc_scale(double *a, size_t n, double scale) {
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)n; t++) a[t] *= scale;
    return FALM_OK;
}

/* FalmCFD::MACCalcPoissonRHS, src/FalmCFD.h:96-109 */
int orc_poisson_rhs(const falm_domain *dom, const double *uu, double *rhs, const double *ja,
                    double dt, double maxdiag) {
    Dom D(dom);
    orc_divergence(dom, uu, rhs, ja);
    return orc_scale(rhs, D.ncell, 1.0 / (dt * maxdiag));
}

/* FalmCFD::ProjectP, src/FalmCFD.cpp:131-185: ProjectPGrid on the interior
 * (src/dev/FalmCFDDevCall.cu:351-376) then ProjectPFace on the face region
 * (:378-407).  u == ua and uu == uua are allowed (point-wise). */
int orc_project_p(const falm_domain *dom, double dt, double *u_, const double *ua_, double *uu_,
                  const double *uua_, const double *p_, const double *kx_, const double *g_) {
    Dom D(dom);
    Col ua{ua_, D.ncell}, uua{uua_, D.ncell}, kx{kx_, D.ncell}, g{g_, D.ncell};
    ColW u{u_, D.ncell}, uu{uu_, D.ncell};
    const double *p = p_;
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                double dpx = 0.5 * kx(c, 0) * (p[D.id(i + 1, j, k)] - p[D.id(i - 1, j, k)]);
                double dpy = 0.5 * kx(c, 1) * (p[D.id(i, j + 1, k)] - p[D.id(i, j - 1, k)]);
                double dpz = 0.5 * kx(c, 2) * (p[D.id(i, j, k + 1)] - p[D.id(i, j, k - 1)]);
                double a0 = ua(c, 0), a1 = ua(c, 1), a2 = ua(c, 2);
                u(c, 0) = a0 - dt * dpx;
                u(c, 1) = a1 - dt * dpy;
                u(c, 2) = a2 - dt * dpz;
            }
    const int lo = gc - 1;
#pragma omp parallel for schedule(static)
    for (int k = lo; k < D.nz - gc; k++)
        for (int j = lo; j < D.ny - gc; j++)
            for (int i = lo; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k), e1 = D.id(i + 1, j, k), n1 = D.id(i, j + 1, k), t1 = D.id(i, j, k + 1);
                double pcc = p[c];
                double dpx = 0.5 * (g(c, 0) + g(e1, 0)) * (p[e1] - pcc);
                double dpy = 0.5 * (g(c, 1) + g(n1, 1)) * (p[n1] - pcc);
                double dpz = 0.5 * (g(c, 2) + g(t1, 2)) * (p[t1] - pcc);
                double a0 = uua(c, 0), a1 = uua(c, 1), a2 = uua(c, 2);
                uu(c, 0) = a0 - dt * dpx;
                uu(c, 1) = a1 - dt * dpy;
                uu(c, 2) = a2 - dt * dpz;
            }
    return FALM_OK;
}

/* kernel_Cartesian_Smagorinsky, src/dev/FalmCFDDevCall.cu:409-476; dispatch
 * :774-797 (Empty does nothing; CSM is out of scope). */
int orc_sgs(const falm_domain *dom, int model, double Cs, const double *u_, double *nut_,
            const double *kx_, const double *ja_) {
    Dom D(dom);
    if (model == FALM_SGS_EMPTY) return FALM_OK;
    if (model != FALM_SGS_SMAGORINSKY) return FALM_ERR_UNSUPPORTED;
    Col u{u_, D.ncell}, kx{kx_, D.ncell}, ja{ja_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                size_t e1 = D.id(i + 1, j, k), w1 = D.id(i - 1, j, k), n1 = D.id(i, j + 1, k),
                       s1 = D.id(i, j - 1, k), t1 = D.id(i, j, k + 1), b1 = D.id(i, j, k - 1);
                double kxx = kx(c, 0), kxy = kx(c, 1), kxz = kx(c, 2);
                double dux = 0.5 * kxx * (u(e1, 0) - u(w1, 0));
                double duy = 0.5 * kxy * (u(n1, 0) - u(s1, 0));
                double duz = 0.5 * kxz * (u(t1, 0) - u(b1, 0));
                double dvx = 0.5 * kxx * (u(e1, 1) - u(w1, 1));
                double dvy = 0.5 * kxy * (u(n1, 1) - u(s1, 1));
                double dvz = 0.5 * kxz * (u(t1, 1) - u(b1, 1));
                double dwx = 0.5 * kxx * (u(e1, 2) - u(w1, 2));
                double dwy = 0.5 * kxy * (u(n1, 2) - u(s1, 2));
                double dwz = 0.5 * kxz * (u(t1, 2) - u(b1, 2));
                double d1 = 2 * SQ(dux), d2 = 2 * SQ(dvy), d3 = 2 * SQ(dwz);
                double d4 = SQ(duy + dvx), d5 = SQ(dvz + dwy), d6 = SQ(duz + dwx);
                double Du = sqrt(d1 + d2 + d3 + d4 + d5 + d6);
                double De = cbrt(ja(c));
                double lc = Cs * De;
                nut_[c] = SQ(lc) * Du;
            }
    return FALM_OK;
}

/* ============================================================ FalmMV */

/* kernel_MV, src/dev/MVDevCall.cu:10-41 (coefficients [c,w,e,s,n,b,t]) */
static inline double stencil7(const Dom &D, const Col &a, const double *x, int i, int j, int k, size_t c) {
    double ac = a(c, 0), aw = a(c, 1), ae = a(c, 2), as = a(c, 3), an = a(c, 4), ab = a(c, 5), at = a(c, 6);
    double xc = x[c], xe = x[D.id(i + 1, j, k)], xw = x[D.id(i - 1, j, k)], xn = x[D.id(i, j + 1, k)],
           xs = x[D.id(i, j - 1, k)], xt = x[D.id(i, j, k + 1)], xb = x[D.id(i, j, k - 1)];
    return ac * xc + ae * xe + aw * xw + an * xn + as * xs + at * xt + ab * xb;
}

int orc_mv(const falm_domain *dom, const double *a_, const double *x, double *ax) {
    Dom D(dom);
    Col a{a_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                ax[c] = stencil7(D, a, x, i, j, k, c);
            }
    return FALM_OK;
}

/* kernel_Struct3d7p_Res, src/dev/FalmEqDevCall.cu:54-85 */
int orc_residual(const falm_domain *dom, const double *a_, const double *x, const double *b, double *r) {
    Dom D(dom);
    Col a{a_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                r[c] = b[c] - stencil7(D, a, x, i, j, k, c);
            }
    return FALM_OK;
}

/* FalmMV::DotProduct, src/MV.h:47-55, kernel src/dev/MVDevCall.cu:57-116 */
double orc_dot(const falm_domain *dom, const double *a, const double *b) {
    Dom D(dom);
    return plane_sum(D, [&](size_t c) { return a[c] * b[c]; });
}

/* FalmMV::EuclideanNormSq, src/MV.h:57-65, kernel src/dev/MVDevCall.cu:118-179
 * (sum over all ncol columns) */
double orc_norm2(const falm_domain *dom, const double *a, int ncol) {
    Dom D(dom);
    return plane_sum(D, [&](size_t c) {
        double t = 0;
        for (int d = 0; d < ncol; d++) t += a[c + d * D.ncell] * a[c + d * D.ncell];
        return t;
    });
}

/* MatColMax/Min/AbsMax/AbsMin, VecMax/Min (src/MV.h:67-136,
 * src/dev/MVDevCall.cu:216-598): the extremum over the interior cells.  The
 * reference seeds the block cache of threads outside the map with 0
 * (src/dev/MVDevCall.cu:251-283), so when the interior does not fill its last
 * thread blocks its MatColMax is max(true max, 0), MatColMin min(true min,
 * 0) and MatColAbsMin 0; its result equals this one when the interior is a
 * multiple of the caller's block shape (a launch hint this API does not
 * take), and always for MatColAbsMax (MaxDiag), the one the step uses. */
int orc_reduce(const falm_domain *dom, int op, const double *a, int ncol, int col, double *out) {
    Dom D(dom);
    const int gc = D.gc;
    bool first = true;
    double r = 0.0;
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                double v;
                if (op == FALM_RED_VEC_MAX || op == FALM_RED_VEC_MIN) {
                    double t = 0;
                    for (int d = 0; d < ncol; d++) t += a[c + d * D.ncell] * a[c + d * D.ncell];
                    v = sqrt(t);
                } else {
                    v = a[c + (size_t)col * D.ncell];
                    if (op == FALM_RED_COL_ABSMAX || op == FALM_RED_COL_ABSMIN) v = fabs(v);
                }
                bool take_max = (op == FALM_RED_COL_MAX || op == FALM_RED_COL_ABSMAX || op == FALM_RED_VEC_MAX);
                if (first || (take_max ? v > r : v < r)) r = v;
                first = false;
            }
    *out = r;
    return FALM_OK;
}

/* kernel_MatrixAdd / kernel_Vecaxby, src/dev/MVDevCall.cu:660-695 */
int orc_matrix_add(double *a, const double *b, size_t n) {
    for (size_t t = 0; t < n; t++) a[t] += b[t];
    return FALM_OK;
}
int orc_vecaxby(double a, const double *x, double b, const double *y, double *res, size_t n) {
    for (size_t t = 0; t < n; t++) res[t] = a * x[t] + b * y[t];
    return FALM_OK;
}

/* ============================================================ FalmEq */

/* kernel_Struct3d7p_SOR, src/dev/FalmEqDevCall.cu:175-221.  Cells of the
 * other colour are rewritten with their own value in the reference, so a
 * half sweep is an in-place update of one colour only. */
int orc_sor_sweep(const falm_domain *dom, const double *a_, double *x, const double *b, double omega, int color) {
    Dom D(dom);
    Col a{a_, D.ncell};
    const int gc = D.gc;
    const int osum = D.ox + D.oy + D.oz;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                if ((i + j + k + osum) % 2 != color) continue;
                size_t c = D.id(i, j, k);
                double xc = x[c];
                double cc = (b[c] - stencil7(D, a, x, i, j, k, c)) / a(c, 0);
                x[c] = xc + omega * cc;
            }
    return FALM_OK;
}

/* kernel_Struct3d7p_Jacobi, src/dev/FalmEqDevCall.cu:101-132 */
int orc_jacobi_sweep(const falm_domain *dom, const double *a_, double *x, const double *xp, const double *b) {
    Dom D(dom);
    Col a{a_, D.ncell};
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                double xc = xp[c];
                x[c] = xc + (b[c] - stencil7(D, a, xp, i, j, k, c)) / a(c, 0);
            }
    return FALM_OK;
}

struct OrcSolver {
    falm_domain dom;
    falm_solver_config cfg;
    std::vector<double> rr, p, q, s, pp, ss, t, xp;
};

/* element-wise update over interior cells: kernel_PBiCGStab_1..4,
 * src/dev/FalmEqDevCall.cu:252-342 */
static void each_interior(const Dom &D, double *dst, const std::function<double(size_t)> &f) {
    const int gc = D.gc;
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                dst[c] = f(c);
            }
}

/* FalmEq::JacobiPC (src/FalmEq.cpp:129-171) and FalmEq::SORPC (:242-303);
 * pc_type Empty is taken as the identity (see file header). */
static void orc_precondition(OrcSolver *S, const double *a, double *x, const double *b) {
    Dom D(&S->dom);
    const falm_solver_config &c = S->cfg;
    if (c.pc_type == FALM_LS_JACOBI) {
        for (int it = 0; it < c.pc_maxit; it++) {
            std::copy(x, x + D.ncell, S->xp.begin());
            orc_jacobi_sweep(&S->dom, a, x, S->xp.data(), b);
        }
    } else if (c.pc_type == FALM_LS_SOR) {
        for (int it = 0; it < c.pc_maxit; it++) {
            orc_sor_sweep(&S->dom, a, x, b, c.pc_relax, FALM_COLOR_BLACK);
            orc_sor_sweep(&S->dom, a, x, b, c.pc_relax, FALM_COLOR_RED);
        }
    } else {
        each_interior(D, x, [&](size_t cc) { return b[cc]; });
    }
}

int orc_solver_create(const falm_domain *dom, const falm_solver_config *cfg, void **out) {
    OrcSolver *S = new OrcSolver();
    S->dom = *dom;
    S->cfg = *cfg;
    size_t n = (size_t)dom->shape[0] * dom->shape[1] * dom->shape[2];
    S->rr.assign(n, 0.0); S->p.assign(n, 0.0); S->q.assign(n, 0.0); S->s.assign(n, 0.0);
    S->pp.assign(n, 0.0); S->ss.assign(n, 0.0); S->t.assign(n, 0.0); S->xp.assign(n, 0.0);
    *out = S;
    return FALM_OK;
}

int orc_solver_destroy(void *s) {
    delete (OrcSolver *)s;
    return FALM_OK;
}

/* FalmEq::Solve (src/FalmEq.h:56-64): Jacobi (src/FalmEq.cpp:81-127),
 * SOR (:173-240), PBiCGStab (:305-410). */
int orc_solver_solve(void *handle, const double *a, double *x, const double *b, double *r,
                     falm_solve_info *info) {
    OrcSolver *S = (OrcSolver *)handle;
    Dom D(&S->dom);
    const falm_solver_config &c = S->cfg;
    const double gN = D.n_interior_global();
    int it = 0;
    double err = 0.0;
    int status = FALM_OK;
    if (c.type == FALM_LS_JACOBI) {
        do {
            std::copy(x, x + D.ncell, S->xp.begin());
            orc_jacobi_sweep(&S->dom, a, x, S->xp.data(), b);
            orc_residual(&S->dom, a, x, b, r);
            err = sqrt(orc_norm2(&S->dom, r, 1)) / gN;
            it++;
        } while (it < c.maxit && err > c.tol);
    } else if (c.type == FALM_LS_SOR) {
        do {
            orc_sor_sweep(&S->dom, a, x, b, c.relax, FALM_COLOR_BLACK);
            orc_sor_sweep(&S->dom, a, x, b, c.relax, FALM_COLOR_RED);
            orc_residual(&S->dom, a, x, b, r);
            err = sqrt(orc_norm2(&S->dom, r, 1)) / gN;
            it++;
        } while (it < c.maxit && err > c.tol);
    } else if (c.type == FALM_LS_PBICGSTAB) {
        double *rr = S->rr.data(), *p = S->p.data(), *q = S->q.data(), *s = S->s.data();
        double *pp = S->pp.data(), *ss = S->ss.data(), *t = S->t.data();
        orc_residual(&S->dom, a, x, b, r);
        err = sqrt(orc_norm2(&S->dom, r, 1)) / gN;
        std::copy(r, r + D.ncell, rr);
        double rho, rrho = 1.0, alpha = 0.0, beta, omega = 1.0;
        do {
            rho = orc_dot(&S->dom, r, rr);
            if (fabs(rho) < FLT_MIN) {
                err = rho;
                status = FALM_ERR_BREAKDOWN;
                break;
            }
            if (it == 0) {
                std::copy(r, r + D.ncell, p);
            } else {
                beta = (rho * alpha) / (rrho * omega);
                each_interior(D, p, [&](size_t cc) { return r[cc] + beta * (p[cc] - omega * q[cc]); });
            }
            std::fill(pp, pp + D.ncell, 0.0);
            orc_precondition(S, a, pp, p);
            orc_mv(&S->dom, a, pp, q);
            alpha = rho / orc_dot(&S->dom, rr, q);
            each_interior(D, s, [&](size_t cc) { return r[cc] - alpha * q[cc]; });
            std::fill(ss, ss + D.ncell, 0.0);
            orc_precondition(S, a, ss, s);
            orc_mv(&S->dom, a, ss, t);
            omega = orc_dot(&S->dom, t, s) / orc_dot(&S->dom, t, t);
            each_interior(D, x, [&](size_t cc) { return x[cc] + (alpha * pp[cc] + omega * ss[cc]); });
            each_interior(D, r, [&](size_t cc) { return s[cc] - omega * t[cc]; });
            rrho = rho;
            err = sqrt(orc_norm2(&S->dom, r, 1)) / gN;
            it++;
        } while (it < c.maxit && err > c.tol);
    } else {
        return FALM_ERR_UNSUPPORTED;
    }
    if (status == FALM_OK && err > c.tol) status = FALM_ERR_NOT_CONVERGED;
    if (info) {
        info->it = it;
        info->err = err;
        info->outer_it = 1;
        info->status = status;
    }
    return FALM_OK;
}

/* ======================================================= boundary fill */

static inline void face_axes(int f, int &d, int &a, int &b) {
    d = f / 2;
    a = (d == 0) ? 1 : 0;
    b = (d == 2) ? 1 : 2;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and a Box-Muller pair:
 * the inlet perturbation generator shared with the GPU path (no reference
 * counterpart). */
static void philox4x32_10(uint32_t ctr[4], const uint32_t key_in[2]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ ctr[3] ^ k1;
        uint32_t n3 = lo0;
        ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
        k0 += W0; k1 += W1;
    }
}

/* exported for the known-answer test of the generator */
void orc_philox4x32_10(uint32_t ctr[4], const uint32_t key[2]) { philox4x32_10(ctr, key); }

/* three N(0,1) deviates for inlet cell (a, b) of the given step */
void orc_inlet_normals(uint64_t seed, int64_t step, int a, int b, double out[3]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)a, (uint32_t)b};
    philox4x32_10(ctr, key);
    /* 2^-32 spaced uniforms in (0,1] and [0,1) */
    double u1 = ((double)ctr[0] + 1.0) * 2.3283064365386963e-10;
    double u2 = (double)ctr[1] * 2.3283064365386963e-10;
    double u3 = ((double)ctr[2] + 1.0) * 2.3283064365386963e-10;
    double u4 = (double)ctr[3] * 2.3283064365386963e-10;
    const double TWO_PI = 6.283185307179586;
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(TWO_PI * u2);
    out[1] = r1 * sin(TWO_PI * u2);
    out[2] = r2 * cos(TWO_PI * u4);
}

/* ubc for windtunnel/nedo (app/windtunnel/bcdevcall.cu:8-115), lid3d
 * (app/lid3d/bcdev.cu:11-105) and ld2d (app/ld2d/bcdevcall.cu); each face
 * covers the interior range of the two in-face indices. */
int orc_bc_velocity(const falm_domain *dom, const falm_face_bc faces[6], double *u_, const double *uprev_,
                    const double *xyz_, double dt, const falm_inlet_noise *noise) {
    Dom D(dom);
    const int gc = D.gc;
    ColW u{u_, D.ncell};
    Col uprev{uprev_, D.ncell}, xyz{xyz_, D.ncell};
    const int shp[3] = {D.nx, D.ny, D.nz};
    for (int f = 0; f < 6; f++) {
        const falm_face_bc &bc = faces[f];
        if (bc.type == FALM_BC_NONE) continue;
        int d, ax, bx;
        face_axes(f, d, ax, bx);
        const int plus = f & 1;
        if (bc.type == FALM_BC_CONVECTIVE && f != 1) return FALM_ERR_UNSUPPORTED;
        for (int kb = gc; kb < shp[bx] - gc; kb++)
            for (int ka = gc; ka < shp[ax] - gc; ka++) {
                int o1[3], o2[3], i1[3], i2[3];
                for (int m = 0; m < 3; m++) { o1[m] = o2[m] = i1[m] = i2[m] = 0; }
                o1[ax] = o2[ax] = i1[ax] = i2[ax] = ka;
                o1[bx] = o2[bx] = i1[bx] = i2[bx] = kb;
                if (plus) {
                    o1[d] = shp[d] - gc; o2[d] = shp[d] - gc + 1; i1[d] = shp[d] - gc - 1; i2[d] = shp[d] - gc - 2;
                } else {
                    o1[d] = gc - 1; o2[d] = gc - 2; i1[d] = gc; i2[d] = gc + 1;
                }
                size_t co1 = D.id(o1[0], o1[1], o1[2]), co2 = D.id(o2[0], o2[1], o2[2]);
                size_t ci1 = D.id(i1[0], i1[1], i1[2]), ci2 = D.id(i2[0], i2[1], i2[2]);
                if (bc.type == FALM_BC_INLET) {
                    double v[3] = {bc.value[0], bc.value[1], bc.value[2]};
                    if (noise && noise->intensity != 0.0) {
                        double z[3];
                        orc_inlet_normals(noise->seed, noise->step, ka, kb, z);
                        for (int m = 0; m < 3; m++) v[m] = v[m] + noise->intensity * z[m];
                    }
                    for (int m = 0; m < 3; m++) { u(co1, m) = v[m]; u(co2, m) = v[m]; }
                } else if (bc.type == FALM_BC_CONVECTIVE) {
                    /* kernel_ubc_xplus, app/windtunnel/bcdevcall.cu:25-43 */
                    double dxi = 1.0 / (xyz(co1, 0) - xyz(ci1, 0));
                    for (int m = 0; m < 3; m++) {
                        double du = 0.5 * dxi * (3 * uprev(co1, m) - 4 * uprev(ci1, m) + uprev(ci2, m));
                        double ub = uprev(co1, m) - dt * du;
                        u(co1, m) = ub;
                        u(co2, m) = 2 * ub - u(ci1, m);
                    }
                } else if (bc.type == FALM_BC_SLIP) {
                    double ub[3] = {u(ci1, 0), u(ci1, 1), u(ci1, 2)};
                    ub[d] = 0.0;
                    for (int m = 0; m < 3; m++) {
                        u(co1, m) = ub[m];
                        u(co2, m) = 2 * ub[m] - u(ci1, m);
                    }
                } else if (bc.type == FALM_BC_WALL0) {
                    for (int m = 0; m < 3; m++) {
                        u(co1, m) = 0.0;
                        u(co2, m) = 2 * 0.0 - u(ci1, m);
                    }
                } else if (bc.type == FALM_BC_REFLECT) {
                    for (int m = 0; m < 3; m++) {
                        u(co1, m) = 2 * bc.value[m] - u(ci1, m);
                        u(co2, m) = 2 * bc.value[m] - u(ci2, m);
                    }
                } else {
                    return FALM_ERR_INVALID;
                }
            }
    }
    return FALM_OK;
}

/* pbc (app/windtunnel/bcdevcall.cu:153-211, app/lid3d/bcdev.cu:143-213) */
int orc_bc_pressure(const falm_domain *dom, const int faces[6], double *p) {
    Dom D(dom);
    const int gc = D.gc;
    const int shp[3] = {D.nx, D.ny, D.nz};
    for (int f = 0; f < 6; f++) {
        if (faces[f] == FALM_PBC_NONE) continue;
        int d, ax, bx;
        face_axes(f, d, ax, bx);
        for (int kb = gc; kb < shp[bx] - gc; kb++)
            for (int ka = gc; ka < shp[ax] - gc; ka++) {
                int o1[3], i1[3];
                o1[ax] = i1[ax] = ka;
                o1[bx] = i1[bx] = kb;
                if (f & 1) { o1[d] = shp[d] - gc; i1[d] = shp[d] - gc - 1; }
                else       { o1[d] = gc - 1;      i1[d] = gc; }
                size_t co = D.id(o1[0], o1[1], o1[2]), ci = D.id(i1[0], i1[1], i1[2]);
                p[co] = (faces[f] == FALM_PBC_ZERO) ? 0.0 : p[ci];
            }
    }
    return FALM_OK;
}

/* uubc (app/lid3d/bcdev.cu:251-321) */
int orc_bc_face_flux(const falm_domain *dom, const int faces[6], double *uu) {
    Dom D(dom);
    const int gc = D.gc;
    const int shp[3] = {D.nx, D.ny, D.nz};
    for (int f = 0; f < 6; f++) {
        if (!faces[f]) continue;
        int d, ax, bx;
        face_axes(f, d, ax, bx);
        for (int kb = gc; kb < shp[bx] - gc; kb++)
            for (int ka = gc; ka < shp[ax] - gc; ka++) {
                int o[3];
                o[ax] = ka;
                o[bx] = kb;
                o[d] = (f & 1) ? shp[d] - gc - 1 : gc - 1;
                uu[D.id(o[0], o[1], o[2]) + (size_t)d * D.ncell] = 0.0;
            }
    }
    return FALM_OK;
}

/* copy_z5, app/ld2d/bcdevcall.cu:240-256 */
int orc_copy_z5(const falm_domain *dom, double *v, int ncol) {
    Dom D(dom);
    const int gc = D.gc;
    for (int j = 0; j < D.ny; j++)
        for (int i = 0; i < D.nx; i++)
            for (int n = 0; n < ncol; n++) {
                double val = v[D.id(i, j, gc) + (size_t)n * D.ncell];
                for (int dk = -2; dk <= 2; dk++)
                    if (dk) v[D.id(i, j, gc + dk) + (size_t)n * D.ncell] = val;
            }
    return FALM_OK;
}

/* ================================================== metrics / matrix */

/* FalmCore::computation_init metric loop, src/falm.h:383-404 */
int orc_metrics(const falm_domain *dom, const double *x, const double *y, const double *z,
                const double *hx, const double *hy, const double *hz,
                double *xyz, double *kx, double *g, double *ja) {
    Dom D(dom);
    for (int k = 0; k < D.nz; k++)
        for (int j = 0; j < D.ny; j++)
            for (int i = 0; i < D.nx; i++) {
                size_t c = D.id(i, j, k);
                xyz[c] = x[i]; xyz[c + D.ncell] = y[j]; xyz[c + 2 * D.ncell] = z[k];
                double pitch[3] = {hx[i], hy[j], hz[k]};
                double volume = pitch[0] * pitch[1] * pitch[2];
                double dkdx[3] = {1.0 / pitch[0], 1.0 / pitch[1], 1.0 / pitch[2]};
                ja[c] = volume;
                for (int d = 0; d < 3; d++) {
                    g[c + d * D.ncell] = volume * dkdx[d] * dkdx[d];
                    kx[c + d * D.ncell] = dkdx[d];
                }
            }
    return FALM_OK;
}

/* makePoissonCoefficientMatrix (app/windtunnel/main.cpp:20-92; lid3d
 * app/lid3d/main.cpp:22-92; ld2d app/ld2d/main.cpp:14-68), then MaxDiag
 * (src/MV.h:67-76) and ScaleMatrix (src/dev/MVDevCall.cu:506-522). */
int orc_poisson_matrix(const falm_domain *dom, int kind, const double *g_, const double *ja_, double *a_,
                       double *maxdiag_out) {
    Dom D(dom);
    Col g{g_, D.ncell}, ja{ja_, D.ncell};
    ColW a{a_, D.ncell};
    const int gc = D.gc;
    const int xe_lim = (kind == FALM_POISSON_OUTLET_XPLUS) ? D.gnx - gc : D.gnx - gc - 1;
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                int gi = D.ox + i, gj = D.oy + j, gk = D.oz + k;
                double ac = 0, ae = 0, aw = 0, an = 0, as = 0, at = 0, ab = 0;
                size_t cc = D.id(i, j, k);
                double gxcc = g(cc, 0), gxe1 = g(D.id(i + 1, j, k), 0), gxw1 = g(D.id(i - 1, j, k), 0);
                double gycc = g(cc, 1), gyn1 = g(D.id(i, j + 1, k), 1), gys1 = g(D.id(i, j - 1, k), 1);
                double gzcc = g(cc, 2), gzt1 = g(D.id(i, j, k + 1), 2), gzb1 = g(D.id(i, j, k - 1), 2);
                double jacob = ja(cc), coef;
                coef = 0.5 * (gxcc + gxe1) / jacob;
                if (gi < xe_lim) { ac -= coef; ae = coef; }
                coef = 0.5 * (gxcc + gxw1) / jacob;
                if (gi > gc) { ac -= coef; aw = coef; }
                coef = 0.5 * (gycc + gyn1) / jacob;
                if (gj < D.gny - gc - 1) { ac -= coef; an = coef; }
                coef = 0.5 * (gycc + gys1) / jacob;
                if (gj > gc) { ac -= coef; as = coef; }
                if (kind != FALM_POISSON_2D) {
                    coef = 0.5 * (gzcc + gzt1) / jacob;
                    if (gk < D.gnz - gc - 1) { ac -= coef; at = coef; }
                    coef = 0.5 * (gzcc + gzb1) / jacob;
                    if (gk > gc) { ac -= coef; ab = coef; }
                }
                a(cc, 0) = ac; a(cc, 1) = aw; a(cc, 2) = ae; a(cc, 3) = as;
                a(cc, 4) = an; a(cc, 5) = ab; a(cc, 6) = at;
            }
    double maxdiag;
    orc_reduce(dom, FALM_RED_COL_ABSMAX, a_, 7, 0, &maxdiag);
    orc_scale(a_, 7 * D.ncell, 1.0 / maxdiag);
    if (maxdiag_out) *maxdiag_out = maxdiag;
    return FALM_OK;
}

/* ================================================================ ALM */

static const double PI = 3.14159265358979323846; /* Falm::Pi = M_PI, src/falmath.h:10 */
static inline double rad2deg(double r) { return r * 180. / PI; }
static inline double deg2rad(double d) { return d * PI / 180.; }

/* one_angle_frame_rotation, src/falmath.h:55-81 */
static void rot(const double x[3], const double ang[3], int type, double out[3]) {
    const double X = x[0], Y = x[1], Z = x[2];
    if (type == 1) {
        double C = cos(ang[0]), S = sin(ang[0]);
        out[0] = X; out[1] = C * Y + S * Z; out[2] = -S * Y + C * Z;
    } else if (type == 2) {
        double C = cos(ang[1]), S = sin(ang[1]);
        out[0] = C * X - S * Z; out[1] = Y; out[2] = S * X + C * Z;
    } else if (type == 3) {
        double C = cos(ang[2]), S = sin(ang[2]);
        out[0] = C * X + S * Y; out[1] = -S * X + C * Y; out[2] = Z;
    } else {
        out[0] = X; out[1] = Y; out[2] = Z;
    }
}

/* one_angle_frame_rotation_dt, src/falmath.h:83-131 */
static void rot_dt(const double x[3], const double v[3], const double ang[3], const double om[3], int type,
                   double out[3]) {
    const double X = x[0], Y = x[1], Z = x[2], U = v[0], V = v[1], W = v[2];
    if (type == 1) {
        double C = cos(ang[0]), S = sin(ang[0]), O = om[0];
        double a[3] = {0., O * (-S * Y + C * Z), O * (-C * Y - S * Z)};
        double b[3] = {U, C * V + S * W, -S * V + C * W};
        for (int m = 0; m < 3; m++) out[m] = a[m] + b[m];
    } else if (type == 2) {
        double C = cos(ang[1]), S = sin(ang[1]), O = om[1];
        double a[3] = {O * (-S * X - C * Z), 0., O * (C * X - S * Z)};
        double b[3] = {C * U - S * W, V, S * U + C * W};
        for (int m = 0; m < 3; m++) out[m] = a[m] + b[m];
    } else if (type == 3) {
        double C = cos(ang[2]), S = sin(ang[2]), O = om[2];
        double a[3] = {O * (-S * X + C * Y), O * (-C * X - S * Y), 0.};
        double b[3] = {C * U + S * V, -S * U + C * V, W};
        for (int m = 0; m < 3; m++) out[m] = a[m] + b[m];
    } else {
        out[0] = U; out[1] = V; out[2] = W;
    }
}

/* A yawed rotor on a moving platform (no reference counterpart; yaw 0 is
 * the reference): rotor frame -> inertial frame turns x by -yaw about z and
 * then by -ang about the platform axis; the platform frame's relative
 * velocity is turned by +yaw about z into the rotor frame. */
static void yawed_to_inertial(const double x[3], double yaw_deg, const double nang[3], int type, double out[3]) {
    if (yaw_deg != 0.0) {
        const double ny[3] = {0.0, 0.0, -deg2rad(yaw_deg)};
        double y[3];
        rot(x, ny, 3, y);
        rot(y, nang, type, out);
    } else {
        rot(x, nang, type, out);
    }
}
static void platform_to_rotor(double v[3], double yaw_deg) {
    if (yaw_deg != 0.0) {
        const double py[3] = {0.0, 0.0, deg2rad(yaw_deg)};
        const double w[3] = {v[0], v[1], v[2]};
        rot(w, py, 3, v);
    }
}

/* find_index, src/alm/almDevCall.cu:12-26 */
static int find_index(const double *seq, double v, int size) {
    if (v < seq[0]) return -1;
    if (v >= seq[size - 1]) return size - 1;
    for (int i = 0; i < size - 1; i++)
        if (seq[i] <= v && v < seq[i + 1]) return i;
    return size - 1;
}

/* trilinear_interpolation, src/alm/almDevCall.cu:28-41 */
static double trilinear(double xd, double yd, double zd, double c0, double c1, double c2, double c3,
                        double c4, double c5, double c6, double c7) {
    double c01 = c0 * (1. - xd) + c1 * xd;
    double c23 = c2 * (1. - xd) + c3 * xd;
    double c45 = c4 * (1. - xd) + c5 * xd;
    double c67 = c6 * (1. - xd) + c7 * xd;
    double c0123 = c01 * (1. - yd) + c23 * yd;
    double c4567 = c45 * (1. - yd) + c67 * yd;
    return c0123 * (1. - zd) + c4567 * zd;
}

/* APFrame::get_airfoil_params, src/alm/apHandler.h:121-141: linear
 * interpolation over the attack angle (degrees), clamped at both ends;
 * cl/cd laid out [attack][ap] (APFrame::id, :52-54). */
void orc_polar_lookup(const double *attack, int nat, const double *cl, const double *cd, int nap, int ap,
                      double att, double *out_cl, double *out_cd) {
    double c = 0, d = 0;
    if (att < attack[0]) {
        c = cl[ap]; d = cd[ap];
    } else if (att >= attack[nat - 1]) {
        c = cl[ap + (size_t)(nat - 1) * nap]; d = cd[ap + (size_t)(nat - 1) * nap];
    } else {
        for (int ak = 0; ak < nat - 1; ak++) {
            if (attack[ak] <= att && attack[ak + 1] > att) {
                double p = (att - attack[ak]) / (attack[ak + 1] - attack[ak]);
                c = (1. - p) * cl[ap + (size_t)ak * nap] + p * cl[ap + (size_t)(ak + 1) * nap];
                d = (1. - p) * cd[ap + (size_t)ak * nap] + p * cd[ap + (size_t)(ak + 1) * nap];
                break;
            }
        }
    }
    *out_cl = c;
    *out_cd = d;
}

/* kernel_DistributeAPForce, src/alm/almDevCall.cu:255-294: all-pairs
 * Gaussian projection of point forces onto the interior cells */
void orc_alm_project(const falm_domain *dom, const double *x, const double *y, const double *z, int nap,
                     const double *ap_xyz, const double *ap_force, double eps, double *ff_) {
    Dom D(dom);
    const int gc = D.gc;
    ColW ff{ff_, D.ncell};
    const double eta = 1. / ((eps * sqrt(PI)) * (eps * sqrt(PI)) * (eps * sqrt(PI)));
#pragma omp parallel for schedule(static)
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                double cx = x[i], cy = y[j], cz = z[k];
                double fx = 0., fy = 0., fz = 0.;
                for (int ap = 0; ap < nap; ap++) {
                    const double *P = &ap_xyz[3 * ap];
                    double rr2 = SQ(cx - P[0]) + SQ(cy - P[1]) + SQ(cz - P[2]);
                    double e = exp(-rr2 / SQ(eps));
                    fx += ap_force[3 * ap] * eta * e;
                    fy += ap_force[3 * ap + 1] * eta * e;
                    fz += ap_force[3 * ap + 2] * eta * e;
                }
                size_t c = D.id(i, j, k);
                ff(c, 0) = fx; ff(c, 1) = fy; ff(c, 2) = fz;
            }
}

struct OrcAlm {
    falm_domain dom;
    int nt, nb, napb, nat, nap;
    double radius, hub_radius, eps;
    int flags;
    std::vector<falm_turbine> tb;
    std::vector<double> r, chord, twist, attack, cl, cd, twist_off;
    std::vector<double> x, y, z;
    /* state */
    std::vector<double> angle, angvel, base, base_vel; /* [nt*3] */
    std::vector<double> set_vel; /* [nt*3] base velocity of orc_alm_set_base */
    std::vector<double> ap_xyz, ap_force, ap_torque, ap_thrust;
    std::vector<int> ap_ijk, ap_rank;
    std::vector<double> torque, thrust;
};

int orc_alm_create(const falm_domain *dom, const falm_alm_config *c, const double *x, const double *y,
                   const double *z, void **out) {
    OrcAlm *A = new OrcAlm();
    A->dom = *dom;
    A->nt = c->n_turbine; A->nb = c->n_blade; A->napb = c->n_ap_per_blade; A->nat = c->n_attack;
    A->nap = A->nt * A->nb * A->napb;
    A->radius = c->radius; A->hub_radius = c->hub_radius; A->eps = c->eps; A->flags = c->flags;
    A->tb.assign(c->turbines, c->turbines + A->nt);
    A->r.assign(c->ap_r, c->ap_r + A->nap);
    A->chord.assign(c->ap_chord, c->ap_chord + A->nap);
    A->twist.assign(c->ap_twist, c->ap_twist + A->nap);
    A->attack.assign(c->attack, c->attack + A->nat);
    A->cl.assign(c->cl, c->cl + (size_t)A->nat * A->nap);
    A->cd.assign(c->cd, c->cd + (size_t)A->nat * A->nap);
    A->twist_off.assign(A->napb, 0.0);
    if (c->twist_offset) A->twist_off.assign(c->twist_offset, c->twist_offset + A->napb);
    A->x.assign(x, x + dom->shape[0]);
    A->y.assign(y, y + dom->shape[1]);
    A->z.assign(z, z + dom->shape[2]);
    A->angle.assign(3 * A->nt, 0.0); A->angvel.assign(3 * A->nt, 0.0);
    A->base.assign(3 * A->nt, 0.0); A->base_vel.assign(3 * A->nt, 0.0); A->set_vel.assign(3 * A->nt, 0.0);
    A->ap_xyz.assign(3 * A->nap, 0.0); A->ap_force.assign(3 * A->nap, 0.0);
    A->ap_torque.assign(A->nap, 0.0); A->ap_thrust.assign(A->nap, 0.0);
    A->ap_ijk.assign(3 * A->nap, 0); A->ap_rank.assign(A->nap, -1);
    A->torque.assign(A->nt, 0.0); A->thrust.assign(A->nt, 0.0);
    *out = A;
    return FALM_OK;
}

int orc_alm_destroy(void *a) {
    delete (OrcAlm *)a;
    return FALM_OK;
}

int orc_alm_set_twist_offset(void *h, const double *off) {
    OrcAlm *A = (OrcAlm *)h;
    for (int i = 0; i < A->napb; i++) A->twist_off[i] = off ? off[i] : 0.0;
    return FALM_OK;
}

/* AlmHandler::Alm, src/alm/alm.h:66-77 */
int orc_alm_step(void *h, const double *u_, double *ff_, double t) {
    OrcAlm *A = (OrcAlm *)h;
    Dom D(&A->dom);
    const int gc = D.gc;
    Col uvw{u_, D.ncell};
    ColW ff{ff_, D.ncell};
    /* kernel_UpdateTurbineAngles, src/alm/almDevCall.cu:43-61; the base
     * surge term has no reference counterpart (base_velocity is 0 there,
     * src/alm/turbineHandler.h:144-150). */
    for (int tid = 0; tid < A->nt; tid++) {
        const falm_turbine &T = A->tb[tid];
        double ang[3] = {0, 0, 0}, om[3] = {0, 0, 0};
        if (T.angle_type) {
            /* the static part of the platform angle has no reference
             * counterpart (0 by default) */
            const double sa = T.motion[0] * sin(T.motion[1] * t + T.motion[2]);
            ang[T.angle_type - 1] = T.angle_offset_deg != 0.0 ? deg2rad(T.angle_offset_deg) + sa : sa;
            om[T.angle_type - 1] = T.motion[0] * T.motion[1] * cos(T.motion[1] * t + T.motion[2]);
        }
        for (int m = 0; m < 3; m++) {
            A->angle[3 * tid + m] = ang[m];
            A->angvel[3 * tid + m] = om[m];
            A->base[3 * tid + m] = T.base[m];
            A->base_vel[3 * tid + m] = A->set_vel[3 * tid + m];
        }
        if (T.surge[0] != 0.0) {
            A->base[3 * tid] = T.base[0] + T.surge[0] * sin(T.surge[1] * t + T.surge[2]);
            A->base_vel[3 * tid] = A->set_vel[3 * tid] + T.surge[0] * T.surge[1] * cos(T.surge[1] * t + T.surge[2]);
        }
    }
    const int napt = A->napb * A->nb;
    /* kernel_UpdateAPX, src/alm/almDevCall.cu:63-129 */
    for (int ap = 0; ap < A->nap; ap++) {
        int tid = ap / napt, bid = (ap % napt) / A->napb;
        const falm_turbine &T = A->tb[tid];
        double theta0 = (2 * PI / A->nb) * bid;
        double theta = T.tip_rate * t + theta0;
        double c1[3] = {T.hub[0], T.hub[1], T.hub[2]};
        c1[1] += A->r[ap] * cos(theta);
        c1[2] += A->r[ap] * sin(theta);
        double nang[3] = {-A->angle[3 * tid], -A->angle[3 * tid + 1], -A->angle[3 * tid + 2]};
        double c0[3];
        yawed_to_inertial(c1, T.yaw_deg, nang, T.angle_type, c0);
        for (int m = 0; m < 3; m++) c0[m] += A->base[3 * tid + m];
        for (int m = 0; m < 3; m++) A->ap_xyz[3 * ap + m] = c0[m];
        int I = find_index(A->x.data(), c0[0], D.nx);
        int J = find_index(A->y.data(), c0[1], D.ny);
        int K = find_index(A->z.data(), c0[2], D.nz);
        A->ap_ijk[3 * ap] = I; A->ap_ijk[3 * ap + 1] = J; A->ap_ijk[3 * ap + 2] = K;
        bool in = I >= gc - 1 && I < D.nx - gc - 1 && J >= gc - 1 && J < D.ny - gc - 1 && K >= gc - 1 && K < D.nz - gc - 1;
        A->ap_rank[ap] = in ? 0 : -1;
    }
    /* kernel_CalcAPForce, src/alm/almDevCall.cu:131-253 */
    const double dr = (A->radius - A->hub_radius) / A->napb;
    for (int ap = 0; ap < A->nap; ap++) {
        int tid = ap / napt, bid = (ap % napt) / A->napb, sec = ap % A->napb;
        const falm_turbine &T = A->tb[tid];
        if (A->ap_rank[ap] != 0) {
            for (int m = 0; m < 3; m++) A->ap_force[3 * ap + m] = 0;
            A->ap_torque[ap] = 0;
            A->ap_thrust[ap] = 0;
            continue;
        }
        int i0 = A->ap_ijk[3 * ap], j0 = A->ap_ijk[3 * ap + 1], k0 = A->ap_ijk[3 * ap + 2];
        int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
        double x0 = A->x[i0], x1 = A->x[i1], y0 = A->y[j0], y1 = A->y[j1], z0 = A->z[k0], z1 = A->z[k1];
        double P[3] = {A->ap_xyz[3 * ap], A->ap_xyz[3 * ap + 1], A->ap_xyz[3 * ap + 2]};
        double xd = fabs(P[0] - x0) / fabs(x1 - x0);
        double yd = fabs(P[1] - y0) / fabs(y1 - y0);
        double zd = fabs(P[2] - z0) / fabs(z1 - z0);
        size_t id0 = D.id(i0, j0, k0), id1 = D.id(i1, j0, k0), id2 = D.id(i0, j1, k0), id3 = D.id(i1, j1, k0);
        size_t id4 = D.id(i0, j0, k1), id5 = D.id(i1, j0, k1), id6 = D.id(i0, j1, k1), id7 = D.id(i1, j1, k1);
        double uap[3];
        for (int m = 0; m < 3; m++)
            uap[m] = trilinear(xd, yd, zd, uvw(id0, m), uvw(id1, m), uvw(id2, m), uvw(id3, m), uvw(id4, m),
                               uvw(id5, m), uvw(id6, m), uvw(id7, m));
        double rel[3], vrel[3], uvw_tt[3];
        for (int m = 0; m < 3; m++) {
            rel[m] = P[m] - A->base[3 * tid + m];
            vrel[m] = uap[m] - A->base_vel[3 * tid + m];
        }
        rot_dt(rel, vrel, &A->angle[3 * tid], &A->angvel[3 * tid], T.angle_type, uvw_tt);
        platform_to_rotor(uvw_tt, T.yaw_deg);
        double theta0 = (2 * PI / A->nb) * bid;
        double theta = T.tip_rate * t + theta0;
        double ux = uvw_tt[0];
        double ut = T.tip_rate * A->r[ap] + uvw_tt[1] * sin(theta) - uvw_tt[2] * cos(theta);
        double urel2 = ux * ux + ut * ut;
        double phi = atan(ux / ut);
        /* APFrame::get_airfoil_params, src/alm/apHandler.h:121-141; the blade
         * pitch and per-section twist offset add to the AP twist (no
         * reference counterpart; both 0 by default). */
        double chord = A->chord[ap];
        double twist = A->twist[ap] + T.pitch_deg[bid < 8 ? bid : 7] + A->twist_off[sec];
        double att = rad2deg(phi) - twist;
        double cl = 0, cd = 0;
        orc_polar_lookup(A->attack.data(), A->nat, A->cl.data(), A->cd.data(), A->nap, ap, att, &cl, &cd);
        double fl = .5 * cl * urel2 * chord * dr;
        double fd = .5 * cd * urel2 * chord * dr;
        double fx = fl * cos(phi) + fd * sin(phi);
        double ft = (A->flags & FALM_ALM_FIX_DRAG_TANGENT) ? fl * sin(phi) - fd * cos(phi)
                                                           : fl * sin(phi) - fd * sin(phi);
        int sg = (T.tip_rate > 0) - (T.tip_rate < 0);
        ft *= sg;
        double ftt[3] = {-fx, ft * sin(theta), -ft * cos(theta)};
        double nang[3] = {-A->angle[3 * tid], -A->angle[3 * tid + 1], -A->angle[3 * tid + 2]};
        double F[3];
        yawed_to_inertial(ftt, T.yaw_deg, nang, T.angle_type, F);
        for (int m = 0; m < 3; m++) A->ap_force[3 * ap + m] = F[m];
        A->ap_torque[ap] = fabs(ft) * A->r[ap];
        A->ap_thrust[ap] = fx;
    }
    orc_alm_project(&A->dom, A->x.data(), A->y.data(), A->z.data(), A->nap, A->ap_xyz.data(),
                    A->ap_force.data(), A->eps, ff_);
    (void)ff;
    /* AlmDevCall::CalcTorqueAndThrust, src/alm/almDevCall.cu:362-370 */
    for (int tid = 0; tid < A->nt; tid++) {
        double tq = 0, th = 0;
        for (int a = 0; a < napt; a++) {
            tq += A->ap_torque[tid * napt + a];
            th += A->ap_thrust[tid * napt + a];
        }
        A->torque[tid] = tq;
        A->thrust[tid] = th;
    }
    return FALM_OK;
}

int orc_alm_get(void *h, double *ap_xyz, double *ap_force, double *torque, double *thrust) {
    OrcAlm *A = (OrcAlm *)h;
    if (ap_xyz) std::copy(A->ap_xyz.begin(), A->ap_xyz.end(), ap_xyz);
    if (ap_force) std::copy(A->ap_force.begin(), A->ap_force.end(), ap_force);
    if (torque) std::copy(A->torque.begin(), A->torque.end(), torque);
    if (thrust) std::copy(A->thrust.begin(), A->thrust.end(), thrust);
    return FALM_OK;
}

/* BladeHandler::buildAP, src/alm/bladeHandler.h:16-104 (r column used as an
 * absolute radius, as the reference does) */
int orc_build_ap(int nst, int nat, const double *rr, const double *chord, const double *twist,
                 const double *cl, const double *cd, int nt, int nb, int napb, double radius, double hub,
                 double *ap_r, double *ap_chord, double *ap_twist, double *ap_cl, double *ap_cd) {
    int nap = nt * nb * napb;
    double dr = (radius - hub) / napb;
    for (int ap = 0; ap < nap; ap++) {
        double apr = (ap % napb + 0.5) * dr + hub;
        ap_r[ap] = apr;
        if (apr < rr[0]) {
            ap_chord[ap] = chord[0]; ap_twist[ap] = twist[0];
            for (int a = 0; a < nat; a++) { ap_cl[ap + (size_t)a * nap] = cl[a]; ap_cd[ap + (size_t)a * nap] = cd[a]; }
        } else if (apr >= rr[nst - 1]) {
            ap_chord[ap] = chord[nst - 1]; ap_twist[ap] = twist[nst - 1];
            for (int a = 0; a < nat; a++) {
                ap_cl[ap + (size_t)a * nap] = cl[(size_t)(nst - 1) * nat + a];
                ap_cd[ap + (size_t)a * nap] = cd[(size_t)(nst - 1) * nat + a];
            }
        } else {
            int i = 0;
            double p = 0;
            for (i = 0; i < nst - 1; i++)
                if (rr[i] <= apr && rr[i + 1] > apr) { p = (apr - rr[i]) / (rr[i + 1] - rr[i]); break; }
            ap_chord[ap] = (1. - p) * chord[i] + p * chord[i + 1];
            ap_twist[ap] = (1. - p) * twist[i] + p * twist[i + 1];
            for (int a = 0; a < nat; a++) {
                ap_cl[ap + (size_t)a * nap] = (1. - p) * cl[(size_t)i * nat + a] + p * cl[(size_t)(i + 1) * nat + a];
                ap_cd[ap + (size_t)a * nap] = (1. - p) * cd[(size_t)i * nat + a] + p * cd[(size_t)(i + 1) * nat + a];
            }
        }
    }
    return FALM_OK;
}

/* ========================================================= step driver */

struct OrcCase {
    falm_case_config cfg;
    falm_domain dom;
    size_t ncell;
    std::vector<double> u, uprev, uu, p, nut, ff, div, rhs, res, a, xyz, kx, g, ja;
    std::vector<double> x, y, z, hx, hy, hz;
    double maxdiag;
    void *solver, *alm;
    int64_t it;
    falm_face_bc ubc[6];
    int pbc[6], uubc[6];
};

static void case_faces(OrcCase *C) {
    const falm_case_config &c = C->cfg;
    for (int f = 0; f < 6; f++) {
        C->ubc[f].type = FALM_BC_NONE;
        C->ubc[f].value[0] = C->ubc[f].value[1] = C->ubc[f].value[2] = 0.0;
        C->pbc[f] = FALM_PBC_NONE;
        C->uubc[f] = 0;
    }
    if (c.kind == FALM_CASE_WINDTUNNEL) {
        /* app/windtunnel/bc.h + bcdevcall.cu */
        C->ubc[0].type = FALM_BC_INLET; C->ubc[0].value[0] = c.u_inflow;
        C->ubc[1].type = FALM_BC_CONVECTIVE;
        C->ubc[2].type = FALM_BC_SLIP; C->ubc[3].type = FALM_BC_SLIP;
        C->ubc[4].type = c.zminus_slip ? FALM_BC_SLIP : FALM_BC_WALL0;
        C->ubc[5].type = FALM_BC_SLIP;
        for (int f = 0; f < 6; f++) C->pbc[f] = FALM_PBC_NEUMANN;
        C->pbc[1] = FALM_PBC_ZERO;
    } else {
        /* app/lid3d/bcdev.cu, app/ld2d/bcdevcall.cu: no-slip reflection, lid on y+ */
        int nf = (c.kind == FALM_CASE_LID2D) ? 4 : 6;
        for (int f = 0; f < nf; f++) {
            C->ubc[f].type = FALM_BC_REFLECT;
            C->pbc[f] = FALM_PBC_NEUMANN;
            C->uubc[f] = 1;
        }
        C->ubc[3].value[0] = c.lid_velocity;
    }
}

int orc_case_create(const falm_case_config *cfg, void **out) {
    OrcCase *C = new OrcCase();
    C->cfg = *cfg;
    const int gc = 2;
    falm_domain &d = C->dom;
    for (int m = 0; m < 3; m++) {
        d.shape[m] = cfg->n[m] + 2 * gc;
        d.offset[m] = 0;
        d.gshape[m] = d.shape[m];
    }
    d.gc = gc;
    size_t n = (size_t)d.shape[0] * d.shape[1] * d.shape[2];
    C->ncell = n;
    C->u.assign(3 * n, 0.0); C->uprev.assign(3 * n, 0.0); C->uu.assign(3 * n, 0.0);
    C->p.assign(n, 0.0); C->nut.assign(n, 0.0); C->ff.assign(3 * n, 0.0); C->div.assign(n, 0.0);
    C->rhs.assign(n, 0.0); C->res.assign(n, 0.0); C->a.assign(7 * n, 0.0);
    C->xyz.assign(3 * n, 0.0); C->kx.assign(3 * n, 0.0); C->g.assign(3 * n, 0.0); C->ja.assign(n, 0.0);
    std::vector<double> *cen[3] = {&C->x, &C->y, &C->z}, *wid[3] = {&C->hx, &C->hy, &C->hz};
    for (int m = 0; m < 3; m++) {
        cen[m]->resize(d.shape[m]);
        wid[m]->resize(d.shape[m]);
        for (int i = 0; i < d.shape[m]; i++) {
            (*cen[m])[i] = cfg->cv[2 * m] ? cfg->cv[2 * m][i] : cfg->origin[m] + (i - gc + 0.5) * cfg->h[m];
            (*wid[m])[i] = cfg->cv[2 * m] ? cfg->cv[2 * m + 1][i] : cfg->h[m];
        }
    }
    for (int q = 0; q < 6; q++) C->cfg.cv[q] = nullptr; /* read during create only */
    orc_metrics(&d, C->x.data(), C->y.data(), C->z.data(), C->hx.data(), C->hy.data(), C->hz.data(),
                C->xyz.data(), C->kx.data(), C->g.data(), C->ja.data());
    /* initial field (app/windtunnel/main.cpp:121-129) */
    for (size_t c = 0; c < n; c++) { C->u[c] = cfg->u_inflow; C->u[c + n] = 0.0; C->u[c + 2 * n] = 0.0; }
    orc_u_to_uu(&d, C->u.data(), C->uu.data(), C->kx.data(), C->ja.data());
    orc_sgs(&d, cfg->sgs_model, cfg->Cs, C->u.data(), C->nut.data(), C->kx.data(), C->ja.data());
    int kind = (cfg->kind == FALM_CASE_WINDTUNNEL) ? FALM_POISSON_OUTLET_XPLUS
             : (cfg->kind == FALM_CASE_LID2D) ? FALM_POISSON_2D : FALM_POISSON_NEUMANN;
    orc_poisson_matrix(&d, kind, C->g.data(), C->ja.data(), C->a.data(), &C->maxdiag);
    orc_solver_create(&d, &cfg->solver, &C->solver);
    C->alm = nullptr;
    if (cfg->alm) orc_alm_create(&d, cfg->alm, C->x.data(), C->y.data(), C->z.data(), &C->alm);
    C->it = 0;
    case_faces(C);
    *out = C;
    return FALM_OK;
}

int orc_case_destroy(void *h) {
    OrcCase *C = (OrcCase *)h;
    orc_solver_destroy(C->solver);
    if (C->alm) orc_alm_destroy(C->alm);
    delete C;
    return FALM_OK;
}

/* one time step: app/nedo/main.cpp:150-204 (windtunnel), app/lid3d/main.cpp:122-175,
 * app/ld2d/main.cpp:112-136, with the pseudo-velocity read from the
 * u_previous snapshot (out of place) */
int orc_case_step(void *h, int nsteps, falm_step_info *info) {
    OrcCase *C = (OrcCase *)h;
    const falm_case_config &c = C->cfg;
    falm_domain *d = &C->dom;
    const size_t n = C->ncell;
    const bool is2d = c.kind == FALM_CASE_LID2D;
    falm_solve_info si{};
    double divn = 0;
    for (int s = 0; s < nsteps; s++) {
        C->it++;
        double t = c.dt * (double)C->it;
        std::copy(C->u.begin(), C->u.end(), C->uprev.begin());
        if (C->alm) orc_alm_step(C->alm, C->u.data(), C->ff.data(), t);
        orc_pseudo_u(d, c.adv_scheme, 1.0 / c.Re, c.dt, C->uprev.data(), C->uprev.data(), C->uu.data(),
                     C->u.data(), C->nut.data(), C->kx.data(), C->g.data(), C->ja.data(), C->ff.data());
        orc_u_to_uu(d, C->u.data(), C->uu.data(), C->kx.data(), C->ja.data());
        if (c.kind != FALM_CASE_WINDTUNNEL) orc_bc_face_flux(d, C->uubc, C->uu.data());
        orc_poisson_rhs(d, C->uu.data(), C->rhs.data(), C->ja.data(), c.dt, C->maxdiag);
        orc_solver_solve(C->solver, C->a.data(), C->p.data(), C->rhs.data(), C->res.data(), &si);
        orc_bc_pressure(d, C->pbc, C->p.data());
        if (is2d) orc_copy_z5(d, C->p.data(), 1);
        orc_project_p(d, c.dt, C->u.data(), C->u.data(), C->uu.data(), C->uu.data(), C->p.data(),
                      C->kx.data(), C->g.data());
        falm_inlet_noise nz{c.inlet_intensity, c.seed, C->it};
        orc_bc_velocity(d, C->ubc, C->u.data(), C->uprev.data(), C->xyz.data(), c.dt, &nz);
        if (is2d) orc_copy_z5(d, C->u.data(), 3);
        orc_sgs(d, c.sgs_model, c.Cs, C->u.data(), C->nut.data(), C->kx.data(), C->ja.data());
        if (is2d) orc_copy_z5(d, C->nut.data(), 1);
        orc_divergence(d, C->uu.data(), C->div.data(), C->ja.data());
        double nint = (double)(d->shape[0] - 2 * d->gc) * (d->shape[1] - 2 * d->gc) * (d->shape[2] - 2 * d->gc);
        divn = sqrt(orc_norm2(d, C->div.data(), 1)) / nint;
    }
    (void)n;
    if (info) {
        info->step = C->it;
        info->t = c.dt * (double)C->it;
        info->div = divn;
        info->solver_it = si.it;
        info->solver_err = si.err;
        info->outer_it = si.outer_it;
        info->status = si.status;
    }
    return FALM_OK;
}

static std::vector<double> *case_field(OrcCase *C, int field, int &ncol) {
    ncol = 1;
    switch (field) {
    case FALM_FIELD_U: ncol = 3; return &C->u;
    case FALM_FIELD_UU: ncol = 3; return &C->uu;
    case FALM_FIELD_P: return &C->p;
    case FALM_FIELD_NUT: return &C->nut;
    case FALM_FIELD_FF: ncol = 3; return &C->ff;
    case FALM_FIELD_DIV: return &C->div;
    case FALM_FIELD_RHS: return &C->rhs;
    case FALM_FIELD_POI_A: ncol = 7; return &C->a;
    case FALM_FIELD_XYZ: ncol = 3; return &C->xyz;
    case FALM_FIELD_KX: ncol = 3; return &C->kx;
    case FALM_FIELD_G: ncol = 3; return &C->g;
    case FALM_FIELD_JA: return &C->ja;
    default: return nullptr;
    }
}

int orc_case_get_field(void *h, int field, double *out, size_t n) {
    OrcCase *C = (OrcCase *)h;
    int ncol;
    std::vector<double> *v = case_field(C, field, ncol);
    if (!v || n != v->size()) return FALM_ERR_INVALID;
    std::copy(v->begin(), v->end(), out);
    return FALM_OK;
}

int orc_case_set_field(void *h, int field, const double *in, size_t n) {
    OrcCase *C = (OrcCase *)h;
    int ncol;
    std::vector<double> *v = case_field(C, field, ncol);
    if (!v || n != v->size()) return FALM_ERR_INVALID;
    std::copy(in, in + n, v->begin());
    return FALM_OK;
}

int orc_case_alm(void *h, void **alm) {
    *alm = ((OrcCase *)h)->alm;
    return FALM_OK;
}

/* ============================================================ adjoint
 *
 * Discrete adjoint of the wind-tunnel step with respect to the blade
 * pitch and per-section twist offsets, for the objective J = mean over a
 * window of steps of the rotor power P = sum_t Omega_t * torque_t
 * (AlmHandler::writePowerThrust, src/alm/alm.h:126-138).  No reference
 * counterpart exists: the reference has no adjoint.  Each stage below is
 * the transpose of the forward code above it refers to, written as a
 * scatter over the forward loop (sequential, FP64).  The Poisson solve is
 * taken as exact: its adjoint is a solve with the transposed operator.
 * Scope: wind-tunnel case, Upwind3-family advection, Smagorinsky, FP64. */

/* orc_sgs transposed: nut = (Cs J^(1/3))^2 sqrt(S(grad u)) */
static void adj_sgs(const Dom &D, double Cs, const double *u_, const double *kx_, const double *ja_,
                    const double *nutbar, double *ubar_) {
    Col u{u_, D.ncell}, kx{kx_, D.ncell}, ja{ja_, D.ncell};
    ColW ub{ubar_, D.ncell};
    const int gc = D.gc;
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                size_t c = D.id(i, j, k);
                const double w = nutbar[c];
                if (w == 0.0) continue;
                size_t nb[3][2] = {{D.id(i + 1, j, k), D.id(i - 1, j, k)}, {D.id(i, j + 1, k), D.id(i, j - 1, k)},
                                   {D.id(i, j, k + 1), D.id(i, j, k - 1)}};
                double gr[3][3];
                for (int d = 0; d < 3; d++)
                    for (int e = 0; e < 3; e++) gr[d][e] = 0.5 * kx(c, e) * (u(nb[e][0], d) - u(nb[e][1], d));
                const double dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
                const double dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
                const double dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
                const double S = 2 * SQ(dux) + 2 * SQ(dvy) + 2 * SQ(dwz) + SQ(duy + dvx) + SQ(dvz + dwy) + SQ(duz + dwx);
                const double Du = sqrt(S);
                if (!(Du > 0.0)) continue;  /* sqrt at 0: zero subgradient */
                const double lc = Cs * cbrt(ja(c));
                const double coef = w * SQ(lc) / (2.0 * Du);
                double dS[3][3];
                dS[0][0] = 4 * dux; dS[1][1] = 4 * dvy; dS[2][2] = 4 * dwz;
                dS[0][1] = dS[1][0] = 2 * (duy + dvx);
                dS[1][2] = dS[2][1] = 2 * (dvz + dwy);
                dS[0][2] = dS[2][0] = 2 * (duz + dwx);
                for (int d = 0; d < 3; d++)
                    for (int e = 0; e < 3; e++) {
                        const double gb = coef * dS[d][e] * 0.5 * kx(c, e);
                        ub(nb[e][0], d) += gb;
                        ub(nb[e][1], d) -= gb;
                    }
            }
}

/* orc_pseudo_u transposed (Upwind3 family, un == u): accumulates into
 * ubar, uubar, nutbar; writes ffbar = dt * uabar on the interior */
static void adj_pseudo_u(const Dom &D, double ALPHA, double ReI, double dt, const double *u_, const double *uu_,
                         const double *nut, const double *kx_, const double *g_, const double *ja_,
                         const double *uabar_, double *ubar_, double *uubar_, double *nutbar, double *ffbar_) {
    Col u{u_, D.ncell}, uu{uu_, D.ncell}, kx{kx_, D.ncell}, g{g_, D.ncell}, ja{ja_, D.ncell}, uab{uabar_, D.ncell};
    ColW ub{ubar_, D.ncell}, uub{uubar_, D.ncell}, ffb{ffbar_, D.ncell};
    const int gc = D.gc;
    const long st[3] = {1, (long)D.nx, (long)D.nx * D.ny};
    const double lco[5] = {0.0, 1. / 24., -27. / 24., 27. / 24., -1. / 24.};  /* m = -2..2 in UDL */
    const double rco[5] = {1. / 24., -27. / 24., 27. / 24., -1. / 24., 0.0};  /* in UDR */
    const double dco[5] = {1., -4., 6., -4., 1.};
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                const double J = ja(c);
                for (int d = 0; d < 3; d++) {
                    const double w = uab(c, d);
                    ffb(c, d) = dt * w;
                    if (w == 0.0) continue;
                    ub(c, d) += w;  /* ua = un + ... */
                    const double cw = -dt * w;
                    for (int e = 0; e < 3; e++) {
                        const long s_ = st[e];
                        const double U = u(c, e) * kx(c, e), Uabs = fabs(U);
                        const double JP = uu(c, e), JM = uu(c - s_, e);
                        double v[5];
                        for (int m = 0; m < 5; m++) v[m] = u(c + (m - 2) * s_, d);
                        const double L = (-v[4] + 27 * v[3] - 27 * v[2] + v[1]) / 24.;
                        const double R = (-v[3] + 27 * v[2] - 27 * v[1] + v[0]) / 24.;
                        const double D4 = v[4] - 4 * v[3] + 6 * v[2] - 4 * v[1] + v[0];
                        for (int m = 0; m < 5; m++)
                            ub(c + (m - 2) * s_, d) += cw * (0.5 * (JP * lco[m] + JM * rco[m]) / J + Uabs * ALPHA * dco[m]);
                        uub(c, e) += cw * 0.5 * L / J;
                        uub(c - s_, e) += cw * 0.5 * R / J;
                        const double sg = (U > 0) - (U < 0);
                        ub(c, e) += cw * ALPHA * D4 * sg * kx(c, e);
                    }
                    const double cv = dt * w;
                    for (int e = 0; e < 3; e++)
                        for (int sgn = -1; sgn <= 1; sgn += 2) {
                            const size_t f = c + sgn * st[e];
                            const double a = (g(c, e) + g(f, e)) / (2 * J);
                            const double b = ReI + 0.5 * (nut[c] + nut[f]);
                            const double diff = u(f, d) - u(c, d);
                            ub(f, d) += cv * a * b;
                            ub(c, d) -= cv * a * b;
                            nutbar[c] += cv * a * 0.5 * diff;
                            nutbar[f] += cv * a * 0.5 * diff;
                        }
                }
            }
}

/* orc_u_to_uu transposed: uu_d(c) = (ja kx_d ua_d)(c)/2 + (ja kx_d ua_d)(c+e_d)/2 on the face region */
static void adj_u_to_uu(const Dom &D, const double *kx_, const double *ja_, const double *uubar_, double *uabar_) {
    Col kx{kx_, D.ncell}, ja{ja_, D.ncell}, uub{uubar_, D.ncell};
    ColW uab{uabar_, D.ncell};
    const long st[3] = {1, (long)D.nx, (long)D.nx * D.ny};
    const int lo = D.gc - 1;
    for (int k = lo; k < D.nz - D.gc; k++)
        for (int j = lo; j < D.ny - D.gc; j++)
            for (int i = lo; i < D.nx - D.gc; i++) {
                const size_t c = D.id(i, j, k);
                for (int d = 0; d < 3; d++) {
                    const double w = uub(c, d);
                    const size_t cn = c + st[d];
                    uab(c, d) += 0.5 * ja(c) * kx(c, d) * w;
                    uab(cn, d) += 0.5 * ja(cn) * kx(cn, d) * w;
                }
            }
}

/* orc_poisson_rhs transposed (interior rhs; ghost rhs are 0 * scale) */
static void adj_rhs(const Dom &D, const double *ja_, double scale, const double *rhsbar, double *uubar_) {
    Col ja{ja_, D.ncell};
    ColW uub{uubar_, D.ncell};
    const long st[3] = {1, (long)D.nx, (long)D.nx * D.ny};
    const int gc = D.gc;
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                const double w = rhsbar[c] * scale / ja(c);
                for (int d = 0; d < 3; d++) {
                    uub(c, d) += w;
                    uub(c - st[d], d) -= w;
                }
            }
}

/* orc_project_p transposed: the cell part (interior) and the face part
 * (face region); pressure adjoint accumulates in pbar */
static void adj_project(const Dom &D, double dt, const double *kx_, const double *g_, const double *ustarbar_,
                        const double *uunbar_, double *uabar_, double *uustarbar_, double *pbar) {
    Col kx{kx_, D.ncell}, g{g_, D.ncell}, usb{ustarbar_, D.ncell}, uunb{uunbar_, D.ncell};
    ColW uab{uabar_, D.ncell}, uusb{uustarbar_, D.ncell};
    const long st[3] = {1, (long)D.nx, (long)D.nx * D.ny};
    const int gc = D.gc;
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                for (int d = 0; d < 3; d++) {
                    const double w = usb(c, d);
                    uab(c, d) += w;
                    const double q = dt * 0.5 * kx(c, d) * w;
                    pbar[c + st[d]] -= q;
                    pbar[c - st[d]] += q;
                }
            }
    const int lo = gc - 1;
    for (int k = lo; k < D.nz - gc; k++)
        for (int j = lo; j < D.ny - gc; j++)
            for (int i = lo; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                for (int d = 0; d < 3; d++) {
                    const double w = uunb(c, d);
                    uusb(c, d) += w;
                    const double q = dt * 0.5 * (g(c, d) + g(c + st[d], d)) * w;
                    pbar[c + st[d]] -= q;
                    pbar[c] += q;
                }
            }
}

/* orc_bc_pressure transposed: Neumann ghost = interior neighbour, ZERO ghost = 0 */
static void adj_pbc(const Dom &D, const int faces[6], double *pbar) {
    const int gc = D.gc;
    const int shp[3] = {D.nx, D.ny, D.nz};
    for (int f = 0; f < 6; f++) {
        if (faces[f] == FALM_PBC_NONE) continue;
        int d, ax, bx;
        face_axes(f, d, ax, bx);
        for (int kb = gc; kb < shp[bx] - gc; kb++)
            for (int ka = gc; ka < shp[ax] - gc; ka++) {
                int o1[3], i1[3];
                o1[ax] = i1[ax] = ka;
                o1[bx] = i1[bx] = kb;
                if (f & 1) { o1[d] = shp[d] - gc; i1[d] = shp[d] - gc - 1; }
                else       { o1[d] = gc - 1;      i1[d] = gc; }
                const size_t co = D.id(o1[0], o1[1], o1[2]), ci = D.id(i1[0], i1[1], i1[2]);
                if (faces[f] == FALM_PBC_NEUMANN) pbar[ci] += pbar[co];
                pbar[co] = 0.0;
            }
    }
}

/* orc_bc_velocity transposed for the wind-tunnel face types.  unp1bar: the
 * adjoint of u^{n+1} (all cells); ustarbar: out, its interior part plus the
 * ghost cells' dependence on the interior; unbar: accumulates the
 * convective outlet's dependence on u^n.  Ghost cells no face writes keep
 * their initial value and carry no dependence. */
static void adj_ubc(const Dom &D, const falm_face_bc faces[6], const double *xyz_, double dt, const double *unp1bar_,
                    double *ustarbar_, double *unbar_) {
    Col xyz{xyz_, D.ncell}, unb1{unp1bar_, D.ncell};
    ColW usb{ustarbar_, D.ncell}, unb{unbar_, D.ncell};
    const int gc = D.gc;
    const int shp[3] = {D.nx, D.ny, D.nz};
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                for (int m = 0; m < 3; m++) usb(c, m) = unb1(c, m);
            }
    for (int f = 0; f < 6; f++) {
        const falm_face_bc &bc = faces[f];
        if (bc.type == FALM_BC_NONE || bc.type == FALM_BC_INLET) continue;
        int d, ax, bx;
        face_axes(f, d, ax, bx);
        const int plus = f & 1;
        for (int kb = gc; kb < shp[bx] - gc; kb++)
            for (int ka = gc; ka < shp[ax] - gc; ka++) {
                int o1[3], o2[3], i1[3], i2[3];
                o1[ax] = o2[ax] = i1[ax] = i2[ax] = ka;
                o1[bx] = o2[bx] = i1[bx] = i2[bx] = kb;
                if (plus) {
                    o1[d] = shp[d] - gc; o2[d] = shp[d] - gc + 1; i1[d] = shp[d] - gc - 1; i2[d] = shp[d] - gc - 2;
                } else {
                    o1[d] = gc - 1; o2[d] = gc - 2; i1[d] = gc; i2[d] = gc + 1;
                }
                const size_t co1 = D.id(o1[0], o1[1], o1[2]), co2 = D.id(o2[0], o2[1], o2[2]);
                const size_t ci1 = D.id(i1[0], i1[1], i1[2]), ci2 = D.id(i2[0], i2[1], i2[2]);
                for (int m = 0; m < 3; m++) {
                    const double w1 = unb1(co1, m), w2 = unb1(co2, m);
                    if (bc.type == FALM_BC_CONVECTIVE) {
                        const double dxi = 1.0 / (xyz(co1, 0) - xyz(ci1, 0));
                        const double W = w1 + 2 * w2;   /* u(o2) = 2 u(o1) - u*(i1) */
                        usb(ci1, m) -= w2;
                        unb(co1, m) += (1.0 - dt * 0.5 * dxi * 3) * W;
                        unb(ci1, m) += (dt * 0.5 * dxi * 4) * W;
                        unb(ci2, m) += (-dt * 0.5 * dxi) * W;
                    } else if (bc.type == FALM_BC_SLIP) {
                        if (m == d) usb(ci1, m) -= w2;      /* u(o1) = 0, u(o2) = -u*(i1) */
                        else usb(ci1, m) += w1 + w2;        /* u(o1) = u(o2) = u*(i1) */
                    } else if (bc.type == FALM_BC_WALL0) {
                        usb(ci1, m) -= w2;
                    }
                }
            }
    }
}

static void rot_T(const double y[3], const double ang[3], int type, double x[3]) {
    /* x = R^T y for the rotation of rot() */
    if (type == 1) {
        double C = cos(ang[0]), S = sin(ang[0]);
        x[0] = y[0]; x[1] = C * y[1] - S * y[2]; x[2] = S * y[1] + C * y[2];
    } else if (type == 2) {
        double C = cos(ang[1]), S = sin(ang[1]);
        x[0] = C * y[0] + S * y[2]; x[1] = y[1]; x[2] = -S * y[0] + C * y[2];
    } else if (type == 3) {
        double C = cos(ang[2]), S = sin(ang[2]);
        x[0] = C * y[0] - S * y[1]; x[1] = S * y[0] + C * y[1]; x[2] = y[2];
    } else {
        x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
    }
}

/* orc_alm_step transposed at state u, time t: ffbar is the adjoint of the
 * projected force field, wP the weight of this step's power in J.
 * Accumulates ubar and the parameter adjoints (per degree). */
static void adj_alm(OrcAlm *A, const double *u_, double t, const double *ffbar_, double wP, double *ubar_,
                    double *gpitch /* [nt*8] */, double *gtwist /* [napb] */) {
    Dom D(&A->dom);
    const int gc = D.gc;
    Col uvw{u_, D.ncell}, ffb{ffbar_, D.ncell};
    ColW ub{ubar_, D.ncell};
    const int napt = A->napb * A->nb;
    const double eps = A->eps;
    const double eta = 1. / ((eps * sqrt(PI)) * (eps * sqrt(PI)) * (eps * sqrt(PI)));
    const double dr = (A->radius - A->hub_radius) / A->napb;
    for (int ap = 0; ap < A->nap; ap++) {
        const int tid = ap / napt, bid = (ap % napt) / A->napb, sec = ap % A->napb;
        const falm_turbine &T = A->tb[tid];
        double ang[3] = {0, 0, 0}, om[3] = {0, 0, 0};
        if (T.angle_type) {
            const double sa = T.motion[0] * sin(T.motion[1] * t + T.motion[2]);
            ang[T.angle_type - 1] = T.angle_offset_deg != 0.0 ? deg2rad(T.angle_offset_deg) + sa : sa;
            om[T.angle_type - 1] = T.motion[0] * T.motion[1] * cos(T.motion[1] * t + T.motion[2]);
        }
        const double *sv = &A->set_vel[3 * tid];
        double base[3] = {T.base[0], T.base[1], T.base[2]}, bvel[3] = {sv[0], sv[1], sv[2]};
        if (T.surge[0] != 0.0) {
            base[0] = T.base[0] + T.surge[0] * sin(T.surge[1] * t + T.surge[2]);
            bvel[0] = sv[0] + T.surge[0] * T.surge[1] * cos(T.surge[1] * t + T.surge[2]);
        }
        const double nang[3] = {-ang[0], -ang[1], -ang[2]};
        const double theta = T.tip_rate * t + (2 * PI / A->nb) * bid;
        double c1[3] = {T.hub[0], T.hub[1] + A->r[ap] * cos(theta), T.hub[2] + A->r[ap] * sin(theta)};
        double P[3];
        yawed_to_inertial(c1, T.yaw_deg, nang, T.angle_type, P);
        for (int m = 0; m < 3; m++) P[m] += base[m];
        const int I = find_index(A->x.data(), P[0], D.nx), Jj = find_index(A->y.data(), P[1], D.ny),
                  K = find_index(A->z.data(), P[2], D.nz);
        const bool in = I >= gc - 1 && I < D.nx - gc - 1 && Jj >= gc - 1 && Jj < D.ny - gc - 1 && K >= gc - 1 &&
                        K < D.nz - gc - 1;
        if (!in) continue;
        /* forward values */
        const int i0 = I, i1 = I + 1, j0 = Jj, j1 = Jj + 1, k0 = K, k1 = K + 1;
        const double xd = fabs(P[0] - A->x[i0]) / fabs(A->x[i1] - A->x[i0]);
        const double yd = fabs(P[1] - A->y[j0]) / fabs(A->y[j1] - A->y[j0]);
        const double zd = fabs(P[2] - A->z[k0]) / fabs(A->z[k1] - A->z[k0]);
        const size_t id[8] = {D.id(i0, j0, k0), D.id(i1, j0, k0), D.id(i0, j1, k0), D.id(i1, j1, k0),
                              D.id(i0, j0, k1), D.id(i1, j0, k1), D.id(i0, j1, k1), D.id(i1, j1, k1)};
        const double wt[8] = {(1 - xd) * (1 - yd) * (1 - zd), xd * (1 - yd) * (1 - zd), (1 - xd) * yd * (1 - zd),
                              xd * yd * (1 - zd), (1 - xd) * (1 - yd) * zd, xd * (1 - yd) * zd,
                              (1 - xd) * yd * zd, xd * yd * zd};
        double uap[3];
        for (int m = 0; m < 3; m++)
            uap[m] = trilinear(xd, yd, zd, uvw(id[0], m), uvw(id[1], m), uvw(id[2], m), uvw(id[3], m),
                               uvw(id[4], m), uvw(id[5], m), uvw(id[6], m), uvw(id[7], m));
        double rel[3], vrel[3], utt[3];
        for (int m = 0; m < 3; m++) {
            rel[m] = P[m] - base[m];
            vrel[m] = uap[m] - bvel[m];
        }
        rot_dt(rel, vrel, ang, om, T.angle_type, utt);
        platform_to_rotor(utt, T.yaw_deg);
        const double ux = utt[0];
        const double ut = T.tip_rate * A->r[ap] + utt[1] * sin(theta) - utt[2] * cos(theta);
        const double urel2 = ux * ux + ut * ut;
        const double phi = atan(ux / ut);
        const double chord = A->chord[ap];
        const double twist = A->twist[ap] + T.pitch_deg[bid < 8 ? bid : 7] + A->twist_off[sec];
        const double att = rad2deg(phi) - twist;
        double cl = 0, cd = 0, scl = 0, scd = 0;
        orc_polar_lookup(A->attack.data(), A->nat, A->cl.data(), A->cd.data(), A->nap, ap, att, &cl, &cd);
        if (att >= A->attack[0] && att < A->attack[A->nat - 1])
            for (int a = 0; a < A->nat - 1; a++)
                if (A->attack[a] <= att && A->attack[a + 1] > att) {
                    const double h = A->attack[a + 1] - A->attack[a];
                    scl = (A->cl[ap + (size_t)(a + 1) * A->nap] - A->cl[ap + (size_t)a * A->nap]) / h;
                    scd = (A->cd[ap + (size_t)(a + 1) * A->nap] - A->cd[ap + (size_t)a * A->nap]) / h;
                    break;
                }
        const double fl = .5 * cl * urel2 * chord * dr;
        const double fd = .5 * cd * urel2 * chord * dr;
        const bool fix = (A->flags & FALM_ALM_FIX_DRAG_TANGENT) != 0;
        const double ft0 = fix ? fl * sin(phi) - fd * cos(phi) : fl * sin(phi) - fd * sin(phi);
        const int sg = (T.tip_rate > 0) - (T.tip_rate < 0);
        const double ft = sg * ft0;
        /* adjoint of the projected force: F_bar = sum_c ff_bar(c) eta exp(-r^2/eps^2) */
        double Fb[3] = {0, 0, 0};
        for (int k = gc; k < D.nz - gc; k++)
            for (int j = gc; j < D.ny - gc; j++)
                for (int i = gc; i < D.nx - gc; i++) {
                    const size_t c = D.id(i, j, k);
                    const double rr2 = SQ(A->x[i] - P[0]) + SQ(A->y[j] - P[1]) + SQ(A->z[k] - P[2]);
                    const double e = eta * exp(-rr2 / SQ(eps));
                    for (int m = 0; m < 3; m++) Fb[m] += ffb(c, m) * e;
                }
        /* F = R(nang) (-fx, ft sin(theta), -ft cos(theta)) */
        double fttb[3];
        rot_T(Fb, nang, T.angle_type, fttb);
        if (T.yaw_deg != 0.0) { /* transpose of the yaw turn */
            const double ny[3] = {0.0, 0.0, -deg2rad(T.yaw_deg)};
            const double w[3] = {fttb[0], fttb[1], fttb[2]};
            rot_T(w, ny, 3, fttb);
        }
        double fxb = -fttb[0];
        double ftb = fttb[1] * sin(theta) - fttb[2] * cos(theta);
        ftb += wP * T.tip_rate * A->r[ap] * ((ft > 0) - (ft < 0));  /* power: Omega |ft| r */
        const double ft0b = sg * ftb;
        const double flb = cos(phi) * fxb + sin(phi) * ft0b;
        const double fdb = sin(phi) * fxb + (fix ? -cos(phi) : -sin(phi)) * ft0b;
        double phib = (-fl * sin(phi) + fd * cos(phi)) * fxb + (fix ? (fl * cos(phi) + fd * sin(phi))
                                                                    : (fl * cos(phi) - fd * cos(phi))) * ft0b;
        const double clb = .5 * urel2 * chord * dr * flb, cdb = .5 * urel2 * chord * dr * fdb;
        const double urel2b = .5 * cl * chord * dr * flb + .5 * cd * chord * dr * fdb;
        const double attb = scl * clb + scd * cdb;
        phib += attb * 180. / PI;
        gpitch[tid * 8 + (bid < 8 ? bid : 7)] += -attb;
        gtwist[sec] += -attb;
        const double den = ux * ux + ut * ut;
        double uxb = phib * ut / den + 2 * ux * urel2b;
        double utb = -phib * ux / den + 2 * ut * urel2b;
        double uttb[3] = {uxb, sin(theta) * utb, -cos(theta) * utb};
        if (T.yaw_deg != 0.0) { /* transpose of platform_to_rotor */
            const double py[3] = {0.0, 0.0, deg2rad(T.yaw_deg)};
            const double w[3] = {uttb[0], uttb[1], uttb[2]};
            rot_T(w, py, 3, uttb);
        }
        double vrelb[3];
        rot_T(uttb, ang, T.angle_type, vrelb);
        for (int q = 0; q < 8; q++)
            for (int m = 0; m < 3; m++) ub(id[q], m) += wt[q] * vrelb[m];
    }
}


/* stage adjoints with C signatures (tests) */
int orc_adj_sgs(const falm_domain *dom, double Cs, const double *u, const double *kx, const double *ja,
                const double *nutbar, double *ubar) {
    adj_sgs(Dom(dom), Cs, u, kx, ja, nutbar, ubar);
    return FALM_OK;
}
int orc_adj_pseudo_u(const falm_domain *dom, double ALPHA, double ReI, double dt, const double *u, const double *uu,
                     const double *nut, const double *kx, const double *g, const double *ja, const double *uabar,
                     double *ubar, double *uubar, double *nutbar, double *ffbar) {
    adj_pseudo_u(Dom(dom), ALPHA, ReI, dt, u, uu, nut, kx, g, ja, uabar, ubar, uubar, nutbar, ffbar);
    return FALM_OK;
}
int orc_adj_u_to_uu(const falm_domain *dom, const double *kx, const double *ja, const double *uubar, double *uabar) {
    adj_u_to_uu(Dom(dom), kx, ja, uubar, uabar);
    return FALM_OK;
}
int orc_adj_project(const falm_domain *dom, double dt, const double *kx, const double *g, const double *ustarbar,
                    const double *uunbar, double *uabar, double *uustarbar, double *pbar) {
    adj_project(Dom(dom), dt, kx, g, ustarbar, uunbar, uabar, uustarbar, pbar);
    return FALM_OK;
}
int orc_adj_ubc(const falm_domain *dom, const falm_face_bc faces[6], const double *xyz, double dt,
                const double *unp1bar, double *ustarbar, double *unbar) {
    adj_ubc(Dom(dom), faces, xyz, dt, unp1bar, ustarbar, unbar);
    return FALM_OK;
}
int orc_adj_rhs(const falm_domain *dom, const double *ja, double scale, const double *rhsbar, double *uubar) {
    adj_rhs(Dom(dom), ja, scale, rhsbar, uubar);
    return FALM_OK;
}
int orc_adj_pbc(const falm_domain *dom, const int faces[6], double *pbar) {
    adj_pbc(Dom(dom), faces, pbar);
    return FALM_OK;
}

static double step_power(OrcCase *C) {
    if (!C->alm) return 0.0;
    OrcAlm *A = (OrcAlm *)C->alm;
    double P = 0.0;
    for (int t = 0; t < A->nt; t++) P += A->tb[t].tip_rate * A->torque[t];
    return P;
}

struct CaseSnap {
    std::vector<double> u, uu, p, nut;
    int64_t it;
};
static CaseSnap snap(OrcCase *C) { return CaseSnap{C->u, C->uu, C->p, C->nut, C->it}; }
static void restore(OrcCase *C, const CaseSnap &S) {
    C->u = S.u; C->uu = S.uu; C->p = S.p; C->nut = S.nut; C->it = S.it;
}

/* J = mean rotor power over the next nsteps steps; the case is left as it was */
int orc_case_window(void *h, int nsteps, double *J) {
    OrcCase *C = (OrcCase *)h;
    if (nsteps < 1 || !C->alm) return FALM_ERR_INVALID;
    CaseSnap S0 = snap(C);
    double sum = 0.0;
    for (int s = 0; s < nsteps; s++) {
        orc_case_step(h, 1, nullptr);
        sum += step_power(C);
    }
    restore(C, S0);
    *J = sum / nsteps;
    return FALM_OK;
}

/* the transposed Poisson operator (same [c,w,e,s,n,b,t] layout) */
static void transpose7(const Dom &D, const double *a_, double *at_) {
    Col a{a_, D.ncell};
    ColW at{at_, D.ncell};
    const int gc = D.gc;
    const long st[3] = {1, (long)D.nx, (long)D.nx * D.ny};
    std::fill(at_, at_ + 7 * D.ncell, 0.0);
    auto interior = [&](size_t c) {
        const int k = (int)(c / ((size_t)D.nx * D.ny)), r = (int)(c % ((size_t)D.nx * D.ny));
        const int j = r / D.nx, i = r % D.nx;
        return i >= gc && i < D.nx - gc && j >= gc && j < D.ny - gc && k >= gc && k < D.nz - gc;
    };
    for (int k = gc; k < D.nz - gc; k++)
        for (int j = gc; j < D.ny - gc; j++)
            for (int i = gc; i < D.nx - gc; i++) {
                const size_t c = D.id(i, j, k);
                at(c, 0) = a(c, 0);
                for (int d = 0; d < 3; d++) {
                    /* column (1 + 2d) couples c to c - st (w/s/b), column (2 + 2d) to c + st (e/n/t) */
                    const size_t cm = c - st[d], cp = c + st[d];
                    at(c, 1 + 2 * d) = interior(cm) ? a(cm, 2 + 2 * d) : 0.0;
                    at(c, 2 + 2 * d) = interior(cp) ? a(cp, 1 + 2 * d) : 0.0;
                }
            }
}

/* J and dJ/d(pitch, twist offset) in 1/degree by the checkpointed discrete
 * adjoint over the next nsteps steps (checkpoints every ckpt steps); the
 * adjoint Poisson solves stop at |r|/N <= adj_rtol * |rhs|/N.  The case is
 * left as it was. */
int orc_case_adjoint(void *h, int nsteps, int ckpt, double adj_rtol, double *J, double *g_pitch /* [nt*8] */,
                     double *g_twist /* [napb] */, double *u0bar /* [3*ncell] or NULL: dJ/du at the start */) {
    OrcCase *C = (OrcCase *)h;
    const falm_case_config &c = C->cfg;
    if (nsteps < 1 || ckpt < 1 || !C->alm || c.kind != FALM_CASE_WINDTUNNEL || c.precision != FALM_CASE_F64)
        return FALM_ERR_INVALID;
    if (c.adv_scheme != FALM_ADV_UPWIND3 && c.adv_scheme != FALM_ADV_UTOPIA && c.adv_scheme != FALM_ADV_KK)
        return FALM_ERR_UNSUPPORTED;
    if (c.sgs_model != FALM_SGS_SMAGORINSKY && c.sgs_model != FALM_SGS_EMPTY) return FALM_ERR_UNSUPPORTED;
    const double ALPHA = c.adv_scheme == FALM_ADV_UPWIND3 ? 1. / 24. : c.adv_scheme == FALM_ADV_UTOPIA ? 1. / 12. : 1. / 4.;
    OrcAlm *A = (OrcAlm *)C->alm;
    Dom D(&C->dom);
    const size_t n = C->ncell;
    CaseSnap S0 = snap(C);
    /* forward with checkpoints */
    std::vector<CaseSnap> cps;
    double sum = 0.0;
    for (int s = 0; s < nsteps; s++) {
        if (s % ckpt == 0) cps.push_back(snap(C));
        orc_case_step(h, 1, nullptr);
        sum += step_power(C);
    }
    *J = sum / nsteps;
    const double wP = 1.0 / nsteps;
    /* transposed operator and its solver */
    std::vector<double> at(7 * n);
    transpose7(D, C->a.data(), at.data());
    falm_solver_config scfg = c.solver;
    scfg.type = FALM_LS_PBICGSTAB;
    scfg.maxit = 100000;
    void *solverT = nullptr;
    orc_solver_create(&C->dom, &scfg, &solverT);
    OrcSolver *ST = (OrcSolver *)solverT;
    std::vector<double> gp(A->nt * 8, 0.0), gt(A->napb, 0.0);
    std::vector<double> ub(3 * n, 0.0), uub(3 * n, 0.0);  /* adjoints of u^{n+1}, uu^{n+1} */
    std::vector<double> usb(3 * n), uab(3 * n), uusb(3 * n), pb(n), rhsb(n), lam(n), res(n), nutb(n), ffb(3 * n);
    std::vector<double> unb(3 * n), uunb(3 * n);
    const int nseg = (int)cps.size();
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * ckpt, s1 = std::min(nsteps, s0 + ckpt);
        restore(C, cps[sg]);
        std::vector<CaseSnap> states;
        for (int s = s0; s < s1; s++) {
            states.push_back(snap(C));
            if (s + 1 < s1) orc_case_step(h, 1, nullptr);
        }
        for (int s = s1 - 1; s >= s0; s--) {
            const CaseSnap &Sn = states[s - s0];
            const double t = c.dt * (double)(Sn.it + 1);
            /* nut^n = SGS(u^n) */
            std::vector<double> nutn(n, 0.0);
            orc_sgs(&C->dom, c.sgs_model, c.Cs, Sn.u.data(), nutn.data(), C->kx.data(), C->ja.data());
            std::fill(unb.begin(), unb.end(), 0.0);
            std::fill(uunb.begin(), uunb.end(), 0.0);
            std::fill(usb.begin(), usb.end(), 0.0);
            std::fill(uab.begin(), uab.end(), 0.0);
            std::fill(uusb.begin(), uusb.end(), 0.0);
            std::fill(pb.begin(), pb.end(), 0.0);
            std::fill(nutb.begin(), nutb.end(), 0.0);
            /* ubc, projection, pbc */
            adj_ubc(D, C->ubc, C->xyz.data(), c.dt, ub.data(), usb.data(), unb.data());
            adj_project(D, c.dt, C->kx.data(), C->g.data(), usb.data(), uub.data(), uab.data(), uusb.data(), pb.data());
            adj_pbc(D, C->pbc, pb.data());
            /* Poisson: rhs_bar = A^-T p_bar (interior) */
            double nrm = 0.0;
            for (int k = D.gc; k < D.nz - D.gc; k++)
                for (int j = D.gc; j < D.ny - D.gc; j++)
                    for (int i = D.gc; i < D.nx - D.gc; i++) nrm += SQ(pb[D.id(i, j, k)]);
            std::fill(rhsb.begin(), rhsb.end(), 0.0);
            if (nrm > 0.0) {
                const double nint = (double)(D.nx - 2 * D.gc) * (D.ny - 2 * D.gc) * (D.nz - 2 * D.gc);
                ST->cfg.tol = adj_rtol * sqrt(nrm) / nint;
                std::fill(lam.begin(), lam.end(), 0.0);
                std::vector<double> pbi(n, 0.0);
                for (int k = D.gc; k < D.nz - D.gc; k++)
                    for (int j = D.gc; j < D.ny - D.gc; j++)
                        for (int i = D.gc; i < D.nx - D.gc; i++) pbi[D.id(i, j, k)] = pb[D.id(i, j, k)];
                falm_solve_info si{};
                orc_solver_solve(solverT, at.data(), lam.data(), pbi.data(), res.data(), &si);
                rhsb = lam;
            }
            /* rhs, face flux, ghost copy of ua, pseudo velocity */
            adj_rhs(D, C->ja.data(), 1.0 / (c.dt * C->maxdiag), rhsb.data(), uusb.data());
            adj_u_to_uu(D, C->kx.data(), C->ja.data(), uusb.data(), uab.data());
            for (size_t q = 0; q < n; q++) {
                const int k = (int)(q / ((size_t)D.nx * D.ny)), r = (int)(q % ((size_t)D.nx * D.ny));
                const int j = r / D.nx, i = r % D.nx;
                const bool inner = i >= D.gc && i < D.nx - D.gc && j >= D.gc && j < D.ny - D.gc && k >= D.gc &&
                                   k < D.nz - D.gc;
                if (!inner)
                    for (int m = 0; m < 3; m++) {
                        unb[q + m * n] += uab[q + m * n];  /* ua ghosts are u^n ghosts */
                        uab[q + m * n] = 0.0;
                    }
            }
            adj_pseudo_u(D, ALPHA, 1.0 / c.Re, c.dt, Sn.u.data(), Sn.uu.data(), nutn.data(), C->kx.data(),
                         C->g.data(), C->ja.data(), uab.data(), unb.data(), uunb.data(), nutb.data(), ffb.data());
            /* ALM at u^n and this step's power */
            adj_alm(A, Sn.u.data(), t, ffb.data(), wP, unb.data(), gp.data(), gt.data());
            /* nut^n = SGS(u^n); ghost nut are constant */
            if (c.sgs_model == FALM_SGS_SMAGORINSKY) adj_sgs(D, c.Cs, Sn.u.data(), C->kx.data(), C->ja.data(), nutb.data(), unb.data());
            ub.swap(unb);
            uub.swap(uunb);
        }
    }
    orc_solver_destroy(solverT);
    restore(C, S0);
    if (u0bar) std::copy(ub.begin(), ub.end(), u0bar);
    for (int q = 0; q < A->nt * 8; q++) g_pitch[q] = gp[q];
    for (int q = 0; q < A->napb; q++) g_twist[q] = gt[q];
    return FALM_OK;
}

/* adj_alm with a C signature (tests): accumulates ubar, gpitch, gtwist */
int orc_alm_vjp(void *h, const double *u, double t, const double *ffbar, double wP, double *ubar, double *gpitch,
                double *gtwist) {
    adj_alm((OrcAlm *)h, u, t, ffbar, wP, ubar, gpitch, gtwist);
    return FALM_OK;
}

int orc_alm_set_pitch(void *h, int turbine, const double *pitch_deg) {
    OrcAlm *A = (OrcAlm *)h;
    if (turbine < 0 || turbine >= A->nt) return FALM_ERR_INVALID;
    for (int b = 0; b < 8; b++) A->tb[turbine].pitch_deg[b] = pitch_deg ? pitch_deg[b] : 0.0;
    return FALM_OK;
}

/* static yaw and platform angle offset of one turbine (degrees; no
 * reference counterpart) */
int orc_alm_set_attitude(void *h, int turbine, double yaw_deg, double angle_offset_deg) {
    OrcAlm *A = (OrcAlm *)h;
    if (turbine < 0 || turbine >= A->nt) return FALM_ERR_INVALID;
    A->tb[turbine].yaw_deg = yaw_deg;
    A->tb[turbine].angle_offset_deg = angle_offset_deg;
    return FALM_OK;
}

/* base position and velocity of one turbine from now on (no reference
 * counterpart: the reference parses a file name for a base coordinate but
 * never reads the file, src/alm/turbineHandler.h:144-150) */
int orc_alm_set_base(void *h, int turbine, const double *base, const double *velocity) {
    OrcAlm *A = (OrcAlm *)h;
    if (turbine < 0 || turbine >= A->nt || !base) return FALM_ERR_INVALID;
    for (int m = 0; m < 3; m++) {
        A->tb[turbine].base[m] = base[m];
        A->set_vel[3 * turbine + m] = velocity ? velocity[m] : 0.0;
    }
    return FALM_OK;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

} /* extern "C" */
