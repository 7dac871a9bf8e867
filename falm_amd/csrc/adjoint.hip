// adjoint.hip — discrete adjoint of the device-resident wind-tunnel step
// for the gradient of the window-mean rotor power with respect to blade pitch
// and per-section twist offsets.  No reference counterpart exists (the
// reference has no adjoint).
//
// Every stage is the transpose of the forward kernel it mirrors, computed in
// double.  The linear stages (projection, Poisson right-hand side, face
// flux) gather: each thread takes one adjoint output cell and sums what the
// forward outputs that read it contribute (deterministic, no atomics).  The
// others scatter over the forward's loop: each thread takes one forward
// output cell and adds its contributions to the input adjoints with hardware
// double atomics (so sums are order-independent up to rounding).  The Poisson
// solve is taken as exact; its adjoint is a BiCGStab solve with the
// transposed operator.  Time is handled by checkpoints every `ckpt` steps and
// recomputation of each segment's states during the reverse sweep.
//
// Scope: FALM_CASE_WINDTUNNEL, FALM_CASE_F64, the Upwind3 family of
// advection schemes, Smagorinsky or no SGS model.
#include <cmath>
#include <initializer_list>
#include <utility>
#include <vector>

#include "alm.hpp"
#include "case.hpp"
#include "linalg.hpp"

namespace falm {

// cell of a box from a flat index (x fastest)
__device__ inline bool box_cell(const Box &b, long long t, int &i, int &j, int &k) {
    if (t >= b.size()) return false;
    i = b.lo[0] + (int)(t % b.n[0]);
    const long long r = t / b.n[0];
    j = b.lo[1] + (int)(r % b.n[1]);
    k = b.lo[2] + (int)(r / b.n[1]);
    return true;
}
inline dim3 box_grid(const Box &b) { return ew_grid((size_t)b.size()); }
__device__ inline bool in_box(const Box &b, int i, int j, int k) {
    return i >= b.lo[0] && i < b.lo[0] + b.n[0] && j >= b.lo[1] && j < b.lo[1] + b.n[1] && k >= b.lo[2] &&
           k < b.lo[2] + b.n[2];
}
// the interior grown by one cell on every side: every cell that an interior
// or face-region stencil reaches
inline Box ext_box(const Geom &g) {
    return Box{{g.gc - 1, g.gc - 1, g.gc - 1},
               {g.nx - 2 * g.gc + 2, g.ny - 2 * g.gc + 2, g.nz - 2 * g.gc + 2}};
}
__device__ inline long long gtid() { return blockIdx.x * (long long)blockDim.x + threadIdx.x; }

struct F3 {
    double *c[3];
};
struct C3 {
    const double *c[3];
};

// ------------------------------------------------------------- kernels
// u* adjoint starts as the interior of the u^{n+1} adjoint
__global__ void k_adj_copy_interior(Box ib, Geom g, C3 src, F3 dst) {
    int i, j, k;
    if (!box_cell(ib, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    for (int m = 0; m < 3; m++) dst.c[m][c] = src.c[m][c];
}

struct FaceRange {
    int type[6];
    int start[7];  // flat start index of each face's cells
};

// orc_bc_velocity transposed (wind-tunnel face types)
__global__ void k_adj_ubc(Geom g, FaceRange fr, const double *xyz0, double dt, C3 unp1b, F3 usb, F3 unb) {
    const long long t = gtid();
    int f = 0;
    while (f < 6 && t >= fr.start[f + 1]) f++;
    if (f >= 6) return;
    const int type = fr.type[f];
    if (type != FALM_BC_CONVECTIVE && type != FALM_BC_SLIP && type != FALM_BC_WALL0) return;
    const int d = f / 2, ax = d == 0 ? 1 : 0, bx = d == 2 ? 1 : 2;
    const int shp[3] = {g.nx, g.ny, g.nz};
    const int gc = g.gc;
    const int na = shp[ax] - 2 * gc;
    const long long q = t - fr.start[f];
    const int ka = gc + (int)(q % na), kb = gc + (int)(q / na);
    int o1[3], o2[3], i1[3], i2[3];
    o1[ax] = o2[ax] = i1[ax] = i2[ax] = ka;
    o1[bx] = o2[bx] = i1[bx] = i2[bx] = kb;
    if (f & 1) {
        o1[d] = shp[d] - gc; o2[d] = shp[d] - gc + 1; i1[d] = shp[d] - gc - 1; i2[d] = shp[d] - gc - 2;
    } else {
        o1[d] = gc - 1; o2[d] = gc - 2; i1[d] = gc; i2[d] = gc + 1;
    }
    const int co1 = g.id(o1[0], o1[1], o1[2]), co2 = g.id(o2[0], o2[1], o2[2]);
    const int ci1 = g.id(i1[0], i1[1], i1[2]), ci2 = g.id(i2[0], i2[1], i2[2]);
    for (int m = 0; m < 3; m++) {
        const double w1 = unp1b.c[m][co1], w2 = unp1b.c[m][co2];
        if (type == FALM_BC_CONVECTIVE) {
            const double dxi = 1.0 / (xyz0[co1] - xyz0[ci1]);
            const double W = w1 + 2 * w2;  // u(o2) = 2 u(o1) - u*(i1)
            unsafeAtomicAdd(&usb.c[m][ci1], -w2);
            unsafeAtomicAdd(&unb.c[m][co1], (1.0 - dt * 0.5 * dxi * 3) * W);
            unsafeAtomicAdd(&unb.c[m][ci1], (dt * 0.5 * dxi * 4) * W);
            unsafeAtomicAdd(&unb.c[m][ci2], (-dt * 0.5 * dxi) * W);
        } else if (type == FALM_BC_SLIP) {
            unsafeAtomicAdd(&usb.c[m][ci1], m == d ? -w2 : w1 + w2);
        } else {
            unsafeAtomicAdd(&usb.c[m][ci1], -w2);
        }
    }
}

// orc_project_p transposed, gathered over the interior grown by one cell
// (xb): the forward's cell part on the interior (ib) reads p at c +- e_d
// with weight -+ dt/2 kx_d, its face part on the face region (fb) reads p at
// f + e_d and f with weight -+ dt/2 (g_d(f) + g_d(f + e_d)); p's adjoint at
// c sums those terms of the cells whose stencils reach c
__global__ void k_adj_project(Box xb, Box fb, Box ib, Geom g, double dt, C3 kx, C3 gg, C3 usb, C3 uunb, F3 uab,
                              F3 uusb, double *pb) {
    int i, j, k;
    if (!box_cell(xb, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const int st[3] = {1, g.nx, g.nxy};
    const bool cin = in_box(ib, i, j, k), cf = in_box(fb, i, j, k);
    double acc = 0.0;
    for (int d = 0; d < 3; d++) {
        if (cin) uab.c[d][c] += usb.c[d][c];
        if (cf) uusb.c[d][c] += uunb.c[d][c];
        const int cm = c - st[d], cp = c + st[d];
        const bool mi = in_box(ib, i - (d == 0), j - (d == 1), k - (d == 2));
        const bool pi = in_box(ib, i + (d == 0), j + (d == 1), k + (d == 2));
        const bool mf = in_box(fb, i - (d == 0), j - (d == 1), k - (d == 2));
        if (mi) acc -= dt * 0.5 * kx.c[d][cm] * usb.c[d][cm];
        if (pi) acc += dt * 0.5 * kx.c[d][cp] * usb.c[d][cp];
        if (mf) acc -= dt * 0.5 * (gg.c[d][cm] + gg.c[d][c]) * uunb.c[d][cm];
        if (cf) acc += dt * 0.5 * (gg.c[d][c] + gg.c[d][cp]) * uunb.c[d][c];
    }
    pb[c] += acc;
}

// orc_bc_pressure transposed: Neumann ghost = interior neighbour, ZERO = 0
__global__ void k_adj_pbc(Geom g, FaceRange fr, double *pb) {
    const long long t = gtid();
    int f = 0;
    while (f < 6 && t >= fr.start[f + 1]) f++;
    if (f >= 6) return;
    const int type = fr.type[f];
    if (type == FALM_PBC_NONE) return;
    const int d = f / 2, ax = d == 0 ? 1 : 0, bx = d == 2 ? 1 : 2;
    const int shp[3] = {g.nx, g.ny, g.nz};
    const int gc = g.gc;
    const int na = shp[ax] - 2 * gc;
    const long long q = t - fr.start[f];
    const int ka = gc + (int)(q % na), kb = gc + (int)(q / na);
    int o1[3], i1[3];
    o1[ax] = i1[ax] = ka;
    o1[bx] = i1[bx] = kb;
    o1[d] = (f & 1) ? shp[d] - gc : gc - 1;
    i1[d] = (f & 1) ? shp[d] - gc - 1 : gc;
    const int co = g.id(o1[0], o1[1], o1[2]), ci = g.id(i1[0], i1[1], i1[2]);
    if (type == FALM_PBC_NEUMANN) unsafeAtomicAdd(&pb[ci], pb[co]);
    pb[co] = 0.0;
}

// orc_poisson_rhs transposed, gathered over the face region (fb): the face
// flux uu_d(f) enters the right-hand side of f and of f + e_d (interior
// cells, ib) with weights +-scale/ja
__global__ void k_adj_rhs(Box fb, Box ib, Geom g, const double *ja, double scale, const double *rhsb, F3 uusb) {
    int i, j, k;
    if (!box_cell(fb, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const int st[3] = {1, g.nx, g.nxy};
    const bool cin = in_box(ib, i, j, k);
    const double wc = cin ? rhsb[c] * scale / ja[c] : 0.0;
    for (int d = 0; d < 3; d++) {
        const int cp = c + st[d];
        const bool pi = in_box(ib, i + (d == 0), j + (d == 1), k + (d == 2));
        if (cin || pi) uusb.c[d][c] += wc - (pi ? rhsb[cp] * scale / ja[cp] : 0.0);
    }
}

// orc_u_to_uu transposed, gathered over the interior grown by one cell
// (xb): u_d(c) enters the face fluxes uu_d of the face-region (fb) faces c
// and c - e_d, both with weight ja kx_d / 2 at c
__global__ void k_adj_u_to_uu(Box xb, Box fb, Geom g, C3 kx, const double *ja, C3 uusb, F3 uab) {
    int i, j, k;
    if (!box_cell(xb, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const int st[3] = {1, g.nx, g.nxy};
    const bool cf = in_box(fb, i, j, k);
    for (int d = 0; d < 3; d++) {
        const bool mf = in_box(fb, i - (d == 0), j - (d == 1), k - (d == 2));
        if (!cf && !mf) continue;
        const double w = (cf ? uusb.c[d][c] : 0.0) + (mf ? uusb.c[d][c - st[d]] : 0.0);
        uab.c[d][c] += 0.5 * ja[c] * kx.c[d][c] * w;
    }
}

// ua's ghost cells are u^n's (the swap keeps them): move their adjoint
__global__ void k_adj_fold_ghosts(Geom g, size_t n, F3 uab, F3 unb) {
    const long long t = gtid();
    if (t >= (long long)n) return;
    const int c = (int)t;
    const int k = c / g.nxy, r = c - k * g.nxy, j = r / g.nx, i = r - j * g.nx;
    const bool inner = i >= g.gc && i < g.nx - g.gc && j >= g.gc && j < g.ny - g.gc && k >= g.gc && k < g.nz - g.gc;
    if (inner) return;
    for (int m = 0; m < 3; m++) {
        unb.c[m][c] += uab.c[m][c];
        uab.c[m][c] = 0.0;
    }
}

// orc_pseudo_u transposed (Upwind3 family, un == u)
__global__ void k_adj_pseudo_u(Box ib, Geom g, double ALPHA, double ReI, double dt, C3 u, C3 uu,
                               const double *nut, C3 kx, C3 gg, const double *ja, C3 uab, F3 unb, F3 uunb,
                               double *nutb, F3 ffb) {
    int i, j, k;
    if (!box_cell(ib, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const int st[3] = {1, g.nx, g.nxy};
    const double lco[5] = {0.0, 1. / 24., -27. / 24., 27. / 24., -1. / 24.};
    const double rco[5] = {1. / 24., -27. / 24., 27. / 24., -1. / 24., 0.0};
    const double dco[5] = {1., -4., 6., -4., 1.};
    const double J = ja[c];
    for (int d = 0; d < 3; d++) {
        const double w = uab.c[d][c];
        ffb.c[d][c] = dt * w;
        if (w == 0.0) continue;
        unsafeAtomicAdd(&unb.c[d][c], w);
        const double cw = -dt * w;
        for (int e = 0; e < 3; e++) {
            const int s = st[e];
            const double U = u.c[e][c] * kx.c[e][c], Uabs = fabs(U);
            const double JP = uu.c[e][c], JM = uu.c[e][c - s];
            double v[5];
            for (int m = 0; m < 5; m++) v[m] = u.c[d][c + (m - 2) * s];
            const double L = (-v[4] + 27 * v[3] - 27 * v[2] + v[1]) / 24.;
            const double R = (-v[3] + 27 * v[2] - 27 * v[1] + v[0]) / 24.;
            const double D4 = v[4] - 4 * v[3] + 6 * v[2] - 4 * v[1] + v[0];
            for (int m = 0; m < 5; m++)
                unsafeAtomicAdd(&unb.c[d][c + (m - 2) * s],
                                cw * (0.5 * (JP * lco[m] + JM * rco[m]) / J + Uabs * ALPHA * dco[m]));
            unsafeAtomicAdd(&uunb.c[e][c], cw * 0.5 * L / J);
            unsafeAtomicAdd(&uunb.c[e][c - s], cw * 0.5 * R / J);
            const double sg = (double)((U > 0) - (U < 0));
            unsafeAtomicAdd(&unb.c[e][c], cw * ALPHA * D4 * sg * kx.c[e][c]);
        }
        const double cv = dt * w;
        for (int e = 0; e < 3; e++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const int f = c + sgn * st[e];
                const double a = (gg.c[e][c] + gg.c[e][f]) / (2 * J);
                const double b = ReI + 0.5 * (nut[c] + nut[f]);
                const double diff = u.c[d][f] - u.c[d][c];
                unsafeAtomicAdd(&unb.c[d][f], cv * a * b);
                unsafeAtomicAdd(&unb.c[d][c], -cv * a * b);
                unsafeAtomicAdd(&nutb[c], cv * a * 0.5 * diff);
                unsafeAtomicAdd(&nutb[f], cv * a * 0.5 * diff);
            }
    }
}

// orc_sgs transposed
__global__ void k_adj_sgs(Box ib, Geom g, double Cs, C3 u, C3 kx, const double *ja, const double *nutb, F3 unb) {
    int i, j, k;
    if (!box_cell(ib, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const double w = nutb[c];
    if (w == 0.0) return;
    const int st[3] = {1, g.nx, g.nxy};
    double gr[3][3];
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) gr[d][e] = 0.5 * kx.c[e][c] * (u.c[d][c + st[e]] - u.c[d][c - st[e]]);
    const double dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
    const double dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
    const double dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
    const double S = 2 * dux * dux + 2 * dvy * dvy + 2 * dwz * dwz + (duy + dvx) * (duy + dvx) +
                     (dvz + dwy) * (dvz + dwy) + (duz + dwx) * (duz + dwx);
    const double Du = sqrt(S);
    if (!(Du > 0.0)) return;  // sqrt at 0: zero subgradient
    const double lc = Cs * cbrt(ja[c]);
    const double coef = w * lc * lc / (2.0 * Du);
    double dS[3][3];
    dS[0][0] = 4 * dux; dS[1][1] = 4 * dvy; dS[2][2] = 4 * dwz;
    dS[0][1] = dS[1][0] = 2 * (duy + dvx);
    dS[1][2] = dS[2][1] = 2 * (dvz + dwy);
    dS[0][2] = dS[2][0] = 2 * (duz + dwx);
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) {
            const double gb = coef * dS[d][e] * 0.5 * kx.c[e][c];
            unsafeAtomicAdd(&unb.c[d][c + st[e]], gb);
            unsafeAtomicAdd(&unb.c[d][c - st[e]], -gb);
        }
}

// the transposed Poisson operator in the same [c,w,e,s,n,b,t] layout
__global__ void k_transpose7(Box ib, Geom g, Mat7<double> a, double *at, size_t n) {
    int i, j, k;
    if (!box_cell(ib, gtid(), i, j, k)) return;
    const int c = g.id(i, j, k);
    const int st[3] = {1, g.nx, g.nxy};
    const int ijk[3] = {i, j, k};
    const int lim[3] = {g.nx - g.gc, g.ny - g.gc, g.nz - g.gc};
    at[c] = a.c[0][c];
    for (int d = 0; d < 3; d++) {
        const bool hasm = ijk[d] - 1 >= g.gc, hasp = ijk[d] + 1 < lim[d];
        at[c + (size_t)(1 + 2 * d) * n] = hasm ? a.c[2 + 2 * d][c - st[d]] : 0.0;
        at[c + (size_t)(2 + 2 * d) * n] = hasp ? a.c[1 + 2 * d][c + st[d]] : 0.0;
    }
}

// ------------------------------------------------------------- driver
struct DevSnap {
    double *u = nullptr, *uu = nullptr, *p = nullptr, *nut = nullptr;
    int64_t it = 0;
};

struct AdjWork {
    std::vector<void *> bufs;
    size_t bytes = 0;
    ~AdjWork() {
        for (void *p : bufs) (void)hipFree(p);
    }
    falm_status alloc(double **p, size_t count) {
        FALM_HIP(hipMalloc((void **)p, count * sizeof(double)));
        bufs.push_back(*p);
        bytes += count * sizeof(double);
        return FALM_OK;
    }
};

static falm_status snap_alloc(AdjWork &W, const falm_case *C, DevSnap &S) {
    FALM_TRY(W.alloc(&S.u, 3 * C->n));
    FALM_TRY(W.alloc(&S.uu, 3 * C->n));
    FALM_TRY(W.alloc(&S.p, C->n));
    FALM_TRY(W.alloc(&S.nut, C->n));
    return FALM_OK;
}
static falm_status snap_take(const falm_case *C, DevSnap &S) {
    hipStream_t s = C->stream;
    FALM_HIP(hipMemcpyAsync(S.u, C->u, 3 * C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(S.uu, C->uu, 3 * C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(S.p, C->p, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(S.nut, C->nut, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    S.it = C->it;
    return FALM_OK;
}
static falm_status snap_restore(falm_case *C, const DevSnap &S) {
    hipStream_t s = C->stream;
    FALM_HIP(hipMemcpyAsync(C->u, S.u, 3 * C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(C->uu, S.uu, 3 * C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(C->p, S.p, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FALM_HIP(hipMemcpyAsync(C->nut, S.nut, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    C->it = S.it;
    return FALM_OK;
}

static falm_status check_adjoint_case(const falm_case *C) {
    const falm_case_config &c = C->cfg;
    if (!C->alm) {
        set_error("falm_case_adjoint: the case has no rotor");
        return FALM_ERR_INVALID;
    }
    if (c.kind != FALM_CASE_WINDTUNNEL || C->prec != FALM_F64) {
        set_error("falm_case_adjoint: wind-tunnel cases in FP64 only");
        return FALM_ERR_UNSUPPORTED;
    }
    if (c.adv_scheme != FALM_ADV_UPWIND3 && c.adv_scheme != FALM_ADV_UTOPIA && c.adv_scheme != FALM_ADV_KK) {
        set_error("falm_case_adjoint: Upwind3-family advection only");
        return FALM_ERR_UNSUPPORTED;
    }
    if (c.sgs_model != FALM_SGS_SMAGORINSKY && c.sgs_model != FALM_SGS_EMPTY) {
        set_error("falm_case_adjoint: Smagorinsky or no SGS model only");
        return FALM_ERR_UNSUPPORTED;
    }
    return FALM_OK;
}

static FaceRange velocity_faces(const falm_case *C, const Geom &g, long long &total) {
    FaceRange fr;
    const int shp[3] = {g.nx, g.ny, g.nz};
    fr.start[0] = 0;
    for (int f = 0; f < 6; f++) {
        const int d = f / 2, ax = d == 0 ? 1 : 0, bx = d == 2 ? 1 : 2;
        fr.type[f] = C->ubc[f].type;
        fr.start[f + 1] = fr.start[f] + (shp[ax] - 2 * g.gc) * (shp[bx] - 2 * g.gc);
    }
    total = fr.start[6];
    return fr;
}
static FaceRange pressure_faces(const falm_case *C, const Geom &g, long long &total) {
    FaceRange fr = velocity_faces(C, g, total);
    for (int f = 0; f < 6; f++) fr.type[f] = C->pbc[f];
    return fr;
}

static C3 c3(const void *p, size_t n) {
    C3 v;
    for (int m = 0; m < 3; m++) v.c[m] = (const double *)p + m * n;
    return v;
}
static F3 f3(void *p, size_t n) {
    F3 v;
    for (int m = 0; m < 3; m++) v.c[m] = (double *)p + m * n;
    return v;
}

}  // namespace falm

using namespace falm;

namespace {

falm_status vjp_check(const falm_domain *dom, std::initializer_list<const void *> ptrs) {
    if (!dom) return FALM_ERR_INVALID;
    for (const void *p : ptrs)
        if (!p) return FALM_ERR_INVALID;
    return check_domain(dom);
}

}  // namespace

// The VJP kernels accumulate into their outputs with hardware FP64 atomic
// adds, which are exact only on coarse-grained device memory (hipMalloc):
// on host-pinned, managed or fine-grained memory they can drop updates
// silently.  Host and managed memory are refused here (fine-grained device
// memory from hipExtMallocWithFlags cannot be told apart by its attributes).
falm_status falm::vjp_outputs_on_device(std::initializer_list<const void *> outs) {
    for (const void *p : outs) {
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, p) != hipSuccess) {
            (void)hipGetLastError();
            set_error("VJP output %p is not device memory", p);
            return FALM_ERR_INVALID;
        }
        if (at.type != hipMemoryTypeDevice || at.isManaged) {
            set_error("VJP output %p must be device memory from hipMalloc (the adjoint accumulates with FP64 "
                      "atomics)", p);
            return FALM_ERR_INVALID;
        }
    }
    return FALM_OK;
}

namespace {

FaceRange face_range(const Geom &g, const int type[6], long long &total) {
    FaceRange fr;
    const int shp[3] = {g.nx, g.ny, g.nz};
    fr.start[0] = 0;
    for (int f = 0; f < 6; f++) {
        const int d = f / 2, ax = d == 0 ? 1 : 0, bx = d == 2 ? 1 : 2;
        fr.type[f] = type[f];
        fr.start[f + 1] = fr.start[f] + (shp[ax] - 2 * g.gc) * (shp[bx] - 2 * g.gc);
    }
    total = fr.start[6];
    return fr;
}

}  // namespace

extern "C" {

// ------------------------------------------------ per-stage VJPs (FP64)
falm_status falm_pseudo_u_vjp(const falm_domain *dom, int adv_scheme, double ReI, double dt, const void *u,
                              const void *uu, const void *nut, const void *kx, const void *g, const void *ja,
                              const void *uabar, void *ubar, void *uubar, void *nutbar, void *ffbar, void *stream) {
    FALM_TRY(vjp_check(dom, {u, uu, nut, kx, g, ja, uabar, ubar, uubar, nutbar, ffbar}));
    FALM_TRY(vjp_outputs_on_device({ubar, uubar, nutbar, ffbar}));
    if (adv_scheme != FALM_ADV_UPWIND3 && adv_scheme != FALM_ADV_UTOPIA && adv_scheme != FALM_ADV_KK) {
        set_error("falm_pseudo_u_vjp: Upwind3-family advection only");
        return FALM_ERR_UNSUPPORTED;
    }
    const double ALPHA = adv_scheme == FALM_ADV_UPWIND3 ? 1. / 24. : adv_scheme == FALM_ADV_UTOPIA ? 1. / 12. : 1. / 4.;
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box ib = interior_box(gm);
    hipLaunchKernelGGL(k_adj_pseudo_u, box_grid(ib), dim3(EW_BLOCK), 0, (hipStream_t)stream, ib, gm, ALPHA, ReI, dt,
                       c3(u, n), c3(uu, n), (const double *)nut, c3(kx, n), c3(g, n), (const double *)ja,
                       c3(uabar, n), f3(ubar, n), f3(uubar, n), (double *)nutbar, f3(ffbar, n));
    FALM_CHECK_LAUNCH("k_adj_pseudo_u");
    return FALM_OK;
}

falm_status falm_u_to_uu_vjp(const falm_domain *dom, const void *kx, const void *ja, const void *uubar, void *ubar,
                             void *stream) {
    FALM_TRY(vjp_check(dom, {kx, ja, uubar, ubar}));
    FALM_TRY(vjp_outputs_on_device({ubar}));
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box fb = face_box(gm), xb = ext_box(gm);
    hipLaunchKernelGGL(k_adj_u_to_uu, box_grid(xb), dim3(EW_BLOCK), 0, (hipStream_t)stream, xb, fb, gm, c3(kx, n),
                       (const double *)ja, c3(uubar, n), f3(ubar, n));
    FALM_CHECK_LAUNCH("k_adj_u_to_uu");
    return FALM_OK;
}

falm_status falm_poisson_rhs_vjp(const falm_domain *dom, const void *ja, double dt, double maxdiag,
                                 const void *rhsbar, void *uubar, void *stream) {
    FALM_TRY(vjp_check(dom, {ja, rhsbar, uubar}));
    FALM_TRY(vjp_outputs_on_device({uubar}));
    if (!(dt * maxdiag != 0.0)) return FALM_ERR_INVALID;
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box ib = interior_box(gm), fb = face_box(gm);
    hipLaunchKernelGGL(k_adj_rhs, box_grid(fb), dim3(EW_BLOCK), 0, (hipStream_t)stream, fb, ib, gm,
                       (const double *)ja, 1.0 / (dt * maxdiag), (const double *)rhsbar, f3(uubar, n));
    FALM_CHECK_LAUNCH("k_adj_rhs");
    return FALM_OK;
}

falm_status falm_project_p_vjp(const falm_domain *dom, double dt, const void *kx, const void *g, const void *ubar,
                               const void *uubar, void *uabar, void *uuabar, void *pbar, void *stream) {
    FALM_TRY(vjp_check(dom, {kx, g, ubar, uubar, uabar, uuabar, pbar}));
    FALM_TRY(vjp_outputs_on_device({uabar, uuabar, pbar}));
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box fb = face_box(gm), ib = interior_box(gm), xb = ext_box(gm);
    hipLaunchKernelGGL(k_adj_project, box_grid(xb), dim3(EW_BLOCK), 0, (hipStream_t)stream, xb, fb, ib, gm, dt,
                       c3(kx, n), c3(g, n), c3(ubar, n), c3(uubar, n), f3(uabar, n), f3(uuabar, n),
                       (double *)pbar);
    FALM_CHECK_LAUNCH("k_adj_project");
    return FALM_OK;
}

falm_status falm_sgs_vjp(const falm_domain *dom, double Cs, const void *u, const void *kx, const void *ja,
                         const void *nutbar, void *ubar, void *stream) {
    FALM_TRY(vjp_check(dom, {u, kx, ja, nutbar, ubar}));
    FALM_TRY(vjp_outputs_on_device({ubar}));
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box ib = interior_box(gm);
    hipLaunchKernelGGL(k_adj_sgs, box_grid(ib), dim3(EW_BLOCK), 0, (hipStream_t)stream, ib, gm, Cs, c3(u, n),
                       c3(kx, n), (const double *)ja, (const double *)nutbar, f3(ubar, n));
    FALM_CHECK_LAUNCH("k_adj_sgs");
    return FALM_OK;
}

falm_status falm_bc_velocity_vjp(const falm_domain *dom, const falm_face_bc faces[6], const void *xyz, double dt,
                                 const void *unp1bar, void *usbar, void *unbar, void *stream) {
    FALM_TRY(vjp_check(dom, {faces, xyz, unp1bar, usbar, unbar}));
    FALM_TRY(vjp_outputs_on_device({usbar, unbar}));
    const Geom gm = make_geom(dom);
    const size_t n = ncell_of(dom);
    const Box ib = interior_box(gm);
    int type[6];
    for (int f = 0; f < 6; f++) type[f] = faces[f].type;
    long long nf = 0;
    const FaceRange fr = face_range(gm, type, nf);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_adj_copy_interior, box_grid(ib), dim3(EW_BLOCK), 0, s, ib, gm, c3(unp1bar, n),
                       f3(usbar, n));
    if (nf > 0)
        hipLaunchKernelGGL(k_adj_ubc, ew_grid((size_t)nf), dim3(EW_BLOCK), 0, s, gm, fr, (const double *)xyz, dt,
                           c3(unp1bar, n), f3(usbar, n), f3(unbar, n));
    FALM_CHECK_LAUNCH("k_adj_ubc");
    return FALM_OK;
}

falm_status falm_bc_pressure_vjp(const falm_domain *dom, const int faces[6], void *pbar, void *stream) {
    FALM_TRY(vjp_check(dom, {faces, pbar}));
    FALM_TRY(vjp_outputs_on_device({pbar}));
    const Geom gm = make_geom(dom);
    long long nf = 0;
    const FaceRange fr = face_range(gm, faces, nf);
    if (nf > 0)
        hipLaunchKernelGGL(k_adj_pbc, ew_grid((size_t)nf), dim3(EW_BLOCK), 0, (hipStream_t)stream, gm, fr,
                           (double *)pbar);
    FALM_CHECK_LAUNCH("k_adj_pbc");
    return FALM_OK;
}

falm_status falm_case_window(falm_case *C, int nsteps, double *J) {
    if (!C || !J || nsteps < 1) return FALM_ERR_INVALID;
    if (!C->alm) {
        set_error("falm_case_window: the case has no rotor");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = C->stream;
    AdjWork W;
    DevSnap S0;
    double *acc = nullptr;
    const int prec_ok = (C->prec == FALM_F64);
    if (prec_ok) {
        FALM_TRY(snap_alloc(W, C, S0));
        FALM_TRY(snap_take(C, S0));
    } else {
        // float fields: snapshot as raw bytes
        FALM_TRY(W.alloc(&S0.u, (3 * C->n * C->tb + 7) / 8));
        FALM_TRY(W.alloc(&S0.uu, (3 * C->n * C->tb + 7) / 8));
        FALM_TRY(W.alloc(&S0.p, C->n));
        FALM_TRY(W.alloc(&S0.nut, (C->n * C->tb + 7) / 8));
        FALM_HIP(hipMemcpyAsync(S0.u, C->u, 3 * C->n * C->tb, hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(S0.uu, C->uu, 3 * C->n * C->tb, hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(S0.p, C->p, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(S0.nut, C->nut, C->n * C->tb, hipMemcpyDeviceToDevice, s));
        S0.it = C->it;
    }
    FALM_TRY(W.alloc(&acc, 1));
    FALM_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
    falm_solve_info si{0, 0.0, 0, FALM_OK};
    for (int k = 0; k < nsteps; k++) FALM_TRY(case_step_power(C, &si, acc));
    double h = 0.0;
    FALM_HIP(hipMemcpyAsync(&h, acc, sizeof(double), hipMemcpyDeviceToHost, s));
    if (prec_ok) {
        FALM_TRY(snap_restore(C, S0));
    } else {
        FALM_HIP(hipMemcpyAsync(C->u, S0.u, 3 * C->n * C->tb, hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(C->uu, S0.uu, 3 * C->n * C->tb, hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(C->p, S0.p, C->n * sizeof(double), hipMemcpyDeviceToDevice, s));
        FALM_HIP(hipMemcpyAsync(C->nut, S0.nut, C->n * C->tb, hipMemcpyDeviceToDevice, s));
        C->it = S0.it;
    }
    FALM_HIP(hipStreamSynchronize(s));
    *J = h / nsteps;
    return FALM_OK;
}

falm_status falm_case_adjoint(falm_case *C, int nsteps, int ckpt, double adj_rtol, double *J, double *g_pitch,
                              double *g_twist, size_t *bytes) {
    if (!C || !J || !g_pitch || !g_twist || nsteps < 1 || ckpt < 1 || !(adj_rtol > 0.0)) return FALM_ERR_INVALID;
    FALM_TRY(check_adjoint_case(C));
    const falm_case_config &c = C->cfg;
    hipStream_t s = C->stream;
    const size_t n = C->n;
    const Geom g = make_geom(&C->dom);
    const Box ib = interior_box(g), fbx = face_box(g), xbx = ext_box(g);
    const double ALPHA = c.adv_scheme == FALM_ADV_UPWIND3 ? 1. / 24. : c.adv_scheme == FALM_ADV_UTOPIA ? 1. / 12. : 1. / 4.;
    const int nt = alm_n_turbine(C->alm), napb = alm_n_ap_per_blade(C->alm);
    AdjWork W;
    // forward with checkpoints
    DevSnap S0;
    FALM_TRY(snap_alloc(W, C, S0));
    FALM_TRY(snap_take(C, S0));
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    std::vector<DevSnap> cps(nseg), states(ckpt);
    for (auto &x : cps) FALM_TRY(snap_alloc(W, C, x));
    for (auto &x : states) FALM_TRY(snap_alloc(W, C, x));
    double *acc = nullptr;
    FALM_TRY(W.alloc(&acc, 1));
    FALM_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
    falm_solve_info si{0, 0.0, 0, FALM_OK};
    for (int st = 0; st < nsteps; st++) {
        if (st % ckpt == 0) FALM_TRY(snap_take(C, cps[st / ckpt]));
        FALM_TRY(case_step_power(C, &si, acc));
    }
    double hJ = 0.0;
    FALM_HIP(hipMemcpyAsync(&hJ, acc, sizeof(double), hipMemcpyDeviceToHost, s));
    // adjoint work fields
    double *ub, *uub, *unb, *uunb, *usb, *uab, *uusb, *pb, *lam, *nutb, *ffb, *nutn, *at, *gp, *gt, *ga;
    FALM_TRY(W.alloc(&ub, 3 * n));
    FALM_TRY(W.alloc(&uub, 3 * n));
    FALM_TRY(W.alloc(&unb, 3 * n));
    FALM_TRY(W.alloc(&uunb, 3 * n));
    FALM_TRY(W.alloc(&usb, 3 * n));
    FALM_TRY(W.alloc(&uab, 3 * n));
    FALM_TRY(W.alloc(&uusb, 3 * n));
    FALM_TRY(W.alloc(&pb, n));
    FALM_TRY(W.alloc(&lam, n));
    FALM_TRY(W.alloc(&nutb, n));
    FALM_TRY(W.alloc(&ffb, 3 * n));
    FALM_TRY(W.alloc(&nutn, n));
    FALM_TRY(W.alloc(&at, 7 * n));
    FALM_TRY(W.alloc(&gp, (size_t)nt * 8));
    FALM_TRY(W.alloc(&gt, (size_t)napb));
    FALM_TRY(W.alloc(&ga, 2 * (size_t)nt));
    FALM_HIP(hipMemsetAsync(ub, 0, 3 * n * sizeof(double), s));
    FALM_HIP(hipMemsetAsync(uub, 0, 3 * n * sizeof(double), s));
    FALM_HIP(hipMemsetAsync(at, 0, 7 * n * sizeof(double), s));
    FALM_HIP(hipMemsetAsync(gp, 0, (size_t)nt * 8 * sizeof(double), s));
    FALM_HIP(hipMemsetAsync(gt, 0, (size_t)napb * sizeof(double), s));
    FALM_HIP(hipMemsetAsync(ga, 0, 2 * (size_t)nt * sizeof(double), s));
    hipLaunchKernelGGL(k_transpose7, box_grid(ib), dim3(EW_BLOCK), 0, s, ib, g, mat7<double>(C->a, n), at, n);
    FALM_CHECK_LAUNCH("k_transpose7");
    // the transposed solve: BiCGStab with the case's preconditioner, double
    falm_solver_config scfg = c.solver;
    scfg.type = FALM_LS_PBICGSTAB;
    scfg.precision = FALM_SOLVE_F64;
    scfg.maxit = 100000;
    falm_solver *ST = nullptr;
    FALM_TRY(falm_solver_create(&C->dom, &scfg, &ST));
    struct Guard {
        falm_solver *S;
        ~Guard() { falm_solver_destroy(S); }
    } guard{ST};
    FALM_TRY(falm_solver_set_matrix(ST, at, s));
    long long nvf = 0, npf = 0;
    const FaceRange vfr = velocity_faces(C, g, nvf), pfr = pressure_faces(C, g, npf);
    const double nint = (double)ib.size();
    const double wP = 1.0 / nsteps;
    const C3 kx = c3(C->kx, n), gg = c3(C->g, n);
    const double *ja = (const double *)C->ja;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * ckpt, s1 = (s0 + ckpt < nsteps) ? s0 + ckpt : nsteps;
        FALM_TRY(snap_restore(C, cps[sg]));
        for (int st = s0; st < s1; st++) {
            FALM_TRY(snap_take(C, states[st - s0]));
            if (st + 1 < s1) FALM_TRY(case_step_power(C, &si, nullptr));
        }
        for (int st = s1 - 1; st >= s0; st--) {
            const DevSnap &Sn = states[st - s0];
            const double t = c.dt * (double)(Sn.it + 1);
            // nut^n = SGS(u^n)
            FALM_HIP(hipMemsetAsync(nutn, 0, n * sizeof(double), s));
            if (c.sgs_model == FALM_SGS_SMAGORINSKY) FALM_TRY(sep_sgs(&C->dom, FALM_F64, c.Cs, Sn.u, nutn, C->sm, s));
            for (double *z : {unb, uunb, usb, uab, uusb})
                FALM_HIP(hipMemsetAsync(z, 0, 3 * n * sizeof(double), s));
            FALM_HIP(hipMemsetAsync(pb, 0, n * sizeof(double), s));
            FALM_HIP(hipMemsetAsync(nutb, 0, n * sizeof(double), s));
            // boundary fill, projection, pressure boundary
            hipLaunchKernelGGL(k_adj_copy_interior, box_grid(ib), dim3(EW_BLOCK), 0, s, ib, g, c3(ub, n), f3(usb, n));
            hipLaunchKernelGGL(k_adj_ubc, ew_grid((size_t)nvf), dim3(EW_BLOCK), 0, s, g, vfr, (const double *)C->xyz,
                               c.dt, c3(ub, n), f3(usb, n), f3(unb, n));
            hipLaunchKernelGGL(k_adj_project, box_grid(xbx), dim3(EW_BLOCK), 0, s, xbx, fbx, ib, g, c.dt, kx, gg,
                               c3(usb, n), c3(uub, n), f3(uab, n), f3(uusb, n), pb);
            hipLaunchKernelGGL(k_adj_pbc, ew_grid((size_t)npf), dim3(EW_BLOCK), 0, s, g, pfr, pb);
            FALM_CHECK_LAUNCH("k_adj_pbc");
            // rhs_bar = A^-T p_bar on the interior
            double nrm2 = 0.0;
            FALM_TRY(falm_norm2(&C->dom, FALM_F64, pb, 1, &nrm2, s));
            FALM_HIP(hipMemsetAsync(lam, 0, n * sizeof(double), s));
            if (nrm2 > 0.0) {
                falm_solver_config tcfg = scfg;
                tcfg.tol = adj_rtol * sqrt(nrm2) / nint;
                FALM_TRY(solver_set_tol(ST, tcfg.tol));
                falm_solve_info ti;
                FALM_TRY(falm_solver_solve(ST, at, lam, pb, nullptr, s, &ti));
            }
            // rhs, face flux, ghost cells of ua, pseudo-velocity
            hipLaunchKernelGGL(k_adj_rhs, box_grid(fbx), dim3(EW_BLOCK), 0, s, fbx, ib, g, ja,
                               1.0 / (c.dt * C->maxdiag), (const double *)lam, f3(uusb, n));
            hipLaunchKernelGGL(k_adj_u_to_uu, box_grid(xbx), dim3(EW_BLOCK), 0, s, xbx, fbx, g, kx, ja, c3(uusb, n),
                               f3(uab, n));
            hipLaunchKernelGGL(k_adj_fold_ghosts, ew_grid(n), dim3(EW_BLOCK), 0, s, g, n, f3(uab, n), f3(unb, n));
            hipLaunchKernelGGL(k_adj_pseudo_u, box_grid(ib), dim3(EW_BLOCK), 0, s, ib, g, ALPHA, 1.0 / c.Re, c.dt,
                               c3(Sn.u, n), c3(Sn.uu, n), (const double *)nutn, kx, gg, ja, c3(uab, n), f3(unb, n),
                               f3(uunb, n), nutb, f3(ffb, n));
            FALM_CHECK_LAUNCH("k_adj_pseudo_u");
            // rotor: this step's power and force
            FALM_TRY(alm_adjoint(C->alm, Sn.u, t, ffb, wP, unb, gp, gt, ga, s));
            // nut^n = SGS(u^n)
            if (c.sgs_model == FALM_SGS_SMAGORINSKY) {
                hipLaunchKernelGGL(k_adj_sgs, box_grid(ib), dim3(EW_BLOCK), 0, s, ib, g, c.Cs, c3(Sn.u, n), kx, ja,
                                   (const double *)nutb, f3(unb, n));
                FALM_CHECK_LAUNCH("k_adj_sgs");
            }
            std::swap(ub, unb);
            std::swap(uub, uunb);
        }
    }
    std::vector<double> hgp((size_t)nt * 8), hgt((size_t)napb), hga(2 * (size_t)nt);
    FALM_HIP(hipMemcpyAsync(hgp.data(), gp, hgp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipMemcpyAsync(hgt.data(), gt, hgt.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipMemcpyAsync(hga.data(), ga, hga.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_TRY(snap_restore(C, S0));
    FALM_HIP(hipStreamSynchronize(s));
    *J = hJ / nsteps;
    // the attitude gradient, per degree
    C->g_att.assign(hga.size(), 0.0);
    for (size_t q = 0; q < hga.size(); q++) C->g_att[q] = hga[q] * (3.14159265358979323846 / 180.0);
    for (size_t q = 0; q < hgp.size(); q++) g_pitch[q] = hgp[q];
    for (size_t q = 0; q < hgt.size(); q++) g_twist[q] = hgt[q];
    if (bytes) *bytes = W.bytes + 7 * n * sizeof(double) * 2;  // + the transposed solver's operator copies
    return FALM_OK;
}

falm_status falm_case_adjoint_attitude(falm_case *C, double *g_yaw, double *g_offset) {
    if (!C || !g_yaw || !g_offset) return FALM_ERR_INVALID;
    if (C->g_att.empty()) {
        set_error("falm_case_adjoint_attitude: no falm_case_adjoint call on this case yet");
        return FALM_ERR_SETUP;
    }
    const size_t nt = C->g_att.size() / 2;
    for (size_t t = 0; t < nt; t++) {
        g_yaw[t] = C->g_att[2 * t];
        g_offset[t] = C->g_att[2 * t + 1];
    }
    return FALM_OK;
}

}  // extern "C"
