// cfd.hip — fractional-step operators of FalmCFD on gfx950.
//
// pseudo-velocity (advection + diffusion + body force), contravariant face
// flux, Poisson right-hand side, pressure projection, Smagorinsky eddy
// viscosity and divergence.  All are HBM-bound stencils: one thread per cell
// over 64x4 tiles with an XCD-contiguous tile order; neighbour reuse comes
// from L1/L2 (the x-neighbours of a wavefront are in the same cache lines,
// the y/z-neighbours are re-read by the adjacent rows/planes of the same
// XCD slab while they are still in its L2).
#include "common.hpp"

namespace falm {

// ------------------------------------------------------------ schemes
// src/FalmCFDScheme.h:8-10
template <class T>
__device__ inline T upwind_flux1(T ul, T ur, T UF) {
    return T(0.5) * (UF * (ur + ul) - fabs(UF) * (ur - ul));
}

// One advection term of the Upwind3 family (Upwind3rd, UTOPIA, KK differ
// only in the fourth-difference weight ALPHA; src/FalmCFDScheme.h:68-165).
template <class T>
__device__ inline T u3_dir(T ALPHA, T ucc, T up1, T up2, T um1, T um2, T Uabs, T JP, T JM, T jac) {
    T UDL = JP * (-up2 + T(27) * up1 - T(27) * ucc + um1) / T(24.);
    T UDR = JM * (-up1 + T(27) * ucc - T(27) * um1 + um2) / T(24.);
    T UD4 = Uabs * (up2 - T(4) * up1 + T(6) * ucc - T(4) * um1 + um2) * ALPHA;
    return T(.5) * (UDL + UDR) / jac + UD4;
}

// QUICK, src/FalmCFDScheme.h:33-66 (one direction pair, before /16J)
template <class T>
__device__ inline T quick_dir(T ucc, T up1, T up2, T um1, T um2, T JP, T JM) {
    T adv = T(0);
    adv += JP * (-up2 + T(9) * up1 + T(9) * ucc - um1);
    adv += fabs(JP) * (up2 - T(3) * up1 + T(3) * ucc - um1);
    adv -= JM * (-up1 + T(9) * ucc + T(9) * um1 - um2);
    adv -= fabs(JM) * (up1 - T(3) * ucc + T(3) * um1 - um2);
    return adv;
}

template <int SCHEME, class T>
__device__ inline T advection(T ucc, T ue1, T ue2, T uw1, T uw2, T un1, T un2, T us1, T us2, T ut1,
                              T ut2, T ub1, T ub2, T Uabs, T Vabs, T Wabs, T UE, T UW, T VN, T VS,
                              T WT, T WB, T jac) {
    if (SCHEME == FALM_ADV_UPWIND3 || SCHEME == FALM_ADV_UTOPIA || SCHEME == FALM_ADV_KK) {
        const T ALPHA = SCHEME == FALM_ADV_UPWIND3 ? T(1. / 24.) : SCHEME == FALM_ADV_UTOPIA ? T(1. / 12.) : T(1. / 4.);
        T adv = T(0.0);
        adv += u3_dir<T>(ALPHA, ucc, ue1, ue2, uw1, uw2, Uabs, UE, UW, jac);
        adv += u3_dir<T>(ALPHA, ucc, un1, un2, us1, us2, Vabs, VN, VS, jac);
        adv += u3_dir<T>(ALPHA, ucc, ut1, ut2, ub1, ub2, Wabs, WT, WB, jac);
        return adv;
    } else if (SCHEME == FALM_ADV_QUICK) {
        T adv = T(0.);
        adv += quick_dir<T>(ucc, ue1, ue2, uw1, uw2, UE, UW);
        adv += quick_dir<T>(ucc, un1, un2, us1, us2, VN, VS);
        adv += quick_dir<T>(ucc, ut1, ut2, ub1, ub2, WT, WB);
        adv /= (T(16) * jac);
        return adv;
    } else {  // Upwind1st, src/FalmCFDScheme.h:12-31
        T adv = T(0.0);
        adv += upwind_flux1<T>(ucc, ue1, UE);
        adv -= upwind_flux1<T>(uw1, ucc, UW);
        adv += upwind_flux1<T>(ucc, un1, VN);
        adv -= upwind_flux1<T>(us1, ucc, VS);
        adv += upwind_flux1<T>(ucc, ut1, WT);
        adv -= upwind_flux1<T>(ub1, ucc, WB);
        adv /= jac;
        return adv;
    }
}

// Diffusion, src/FalmCFDScheme.h:167-191
template <class T>
__device__ inline T diffusion(T ReI, T ucc, T ue1, T uw1, T un1, T us1, T ut1, T ub1, T nc, T ne, T nw,
                              T nn, T ns, T nt, T nb, T gxc, T gxe, T gxw, T gyc, T gyn, T gys, T gzc,
                              T gzt, T gzb, T jac) {
    T vis = T(0.0);
    vis += (gxc + gxe) * (ReI + T(0.5) * (nc + ne)) * (ue1 - ucc);
    vis -= (gxc + gxw) * (ReI + T(0.5) * (nc + nw)) * (ucc - uw1);
    vis += (gyc + gyn) * (ReI + T(0.5) * (nc + nn)) * (un1 - ucc);
    vis -= (gyc + gys) * (ReI + T(0.5) * (nc + ns)) * (ucc - us1);
    vis += (gzc + gzt) * (ReI + T(0.5) * (nc + nt)) * (ut1 - ucc);
    vis -= (gzc + gzb) * (ReI + T(0.5) * (nc + nb)) * (ucc - ub1);
    vis /= (T(2) * jac);
    return vis;
}

template <class T>
struct Vec3 {
    const T *__restrict__ c[3];
};
template <class T>
struct Vec3W {
    T *__restrict__ c[3];
};

// kernel_Cartesian_CalcPseudoU, src/dev/FalmCFDDevCall.cu:10-301
template <int SCHEME, class T>
__global__ __launch_bounds__(256) void k_pseudo_u(Tiling tl, Geom g, T ReI, T dt, Vec3<T> un, Vec3<T> u,
                                                  Vec3<T> uu, Vec3W<T> ua, const T *__restrict__ nut,
                                                  Vec3<T> kx, Vec3<T> gg, const T *__restrict__ ja, Vec3<T> ff) {
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int sx = 1, sy = g.nx, sz = g.nxy;
    const int cc = g.id(i, j, k);
    const T uc = u.c[0][cc], vc = u.c[1][cc], wc = u.c[2][cc];
    const T Uabs = fabs(uc * kx.c[0][cc]);
    const T Vabs = fabs(vc * kx.c[1][cc]);
    const T Wabs = fabs(wc * kx.c[2][cc]);
    const T UE = uu.c[0][cc], UW = uu.c[0][cc - sx];
    const T VN = uu.c[1][cc], VS = uu.c[1][cc - sy];
    const T WT = uu.c[2][cc], WB = uu.c[2][cc - sz];
    const T nc = nut[cc], ne = nut[cc + sx], nw = nut[cc - sx], nn = nut[cc + sy], ns = nut[cc - sy],
            nt = nut[cc + sz], nb = nut[cc - sz];
    const T gxc = gg.c[0][cc], gxe = gg.c[0][cc + sx], gxw = gg.c[0][cc - sx];
    const T gyc = gg.c[1][cc], gyn = gg.c[1][cc + sy], gys = gg.c[1][cc - sy];
    const T gzc = gg.c[2][cc], gzt = gg.c[2][cc + sz], gzb = gg.c[2][cc - sz];
    const T jac = ja[cc];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T *__restrict__ q = u.c[d];
        const T ucc = d == 0 ? uc : d == 1 ? vc : wc;
        const T ue1 = q[cc + sx], uw1 = q[cc - sx], un1 = q[cc + sy], us1 = q[cc - sy], ut1 = q[cc + sz],
                ub1 = q[cc - sz];
        const T ue2 = q[cc + 2 * sx], uw2 = q[cc - 2 * sx], un2 = q[cc + 2 * sy], us2 = q[cc - 2 * sy],
                ut2 = q[cc + 2 * sz], ub2 = q[cc - 2 * sz];
        T adv = advection<SCHEME, T>(ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2, Uabs,
                                     Vabs, Wabs, UE, UW, VN, VS, WT, WB, jac);
        T vis = diffusion<T>(ReI, ucc, ue1, uw1, un1, us1, ut1, ub1, nc, ne, nw, nn, ns, nt, nb, gxc, gxe,
                             gxw, gyc, gyn, gys, gzc, gzt, gzb, jac);
        ua.c[d][cc] = un.c[d][cc] + dt * (-adv + vis + ff.c[d][cc]);
    }
}

// UtoCU + InterpolateCU fused (src/dev/FalmCFDDevCall.cu:303-349): the
// face flux of face (i+1/2) is the mean of the two cell-centre fluxes
// ja*kx*u, each evaluated exactly as the reference does.
template <class T>
__global__ __launch_bounds__(256) void k_u_to_uu(Tiling tl, Geom g, Vec3<T> u, Vec3W<T> uu, Vec3<T> kx,
                                                 const T *__restrict__ ja) {
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int cc = g.id(i, j, k);
    const int nb[3] = {1, g.nx, g.nxy};
    const T jc = ja[cc];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int cn = cc + nb[d];
        const T ucc = jc * kx.c[d][cc] * u.c[d][cc];
        const T ucn = ja[cn] * kx.c[d][cn] * u.c[d][cn];
        uu.c[d][cc] = T(0.5) * (ucc + ucn);
    }
}

// Divergence (src/dev/FalmCFDDevCall.cu:557-581), optionally followed by the
// ScaleMatrix pass of MACCalcPoissonRHS (src/FalmCFD.h:96-109): the launch
// covers every padded cell; interior cells get div*scale, the others are
// multiplied by scale as ScaleMatrix does over the whole array.
template <class T, class TO, bool SCALE>
__global__ __launch_bounds__(256) void k_divergence(Tiling tl, Geom g, Vec3<T> uu, TO *__restrict__ div,
                                                    const T *__restrict__ ja, double scale) {
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int cc = g.id(i, j, k);
    const bool inner = i >= g.gc && i < g.nx - g.gc && j >= g.gc && j < g.ny - g.gc && k >= g.gc &&
                       k < g.nz - g.gc;
    if (inner) {
        TO UE = uu.c[0][cc], UW = uu.c[0][cc - 1];
        TO VN = uu.c[1][cc], VS = uu.c[1][cc - g.nx];
        TO WT = uu.c[2][cc], WB = uu.c[2][cc - g.nxy];
        TO d = (UE - UW + VN - VS + WT - WB) / (TO)ja[cc];
        if (SCALE) d *= (TO)scale;
        div[cc] = d;
    } else if (SCALE) {
        div[cc] *= (TO)scale;
    }
}

// ProjectPGrid + ProjectPFace (src/dev/FalmCFDDevCall.cu:351-407) in one
// launch over the face region (which contains the interior): cell update on
// interior cells, face update on every face-region cell.  u may alias ua and
// uu may alias uua.  The pressure may be kept in a wider type P.
template <class T, class P>
__global__ __launch_bounds__(256) void k_project(Tiling tl, Geom g, P dt, T *u0, T *u1, T *u2, const T *ua0,
                                                 const T *ua1, const T *ua2, T *uu0, T *uu1, T *uu2,
                                                 const T *uua0, const T *uua1, const T *uua2,
                                                 const P *__restrict__ p, Vec3<T> kx, Vec3<T> gg) {
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int sy = g.nx, sz = g.nxy;
    const int cc = g.id(i, j, k);
    const P pcc = p[cc];
    if (i >= g.gc && j >= g.gc && k >= g.gc) {  // face region upper bound == interior upper bound
        P dpx = P(0.5) * (P)kx.c[0][cc] * (p[cc + 1] - p[cc - 1]);
        P dpy = P(0.5) * (P)kx.c[1][cc] * (p[cc + sy] - p[cc - sy]);
        P dpz = P(0.5) * (P)kx.c[2][cc] * (p[cc + sz] - p[cc - sz]);
        T a0 = ua0[cc], a1 = ua1[cc], a2 = ua2[cc];
        u0[cc] = (T)((P)a0 - dt * dpx);
        u1[cc] = (T)((P)a1 - dt * dpy);
        u2[cc] = (T)((P)a2 - dt * dpz);
    }
    P dfx = P(0.5) * ((P)gg.c[0][cc] + (P)gg.c[0][cc + 1]) * (p[cc + 1] - pcc);
    P dfy = P(0.5) * ((P)gg.c[1][cc] + (P)gg.c[1][cc + sy]) * (p[cc + sy] - pcc);
    P dfz = P(0.5) * ((P)gg.c[2][cc] + (P)gg.c[2][cc + sz]) * (p[cc + sz] - pcc);
    T b0 = uua0[cc], b1 = uua1[cc], b2 = uua2[cc];
    uu0[cc] = (T)((P)b0 - dt * dfx);
    uu1[cc] = (T)((P)b1 - dt * dfy);
    uu2[cc] = (T)((P)b2 - dt * dfz);
}

// kernel_Cartesian_Smagorinsky, src/dev/FalmCFDDevCall.cu:409-476
template <class T>
__global__ __launch_bounds__(256) void k_smagorinsky(Tiling tl, Geom g, T Cs, Vec3<T> u, T *__restrict__ nut,
                                                     Vec3<T> kx, const T *__restrict__ ja) {
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int sy = g.nx, sz = g.nxy;
    const int cc = g.id(i, j, k);
    const T kxx = kx.c[0][cc], kxy = kx.c[1][cc], kxz = kx.c[2][cc];
    T gr[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T *__restrict__ q = u.c[d];
        gr[d][0] = T(0.5) * kxx * (q[cc + 1] - q[cc - 1]);
        gr[d][1] = T(0.5) * kxy * (q[cc + sy] - q[cc - sy]);
        gr[d][2] = T(0.5) * kxz * (q[cc + sz] - q[cc - sz]);
    }
    const T dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
    const T dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
    const T dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
    T d1 = T(2) * dux * dux, d2 = T(2) * dvy * dvy, d3 = T(2) * dwz * dwz;
    T d4 = (duy + dvx) * (duy + dvx), d5 = (dvz + dwy) * (dvz + dwy), d6 = (duz + dwx) * (duz + dwx);
    T Du = sqrt(d1 + d2 + d3 + d4 + d5 + d6);
    T De = cbrt(ja[cc]);
    T lc = Cs * De;
    nut[cc] = lc * lc * Du;
}

// ------------------------------------------------------------ host side
template <class T>
static Vec3<T> v3(const void *p, size_t n) {
    Vec3<T> v;
    for (int d = 0; d < 3; d++) v.c[d] = col<T>(p, n, d);
    return v;
}
template <class T>
static Vec3W<T> v3w(void *p, size_t n) {
    Vec3W<T> v;
    for (int d = 0; d < 3; d++) v.c[d] = col<T>(p, n, d);
    return v;
}

static bool overlaps(const void *a, size_t abytes, const void *b, size_t bbytes) {
    const char *x = (const char *)a, *y = (const char *)b;
    return x < y + bbytes && y < x + abytes;
}

template <class T>
static falm_status pseudo_u_t(const falm_domain *dom, int scheme, double ReI, double dt, const void *un,
                              const void *u, const void *uu, void *ua, const void *nut, const void *kx,
                              const void *gg, const void *ja, const void *ff, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_PSEUDO_U, s);
#define FALM_PU(S)                                                                                   \
    hipLaunchKernelGGL((k_pseudo_u<S, T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (T)ReI, (T)dt, \
                       v3<T>(un, n), v3<T>(u, n), v3<T>(uu, n), v3w<T>(ua, n), (const T *)nut, v3<T>(kx, n), \
                       v3<T>(gg, n), (const T *)ja, v3<T>(ff, n))
    switch (scheme) {
    case FALM_ADV_UPWIND1: FALM_PU(FALM_ADV_UPWIND1); break;
    case FALM_ADV_UPWIND3: FALM_PU(FALM_ADV_UPWIND3); break;
    case FALM_ADV_QUICK: FALM_PU(FALM_ADV_QUICK); break;
    case FALM_ADV_UTOPIA: FALM_PU(FALM_ADV_UTOPIA); break;
    case FALM_ADV_KK: FALM_PU(FALM_ADV_KK); break;
    default: FALM_PU(FALM_ADV_UPWIND1); break;  // str2advscheme's fallback (src/FalmCFD.h:111-125)
    }
#undef FALM_PU
    FALM_CHECK_LAUNCH("k_pseudo_u");
    tm.done((double)tl.box.size() * 23.0 * sizeof(T));
    return FALM_OK;
}

template <class T>
static falm_status u_to_uu_t(const falm_domain *dom, const void *u, void *uu, const void *kx, const void *ja,
                             hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(face_box(g));
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_u_to_uu<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, v3<T>(u, n), v3w<T>(uu, n),
                       v3<T>(kx, n), (const T *)ja);
    FALM_CHECK_LAUNCH("k_u_to_uu");
    return FALM_OK;
}

template <class T, class TO, bool SCALE>
static falm_status divergence_t(const falm_domain *dom, const void *uu, void *div, const void *ja,
                                double scale, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Box b = SCALE ? Box{{0, 0, 0}, {g.nx, g.ny, g.nz}} : interior_box(g);
    Tiling tl = make_tiling(b);
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_divergence<T, TO, SCALE>), tiling_grid(tl), tiling_block(), 0, s, tl, g, v3<T>(uu, n),
                       (TO *)div, (const T *)ja, scale);
    FALM_CHECK_LAUNCH("k_divergence");
    return FALM_OK;
}

template <class T, class P>
static falm_status project_t(const falm_domain *dom, double dt, void *u, const void *ua, void *uu,
                             const void *uua, const void *p, const void *kx, const void *gg, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(face_box(g));
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_project<T, P>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (P)dt, col<T>(u, n, 0),
                       col<T>(u, n, 1), col<T>(u, n, 2), col<T>(ua, n, 0), col<T>(ua, n, 1), col<T>(ua, n, 2),
                       col<T>(uu, n, 0), col<T>(uu, n, 1), col<T>(uu, n, 2), col<T>(uua, n, 0), col<T>(uua, n, 1),
                       col<T>(uua, n, 2), (const P *)p, v3<T>(kx, n), v3<T>(gg, n));
    FALM_CHECK_LAUNCH("k_project");
    return FALM_OK;
}

template <class T>
static falm_status smag_t(const falm_domain *dom, double Cs, const void *u, void *nut, const void *kx,
                          const void *ja, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_smagorinsky<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (T)Cs, v3<T>(u, n),
                       (T *)nut, v3<T>(kx, n), (const T *)ja);
    FALM_CHECK_LAUNCH("k_smagorinsky");
    return FALM_OK;
}

// Entry points shared with the step driver (mixed-type forms).
falm_status launch_poisson_rhs_mixed(const falm_domain *dom, int prec, const void *uu, double *rhs,
                                     const void *ja, double scale, hipStream_t s) {
    if (prec == FALM_F32) return divergence_t<float, double, true>(dom, uu, rhs, ja, scale, s);
    return divergence_t<double, double, true>(dom, uu, rhs, ja, scale, s);
}
falm_status launch_divergence_mixed(const falm_domain *dom, int prec, const void *uu, void *div,
                                    const void *ja, hipStream_t s) {
    if (prec == FALM_F32) return divergence_t<float, float, false>(dom, uu, div, ja, 1.0, s);
    return divergence_t<double, double, false>(dom, uu, div, ja, 1.0, s);
}
falm_status launch_project_mixed(const falm_domain *dom, int prec, double dt, void *u, const void *ua,
                                 void *uu, const void *uua, const double *p, const void *kx, const void *gg,
                                 hipStream_t s) {
    if (prec == FALM_F32) return project_t<float, double>(dom, dt, u, ua, uu, uua, p, kx, gg, s);
    return project_t<double, double>(dom, dt, u, ua, uu, uua, p, kx, gg, s);
}

}  // namespace falm

using namespace falm;

extern "C" {

falm_status falm_pseudo_u(const falm_domain *dom, int prec, int adv_scheme, double ReI, double dt,
                          const void *un, const void *u, const void *uu, void *ua, const void *nut,
                          const void *kx, const void *g, const void *ja, const void *ff, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!un || !u || !uu || !ua || !nut || !kx || !g || !ja || !ff) {
        set_error("falm_pseudo_u: null field pointer");
        return FALM_ERR_INVALID;
    }
    size_t bytes = 3 * ncell_of(dom) * prec_bytes(prec);
    if (overlaps(ua, bytes, u, bytes) || overlaps(ua, bytes, un, bytes)) {
        set_error("falm_pseudo_u: ua overlaps u or un (the update must be out of place)");
        return FALM_ERR_ALIAS;
    }
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32) return pseudo_u_t<float>(dom, adv_scheme, ReI, dt, un, u, uu, ua, nut, kx, g, ja, ff, s);
    return pseudo_u_t<double>(dom, adv_scheme, ReI, dt, un, u, uu, ua, nut, kx, g, ja, ff, s);
}

falm_status falm_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu, const void *kx,
                         const void *ja, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!u || !uu || !kx || !ja) {
        set_error("falm_u_to_uu: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32) return u_to_uu_t<float>(dom, u, uu, kx, ja, s);
    return u_to_uu_t<double>(dom, u, uu, kx, ja, s);
}

falm_status falm_poisson_rhs(const falm_domain *dom, int prec, const void *uu, void *rhs, const void *ja,
                             double dt, double maxdiag, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!uu || !rhs || !ja) {
        set_error("falm_poisson_rhs: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    double scale = 1.0 / (dt * maxdiag);
    if (prec == FALM_F32) return divergence_t<float, float, true>(dom, uu, rhs, ja, scale, s);
    return divergence_t<double, double, true>(dom, uu, rhs, ja, scale, s);
}

falm_status falm_project_p(const falm_domain *dom, int prec, double dt, void *u, const void *ua, void *uu,
                           const void *uua, const void *p, const void *kx, const void *g, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!u || !ua || !uu || !uua || !p || !kx || !g) {
        set_error("falm_project_p: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32) return project_t<float, float>(dom, dt, u, ua, uu, uua, p, kx, g, s);
    return project_t<double, double>(dom, dt, u, ua, uu, uua, p, kx, g, s);
}

falm_status falm_sgs(const falm_domain *dom, int prec, int model, double Cs, const void *u, void *nut,
                     const void *kx, const void *ja, void *stream) {
    FALM_TRY(check_domain(dom));
    if (model == FALM_SGS_EMPTY) return FALM_OK;
    if (model != FALM_SGS_SMAGORINSKY) {
        set_error("falm_sgs: only the Smagorinsky model is implemented");
        return FALM_ERR_UNSUPPORTED;
    }
    if (!u || !nut || !kx || !ja) {
        set_error("falm_sgs: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32) return smag_t<float>(dom, Cs, u, nut, kx, ja, s);
    return smag_t<double>(dom, Cs, u, nut, kx, ja, s);
}

falm_status falm_divergence(const falm_domain *dom, int prec, const void *uu, void *div, const void *ja,
                            void *stream) {
    FALM_TRY(check_domain(dom));
    if (!uu || !div || !ja) {
        set_error("falm_divergence: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32) return divergence_t<float, float, false>(dom, uu, div, ja, 1.0, s);
    return divergence_t<double, double, false>(dom, uu, div, ja, 1.0, s);
}

}  // extern "C"
