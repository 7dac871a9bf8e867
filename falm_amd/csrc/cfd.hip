// cfd.hip — fractional-step operators of FalmCFD on gfx950.
//
// pseudo-velocity (advection + diffusion + body force), contravariant face
// flux, Poisson right-hand side, pressure projection, Smagorinsky eddy
// viscosity and divergence.  All are HBM-bound stencils: one thread per cell
// over 64x4 tiles with an XCD-contiguous tile order; neighbour reuse comes
// from L1/L2 (the x-neighbours of a wavefront are in the same cache lines,
// the y/z-neighbours are re-read by the adjacent rows/planes of the same
// XCD slab while they are still in its L2).
#include "cfd.hpp"
#include "km.hpp"

namespace falm {

// ------------------------------------------------------------ schemes
// src/FalmCFDScheme.h:8-10
template <class T>
__device__ inline T upwind_flux1(T ul, T ur, T UF) {
    return T(0.5) * (UF * (ur + ul) - fabs(UF) * (ur - ul));
}

// The scheme helpers take rjac = 1/J and multiply where the reference
// divides (by 24, 16, J, 2J): a division costs ~10 VALU instructions and the
// pseudo-velocity kernel is VALU-bound; results move by at most an ulp or two
// per term.

// One advection term of the Upwind3 family (Upwind3rd, UTOPIA, KK differ
// only in the fourth-difference weight ALPHA; src/FalmCFDScheme.h:68-165).
template <class T>
__device__ inline T u3_dir(T ALPHA, T ucc, T up1, T up2, T um1, T um2, T Uabs, T JP, T JM, T rjac) {
    const T r24 = T(1. / 24.);
    T UDL = JP * (-up2 + T(27) * up1 - T(27) * ucc + um1) * r24;
    T UDR = JM * (-up1 + T(27) * ucc - T(27) * um1 + um2) * r24;
    T UD4 = Uabs * (up2 - T(4) * up1 + T(6) * ucc - T(4) * um1 + um2) * ALPHA;
    return T(.5) * (UDL + UDR) * rjac + UD4;
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
                              T WT, T WB, T rjac) {
    if (SCHEME == FALM_ADV_UPWIND3 || SCHEME == FALM_ADV_UTOPIA || SCHEME == FALM_ADV_KK) {
        const T ALPHA = SCHEME == FALM_ADV_UPWIND3 ? T(1. / 24.) : SCHEME == FALM_ADV_UTOPIA ? T(1. / 12.) : T(1. / 4.);
        T adv = T(0.0);
        adv += u3_dir<T>(ALPHA, ucc, ue1, ue2, uw1, uw2, Uabs, UE, UW, rjac);
        adv += u3_dir<T>(ALPHA, ucc, un1, un2, us1, us2, Vabs, VN, VS, rjac);
        adv += u3_dir<T>(ALPHA, ucc, ut1, ut2, ub1, ub2, Wabs, WT, WB, rjac);
        return adv;
    } else if (SCHEME == FALM_ADV_QUICK) {
        T adv = T(0.);
        adv += quick_dir<T>(ucc, ue1, ue2, uw1, uw2, UE, UW);
        adv += quick_dir<T>(ucc, un1, un2, us1, us2, VN, VS);
        adv += quick_dir<T>(ucc, ut1, ut2, ub1, ub2, WT, WB);
        adv *= T(1. / 16.) * rjac;
        return adv;
    } else {  // Upwind1st, src/FalmCFDScheme.h:12-31
        T adv = T(0.0);
        adv += upwind_flux1<T>(ucc, ue1, UE);
        adv -= upwind_flux1<T>(uw1, ucc, UW);
        adv += upwind_flux1<T>(ucc, un1, VN);
        adv -= upwind_flux1<T>(us1, ucc, VS);
        adv += upwind_flux1<T>(ucc, ut1, WT);
        adv -= upwind_flux1<T>(ub1, ucc, WB);
        adv *= rjac;
        return adv;
    }
}

// Diffusion, src/FalmCFDScheme.h:167-191
template <class T>
__device__ inline T diffusion(T ReI, T ucc, T ue1, T uw1, T un1, T us1, T ut1, T ub1, T nc, T ne, T nw,
                              T nn, T ns, T nt, T nb, T gxc, T gxe, T gxw, T gyc, T gyn, T gys, T gzc,
                              T gzt, T gzb, T rjac) {
    T vis = T(0.0);
    vis += (gxc + gxe) * (ReI + T(0.5) * (nc + ne)) * (ue1 - ucc);
    vis -= (gxc + gxw) * (ReI + T(0.5) * (nc + nw)) * (ucc - uw1);
    vis += (gyc + gyn) * (ReI + T(0.5) * (nc + nn)) * (un1 - ucc);
    vis -= (gyc + gys) * (ReI + T(0.5) * (nc + ns)) * (ucc - us1);
    vis += (gzc + gzt) * (ReI + T(0.5) * (nc + nt)) * (ut1 - ucc);
    vis -= (gzc + gzb) * (ReI + T(0.5) * (nc + nb)) * (ucc - ub1);
    vis *= T(0.5) * rjac;
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

// Metric access.  MetField reads the per-cell metric fields kx, g, ja of the
// reference (FalmCore::computation_init, src/falm.h:383-404).  MetSep
// evaluates the same expressions from the 1-D cell widths of the
// tensor-product mesh (volume = hx*hy*hz, kx = 1/h, g = volume*kx*kx, in
// double, then rounded to T), so both give the same bits and MetSep reads
// no per-cell metric array.
template <class T>
struct MetField {
    const T *__restrict__ kx[3];
    const T *__restrict__ g[3];
    const T *__restrict__ ja;
    __device__ inline T KX(int d, int cc, int, int, int) const { return kx[d][cc]; }
    __device__ inline T G(int d, int cc, int, int, int) const { return g[d][cc]; }
    __device__ inline T JA(int cc, int, int, int) const { return ja[cc]; }
    template <class V>
    __device__ inline V div_ja(V x, int cc, int, int, int) const { return x / (V)ja[cc]; }
    static constexpr bool per_cell = true;  // metric arrays count as traffic
};
template <class T>
struct MetSep {
    const double *__restrict__ h[3];   // cell widths along x, y, z
    const double *__restrict__ rh[3];  // 1.0/h along x, y, z
    __device__ inline double vol(int i, int j, int k) const { return h[0][i] * h[1][j] * h[2][k]; }
    __device__ inline double r(int d, int i, int j, int k) const { return rh[d][d == 0 ? i : d == 1 ? j : k]; }
    __device__ inline T KX(int d, int, int i, int j, int k) const { return (T)r(d, i, j, k); }
    __device__ inline T G(int d, int, int i, int j, int k) const {
        const double q = r(d, i, j, k);
        return (T)(vol(i, j, k) * q * q);
    }
    __device__ inline T JA(int, int i, int j, int k) const { return (T)vol(i, j, k); }
    template <class V>
    __device__ inline V div_ja(V x, int cc, int i, int j, int k) const { return x / (V)JA(cc, i, j, k); }
    static constexpr bool per_cell = false;
};

// kernel_Cartesian_CalcPseudoU, src/dev/FalmCFDDevCall.cu:10-301
// ffbox (optional): ff is known to be zero outside this cell box
// (ilo, ihi, jlo, jhi, klo, khi) and is not read there.
template <int SCHEME, class T, class M>
__global__ __launch_bounds__(256) void k_pseudo_u(Tiling tl, Geom g, T ReI, T dt, Vec3<T> un, Vec3<T> u,
                                                  Vec3<T> uu, Vec3W<T> ua, const T *__restrict__ nut, M met,
                                                  Vec3<T> ff, const int *__restrict__ ffbox) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const bool useff = !ffbox || (i >= ffbox[0] && i <= ffbox[1] && j >= ffbox[2] && j <= ffbox[3] &&
                                      k >= ffbox[4] && k <= ffbox[5]);
        const int sx = 1, sy = g.nx, sz = g.nxy;
        const int cc = g.id(i, j, k);
        const T uc = u.c[0][cc], vc = u.c[1][cc], wc = u.c[2][cc];
        const T Uabs = fabs(uc * met.KX(0, cc, i, j, k));
        const T Vabs = fabs(vc * met.KX(1, cc, i, j, k));
        const T Wabs = fabs(wc * met.KX(2, cc, i, j, k));
        const T UE = uu.c[0][cc], UW = uu.c[0][cc - sx];
        const T VN = uu.c[1][cc], VS = uu.c[1][cc - sy];
        const T WT = uu.c[2][cc], WB = uu.c[2][cc - sz];
        const T nc = nut[cc], ne = nut[cc + sx], nw = nut[cc - sx], nn = nut[cc + sy], ns = nut[cc - sy],
                nt = nut[cc + sz], nb = nut[cc - sz];
        const T gxc = met.G(0, cc, i, j, k), gxe = met.G(0, cc + sx, i + 1, j, k), gxw = met.G(0, cc - sx, i - 1, j, k);
        const T gyc = met.G(1, cc, i, j, k), gyn = met.G(1, cc + sy, i, j + 1, k), gys = met.G(1, cc - sy, i, j - 1, k);
        const T gzc = met.G(2, cc, i, j, k), gzt = met.G(2, cc + sz, i, j, k + 1), gzb = met.G(2, cc - sz, i, j, k - 1);
        const T jac = met.JA(cc, i, j, k);
    const T rjac = T(1) / jac;
    #pragma unroll
        for (int d = 0; d < 3; d++) {
            const T *__restrict__ q = u.c[d];
            const T ucc = d == 0 ? uc : d == 1 ? vc : wc;
            const T ue1 = q[cc + sx], uw1 = q[cc - sx], un1 = q[cc + sy], us1 = q[cc - sy], ut1 = q[cc + sz],
                    ub1 = q[cc - sz];
            const T ue2 = q[cc + 2 * sx], uw2 = q[cc - 2 * sx], un2 = q[cc + 2 * sy], us2 = q[cc - 2 * sy],
                    ut2 = q[cc + 2 * sz], ub2 = q[cc - 2 * sz];
            T adv = advection<SCHEME, T>(ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2, Uabs,
                                         Vabs, Wabs, UE, UW, VN, VS, WT, WB, rjac);
            T vis = diffusion<T>(ReI, ucc, ue1, uw1, un1, us1, ut1, ub1, nc, ne, nw, nn, ns, nt, nb, gxc, gxe,
                                 gxw, gyc, gyn, gys, gzc, gzt, gzb, rjac);
            const T f = useff ? ff.c[d][cc] : T(0);
            ua.c[d][cc] = un.c[d][cc] + dt * (-adv + vis + f);
        }
    });
}

// ---------------------------------------------------------------------
// Pseudo-velocity on a tensor-product mesh, k-marching form.  A workgroup
// owns a KM_BX x KM_BY column of cells and walks KM_KC planes in k.  The
// in-plane neighbours (x, y within +-2) come from an LDS copy of the plane
// with a two-cell halo, the k neighbours from a per-thread register queue,
// and the metrics from the 1-D widths (per thread once, per plane as
// workgroup-uniform values).  Each cell then costs ~10 vector loads instead
// of ~60; the arithmetic is the same expression by expression as
// k_pseudo_u, so the results are identical.  The next plane's loads are
// issued before the current plane is computed, and the LDS planes are
// double-buffered, so one barrier per plane suffices.
#ifndef FALM_KM_KC
#define FALM_KM_KC 16
#endif
constexpr int KM_BX = 64, KM_BY = 8, KM_KC = FALM_KM_KC;
constexpr int KM_NT = KM_BX * KM_BY;
constexpr int KM_UX = 4 * KM_BY, KM_UY = 4 * KM_BX;   // u halo cells per component
constexpr int KM_NX = 2 * KM_BY, KM_NY = 2 * KM_BX;   // nut halo cells
constexpr int KM_HALO = 3 * (KM_UX + KM_UY) + KM_NX + KM_NY;
constexpr int KM_HPT = (KM_HALO + KM_NT - 1) / KM_NT;  // halo cells per thread

struct KmTiling {
    Box box;
    int tx, ty, tk;  // column tiles along x, y and k chunks
    int blocks;
};
inline KmTiling make_km_tiling(const Box &b) {
    KmTiling t;
    t.box = b;
    t.tx = (b.n[0] + KM_BX - 1) / KM_BX;
    t.ty = (b.n[1] + KM_BY - 1) / KM_BY;
    t.tk = (b.n[2] + KM_KC - 1) / KM_KC;
    t.blocks = t.tx * t.ty * t.tk;
    return t;
}

// halo slot -> (LDS array: 0..2 u components, 3 nut; LDS row, column)
__device__ inline void km_halo_slot(int s, int &arr, int &row, int &col) {
    if (s < 3 * (KM_UX + KM_UY)) {
        arr = s / (KM_UX + KM_UY);
        int r = s - arr * (KM_UX + KM_UY);
        if (r < KM_UX) {  // x halo: rows 2..BY+1, columns 0,1,BX+2,BX+3
            row = 2 + r / 4;
            const int q = r & 3;
            col = q < 2 ? q : KM_BX + q;
        } else {          // y halo: rows 0,1,BY+2,BY+3, columns 2..BX+1
            r -= KM_UX;
            const int q = r / KM_BX;
            row = q < 2 ? q : KM_BY + q;
            col = 2 + (r - q * KM_BX);
        }
    } else {
        arr = 3;
        int r = s - 3 * (KM_UX + KM_UY);
        if (r < KM_NX) {  // rows 1..BY, columns 0, BX+1
            row = 1 + (r >> 1);
            col = (r & 1) ? KM_BX + 1 : 0;
        } else {          // rows 0, BY+1, columns 1..BX
            r -= KM_NX;
            const int q = r / KM_BX;
            row = q ? KM_BY + 1 : 0;
            col = 1 + (r - q * KM_BX);
        }
    }
}

template <int SCHEME, class T>
__global__ __launch_bounds__(KM_NT) void k_pseudo_u_km(KmTiling kt, Geom g, T ReI, T dt, Vec3<T> u, Vec3<T> uu,
                                                       Vec3W<T> ua, const T *__restrict__ nut, MetSep<T> met,
                                                       Vec3<T> ff, const int *__restrict__ ffbox) {
    __shared__ T su[2][3][KM_BY + 4][KM_BX + 4];
    __shared__ T sn[2][KM_BY + 2][KM_BX + 2];
    const int lane = threadIdx.x & (KM_BX - 1), wy = threadIdx.x / KM_BX;
    const int L = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = L % kt.tx, by = (L / kt.tx) % kt.ty, kc = L / (kt.tx * kt.ty);
    const int i0 = kt.box.lo[0] + bx * KM_BX, j0 = kt.box.lo[1] + by * KM_BY;
    const int i = i0 + lane, j = j0 + wy;
    const bool act = i < kt.box.lo[0] + kt.box.n[0] && j < kt.box.lo[1] + kt.box.n[1];
    const int k0 = kt.box.lo[2] + kc * KM_KC;
    const int k1 = min(k0 + KM_KC, kt.box.lo[2] + kt.box.n[2]);
    const int sy = g.nx, sz = g.nxy;
    // Threads past the box still load their own cell into the LDS plane
    // when it lies in the array: the last active columns and rows read it
    // as a neighbour.  Threads past the array load a valid cell instead.
    const bool ld = i < g.nx && j < g.ny;
    const int col0 = (ld ? i : kt.box.lo[0]) + sy * (ld ? j : kt.box.lo[1]);
    // per-column metric factors (cells of the box only)
    const int ia = act ? i : kt.box.lo[0], ja = act ? j : kt.box.lo[1];
    const double hxc = met.h[0][ia], hxe = met.h[0][ia + 1], hxw = met.h[0][ia - 1];
    const double rxc = met.rh[0][ia], rxe = met.rh[0][ia + 1], rxw = met.rh[0][ia - 1];
    const double hyc = met.h[1][ja], hyn = met.h[1][ja + 1], hys = met.h[1][ja - 1];
    const double ryc = met.rh[1][ja], ryn = met.rh[1][ja + 1], rys = met.rh[1][ja - 1];
    bool inx = true, iny = true;
    if (ffbox) {
        inx = i >= ffbox[0] && i <= ffbox[1];
        iny = j >= ffbox[2] && j <= ffbox[3];
    }
    // halo slots of this thread
    int harr[KM_HPT], hrow[KM_HPT], hcol[KM_HPT], hoff[KM_HPT];
    bool hok[KM_HPT];
#pragma unroll
    for (int q = 0; q < KM_HPT; q++) {
        const int s = threadIdx.x + q * KM_NT;
        hok[q] = false;
        harr[q] = hrow[q] = hcol[q] = hoff[q] = 0;
        if (s < KM_HALO) {
            km_halo_slot(s, harr[q], hrow[q], hcol[q]);
            const int gi = i0 - (harr[q] < 3 ? 2 : 1) + hcol[q];
            const int gj = j0 - (harr[q] < 3 ? 2 : 1) + hrow[q];
            hok[q] = gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
            hoff[q] = gi + sy * gj;
        }
    }
    auto halo_load = [&](int k, T (&hv)[KM_HPT]) {
#pragma unroll
        for (int q = 0; q < KM_HPT; q++) {
            hv[q] = T(0);
            if (hok[q]) {
                const T *src = harr[q] < 3 ? u.c[harr[q]] : nut;
                hv[q] = src[hoff[q] + sz * k];
            }
        }
    };
    auto lds_fill = [&](int b, const T (&hv)[KM_HPT], T u0, T u1, T u2, T nv) {
        su[b][0][wy + 2][lane + 2] = u0;
        su[b][1][wy + 2][lane + 2] = u1;
        su[b][2][wy + 2][lane + 2] = u2;
        sn[b][wy + 1][lane + 1] = nv;
#pragma unroll
        for (int q = 0; q < KM_HPT; q++) {
            if (threadIdx.x + q * KM_NT >= KM_HALO) continue;
            if (harr[q] < 3) su[b][harr[q]][hrow[q]][hcol[q]] = hv[q];
            else sn[b][hrow[q]][hcol[q]] = hv[q];
        }
    };
    // Registers carry everything plane k needs (the k queue, its uu and ff
    // values); the loads for plane k+1 are issued before plane k is computed
    // (halo values first, so waiting for them leaves the rest in flight).
    auto ffin = [&](int k) { return act && inx && iny && (!ffbox || (k >= ffbox[4] && k <= ffbox[5])); };
    T qu[3][5], qn[3], cuu[5], cff[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int d = 0; d < 5; d++) qu[c][d] = u.c[c][col0 + sz * (k0 - 2 + d)];
    qn[0] = nut[col0 + sz * (k0 - 1)];
    qn[1] = nut[col0 + sz * k0];
    qn[2] = nut[col0 + sz * (k0 + 1)];
    T wb = uu.c[2][col0 + sz * (k0 - 1)];
    {
        const int cc = col0 + sz * k0;
        cuu[0] = uu.c[0][cc];
        cuu[1] = uu.c[0][cc - 1];
        cuu[2] = uu.c[1][cc];
        cuu[3] = uu.c[1][cc - sy];
        cuu[4] = uu.c[2][cc];
        const bool f = ffin(k0);
        for (int d = 0; d < 3; d++) cff[d] = f ? ff.c[d][cc] : T(0);
    }
    T hv[KM_HPT];
    halo_load(k0, hv);
    lds_fill(0, hv, qu[0][2], qu[1][2], qu[2][2], qn[1]);
    __syncthreads();
    int b = 0;
    for (int k = k0; k < k1; k++) {
        const int cc = col0 + sz * k;
        const bool more = k + 1 < k1;
        // loads for plane k+1
        T nu[3] = {T(0), T(0), T(0)}, nn = T(0), nuu[5] = {T(0), T(0), T(0), T(0), T(0)};
        T nff[3] = {T(0), T(0), T(0)};
        if (more) {
            halo_load(k + 1, hv);
            const int cn = cc + sz;
            nu[0] = u.c[0][cn + 2 * sz];
            nu[1] = u.c[1][cn + 2 * sz];
            nu[2] = u.c[2][cn + 2 * sz];
            nn = nut[cn + sz];
            nuu[0] = uu.c[0][cn];
            nuu[1] = uu.c[0][cn - 1];
            nuu[2] = uu.c[1][cn];
            nuu[3] = uu.c[1][cn - sy];
            nuu[4] = uu.c[2][cn];
            if (ffin(k + 1)) {
                nff[0] = ff.c[0][cn];
                nff[1] = ff.c[1][cn];
                nff[2] = ff.c[2][cn];
            }
        }
        // plane k
        const T UE = cuu[0], UW = cuu[1], VN = cuu[2], VS = cuu[3], WT = cuu[4], WB = wb;
        const double hzc = ld_uniform(met.h[2], k), hzt = ld_uniform(met.h[2], k + 1), hzb = ld_uniform(met.h[2], k - 1);
        const double rzc = ld_uniform(met.rh[2], k), rzt = ld_uniform(met.rh[2], k + 1), rzb = ld_uniform(met.rh[2], k - 1);
        const double volc = hxc * hyc * hzc;
        const T gxc = (T)(volc * rxc * rxc), gxe = (T)(hxe * hyc * hzc * rxe * rxe), gxw = (T)(hxw * hyc * hzc * rxw * rxw);
        const T gyc = (T)(volc * ryc * ryc), gyn = (T)(hxc * hyn * hzc * ryn * ryn), gys = (T)(hxc * hys * hzc * rys * rys);
        const T gzc = (T)(volc * rzc * rzc), gzt = (T)(hxc * hyc * hzt * rzt * rzt), gzb = (T)(hxc * hyc * hzb * rzb * rzb);
        const T jac = (T)volc;
        const T rjac = T(1) / jac;
        const T Uabs = fabs(qu[0][2] * (T)rxc);
        const T Vabs = fabs(qu[1][2] * (T)ryc);
        const T Wabs = fabs(qu[2][2] * (T)rzc);
        const T nc = qn[1], ne = sn[b][wy + 1][lane + 2], nw = sn[b][wy + 1][lane], nn1 = sn[b][wy + 2][lane + 1],
                ns = sn[b][wy][lane + 1], nt = qn[2], nb = qn[0];
        if (act) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const T ucc = qu[d][2];
                const T ue1 = su[b][d][wy + 2][lane + 3], ue2 = su[b][d][wy + 2][lane + 4];
                const T uw1 = su[b][d][wy + 2][lane + 1], uw2 = su[b][d][wy + 2][lane];
                const T un1 = su[b][d][wy + 3][lane + 2], un2 = su[b][d][wy + 4][lane + 2];
                const T us1 = su[b][d][wy + 1][lane + 2], us2 = su[b][d][wy][lane + 2];
                const T ut1 = qu[d][3], ut2 = qu[d][4], ub1 = qu[d][1], ub2 = qu[d][0];
                T adv = advection<SCHEME, T>(ucc, ue1, ue2, uw1, uw2, un1, un2, us1, us2, ut1, ut2, ub1, ub2, Uabs,
                                             Vabs, Wabs, UE, UW, VN, VS, WT, WB, rjac);
                T vis = diffusion<T>(ReI, ucc, ue1, uw1, un1, us1, ut1, ub1, nc, ne, nw, nn1, ns, nt, nb, gxc, gxe,
                                     gxw, gyc, gyn, gys, gzc, gzt, gzb, rjac);
                ua.c[d][cc] = ucc + dt * (-adv + vis + cff[d]);
            }
        }
        // plane k+1 into the other LDS buffer; rotate the registers
        if (more) lds_fill(b ^ 1, hv, qu[0][3], qu[1][3], qu[2][3], qn[2]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            qu[c][0] = qu[c][1];
            qu[c][1] = qu[c][2];
            qu[c][2] = qu[c][3];
            qu[c][3] = qu[c][4];
            qu[c][4] = nu[c];
        }
        qn[0] = qn[1];
        qn[1] = qn[2];
        qn[2] = nn;
        wb = WT;
#pragma unroll
        for (int d = 0; d < 5; d++) cuu[d] = nuu[d];
#pragma unroll
        for (int d = 0; d < 3; d++) cff[d] = nff[d];
        b ^= 1;
        __syncthreads();
    }
}

// Device view of the coefficient tables in one precision.
template <class T>
struct SepTab {
    const T *__restrict__ t[SEP_NTAB][3];
};

// Metric access from the tables (step driver): division by the cell
// volume becomes a product of the three reciprocal widths.
template <class T>
struct MetTab {
    SepTab<T> tb;
    template <class V>
    __device__ inline V div_ja(V x, int, int i, int j, int k) const {
        return x * (V)(tb.t[SEP_R][0][i] * tb.t[SEP_R][1][j] * tb.t[SEP_R][2][k]);
    }
    static constexpr bool per_cell = false;
};

// ---------------------------------------------------------------------
// Pseudo-velocity for the Upwind3 family on a tensor-product mesh: the same
// k-marching structure as k_pseudo_u_km with the arithmetic reorganised
// around face differences (the scheme is unchanged, its terms are summed
// in another order):
//   along axis e with v(m) = u_d(c + m e) and the face differences
//     a+ = v(1) - v(0),  b+ = v(2) - v(-1),  a- = v(0) - v(-1),  b- = v(1) - v(-2)
//   UDL = JP (27 a+ - b+)/24,  UDR = JM (27 a- - b-)/24   (src/FalmCFDScheme.h:68-165)
//   4th difference = (b+ - 3 a+) - (b- - 3 a-)
//   diffusion = K+ a+ - K- a-, K = (g_c + g_n)/(2J) (ReI + (nut_c + nut_n)/2)
//              (src/FalmCFDScheme.h:167-191)
// and the metric factors come from the per-axis tables (per column once,
// per plane as workgroup-uniform loads).  The k-axis plus-face terms of
// plane k are the minus-face terms of plane k+1 and are carried in
// registers.  Compared with k_pseudo_u_km this removes the double-precision
// metric products, the division by the cell volume and about half of the
// scheme's operations; FP64 results agree with the reference expression
// order to rounding.
template <int SCHEME, class T>
__global__ __launch_bounds__(KM_NT) void k_pseudo_u3_km(KmTiling kt, Geom g, T ReI, T dt, Vec3<T> u, Vec3<T> uu,
                                                        Vec3W<T> ua, const T *__restrict__ nut, SepTab<T> tb,
                                                        Vec3<T> ff, const int *__restrict__ ffbox) {
    constexpr double ALPHA_D = SCHEME == FALM_ADV_UPWIND3 ? 1. / 24. : SCHEME == FALM_ADV_UTOPIA ? 1. / 12. : 1. / 4.;
    const T ALPHA = T(ALPHA_D);
    constexpr unsigned TS = sizeof(T);
    __shared__ T su[2][3][KM_BY + 4][KM_BX + 4];
    __shared__ T sn[2][KM_BY + 2][KM_BX + 2];
    const int lane = threadIdx.x & (KM_BX - 1);
    const int wy = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / KM_BX));
    const int L = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x));
    const int bx = L % kt.tx, by = (L / kt.tx) % kt.ty, kc = L / (kt.tx * kt.ty);
    const int i0 = kt.box.lo[0] + bx * KM_BX, j0 = kt.box.lo[1] + by * KM_BY;
    const int i = i0 + lane, j = j0 + wy;
    const bool act = i < kt.box.lo[0] + kt.box.n[0] && j < kt.box.lo[1] + kt.box.n[1];
    const int k0 = kt.box.lo[2] + kc * KM_KC;
    const int k1 = min(k0 + KM_KC, kt.box.lo[2] + kt.box.n[2]);
    const int sy = g.nx, sz = g.nxy;
    const unsigned syb = (unsigned)sy * TS, szb = (unsigned)sz * TS;
    // as in k_pseudo_u_km: threads past the box load their own cell when it
    // lies in the array, threads past the array a valid cell
    const bool ld = i < g.nx && j < g.ny;
    const int col0 = (ld ? i : kt.box.lo[0]) + sy * (ld ? j : kt.box.lo[1]);
    // per-column coefficients (cells of the box only)
    const int ia = act ? i : kt.box.lo[0], ja = act ? j : kt.box.lo[1];
    const T rx = tb.t[SEP_R][0][ia], ry = tb.t[SEP_R][1][ja];
    const T cxe = tb.t[SEP_CP][0][ia], cxw = tb.t[SEP_CM][0][ia];
    const T cyn = tb.t[SEP_CP][1][ja], cys = tb.t[SEP_CM][1][ja];
    const T arx = ALPHA * rx, ary = ALPHA * ry;
    const T rxy = rx * ry * T(0.5 / 24.0);
    // body force: the column's part of the force box and its k range
    bool fcol = act;
    int fk0 = k0, fk1 = k1 - 1;
    if (ffbox) {
        fcol = fcol && i >= ffbox[0] && i <= ffbox[1] && j >= ffbox[2] && j <= ffbox[3];
        fk0 = ffbox[4];
        fk1 = ffbox[5];
    }
    // halo slots: a base pointer per slot (an in-array cell when the halo
    // cell lies outside the array; its value is then replaced by zero)
    const T *hp[KM_HPT];
    int hl[KM_HPT];  // LDS element index of the slot in plane buffer 0
    bool hok[KM_HPT], hu[KM_HPT];
#pragma unroll
    for (int q = 0; q < KM_HPT; q++) {
        const int s = threadIdx.x + q * KM_NT;
        int arr = 3, row = 0, col = 0;
        hok[q] = false;
        if (s < KM_HALO) km_halo_slot(s, arr, row, col);
        const int gi = i0 - (arr < 3 ? 2 : 1) + col;
        const int gj = j0 - (arr < 3 ? 2 : 1) + row;
        hok[q] = s < KM_HALO && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        hu[q] = s < KM_HALO;
        const T *src = arr == 0 ? u.c[0] : arr == 1 ? u.c[1] : arr == 2 ? u.c[2] : nut;
        hp[q] = src + (hok[q] ? gi + sy * gj : col0);
        hl[q] = arr < 3 ? (&su[0][arr][row][col] - &su[0][0][0][0]) : (&sn[0][row][col] - &sn[0][0][0]);
        if (arr == 3) hl[q] = -1 - hl[q];  // negative: the nut plane
    }
    auto halo_load = [&](int k, T (&hv)[KM_HPT]) {
#pragma unroll
        for (int q = 0; q < KM_HPT; q++) {
            const T v = hp[q][(size_t)sz * k];
            hv[q] = hok[q] ? v : T(0);
        }
    };
    auto lds_fill = [&](int b, const T (&hv)[KM_HPT], T u0, T u1, T u2, T nv) {
        su[b][0][wy + 2][lane + 2] = u0;
        su[b][1][wy + 2][lane + 2] = u1;
        su[b][2][wy + 2][lane + 2] = u2;
        sn[b][wy + 1][lane + 1] = nv;
#pragma unroll
        for (int q = 0; q < KM_HPT; q++) {
            if (!hu[q]) continue;
            if (hl[q] >= 0) (&su[b][0][0][0])[hl[q]] = hv[q];
            else (&sn[b][0][0])[-1 - hl[q]] = hv[q];
        }
    };
    // column queues: u at k-1 .. k+2, nut at k-1 .. k+1, the face fluxes and
    // body force of plane k.  Loads run PD planes ahead: while plane k is
    // computed, the loads of planes k+1 .. k+PD-1 are in flight or in
    // registers and plane k+PD's are issued.  Two planes ahead in float;
    // one in double, where the second stage's registers would halve the
    // waves per SIMD.
    constexpr int PD = sizeof(T) == 4 ? 2 : 1;
    unsigned cb = (unsigned)(col0 + sz * k0) * TS;  // byte offset of the column's plane-k cell
    T qu[3][4], qn[3], cuu[5], cff[3];
    T pzm[3], tzm[3], azm[3];  // k-axis minus-face terms
    // pending loads of planes k+1 .. k+PD (stage 0 = plane k+1): u(kn+2),
    // nut(kn+1), face fluxes, force and halo of plane kn
    T pu[PD][3], pn[PD], puu[PD][5], pff[PD][3], phv[PD][KM_HPT];
    auto issue = [&](int kn, unsigned nb, T (&xu)[3], T &xn, T (&xuu)[5], T (&xff)[3], T (&xhv)[KM_HPT]) {
        // loads of plane kn (byte offset nb): its queue tail u(kn+2), nut(kn+1)
        halo_load(kn, xhv);
        xu[0] = ld_at(u.c[0], nb + 2 * szb);
        xu[1] = ld_at(u.c[1], nb + 2 * szb);
        xu[2] = ld_at(u.c[2], nb + 2 * szb);
        xn = ld_at(nut, nb + szb);
        xuu[0] = ld_at(uu.c[0], nb);
        xuu[1] = ld_at(uu.c[0], nb - TS);
        xuu[2] = ld_at(uu.c[1], nb);
        xuu[3] = ld_at(uu.c[1], nb - syb);
        xuu[4] = ld_at(uu.c[2], nb);
        const bool f = fcol && kn >= fk0 && kn <= fk1;
        for (int d = 0; d < 3; d++) xff[d] = f ? ld_at(ff.c[d], nb) : T(0);
    };
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const T um2 = ld_at(u.c[c], cb - 2 * szb);
#pragma unroll
        for (int d = 0; d < 4; d++) qu[c][d] = ld_at(u.c[c], cb + (unsigned)(d - 1) * szb);
        const T am = qu[c][1] - qu[c][0], bm = qu[c][2] - um2;
        pzm[c] = T(27) * am - bm;
        tzm[c] = bm - T(3) * am;
        azm[c] = am;
    }
    qn[0] = ld_at(nut, cb - szb);
    qn[1] = ld_at(nut, cb);
    qn[2] = ld_at(nut, cb + szb);
    T wb = ld_at(uu.c[2], cb - szb);
    cuu[0] = ld_at(uu.c[0], cb);
    cuu[1] = ld_at(uu.c[0], cb - TS);
    cuu[2] = ld_at(uu.c[1], cb);
    cuu[3] = ld_at(uu.c[1], cb - syb);
    cuu[4] = ld_at(uu.c[2], cb);
    {
        const bool f = fcol && k0 >= fk0 && k0 <= fk1;
        for (int d = 0; d < 3; d++) cff[d] = f ? ld_at(ff.c[d], cb) : T(0);
    }
    // planes past the chunk's end are loaded from its last plane (in the
    // array, values unused) so that the unrolled loop has no branches
    const int klast = k1 - 1;
    {
        T hv0[KM_HPT];
        halo_load(k0, hv0);
#pragma unroll
        for (int st = 0; st + 1 < PD; st++) {
            const int kn = min(k0 + 1 + st, klast);
            issue(kn, (unsigned)(col0 + sz * kn) * TS, pu[st], pn[st], puu[st], pff[st], phv[st]);
        }
        lds_fill(0, hv0, qu[0][1], qu[1][1], qu[2][1], qn[1]);
    }
    __syncthreads();
    // fully unrolled over the chunk's planes: the register queues and the LDS
    // buffer index then rotate by renaming instead of moves
#pragma unroll
    for (int kk = 0; kk < KM_KC; kk++) {
        const int k = k0 + kk;
        if (k >= k1) break;
        const int b = kk & 1;
        {
            const int kn = min(k + PD, klast);
            issue(kn, (unsigned)(col0 + sz * kn) * TS, pu[PD - 1], pn[PD - 1], puu[PD - 1], pff[PD - 1],
                  phv[PD - 1]);
        }
        // plane k
        const T rz = ld_uniform(tb.t[SEP_R][2], k), czt = ld_uniform(tb.t[SEP_CP][2], k),
                czb = ld_uniform(tb.t[SEP_CM][2], k);
        const T rj = rxy * rz;  // (1/2)(1/24)/J
        const T aux = fabs(qu[0][1] * arx), auy = fabs(qu[1][1] * ary), auz = fabs(qu[2][1] * (ALPHA * rz));
        const T nc = qn[1];
        const T ke = cxe * (ReI + T(0.5) * (nc + sn[b][wy + 1][lane + 2]));
        const T kw = cxw * (ReI + T(0.5) * (nc + sn[b][wy + 1][lane]));
        const T kn = cyn * (ReI + T(0.5) * (nc + sn[b][wy + 2][lane + 1]));
        const T ks = cys * (ReI + T(0.5) * (nc + sn[b][wy][lane + 1]));
        const T kt_ = czt * (ReI + T(0.5) * (nc + qn[2]));
        const T kb = czb * (ReI + T(0.5) * (nc + qn[0]));
        const T UE = cuu[0], UW = cuu[1], VN = cuu[2], VS = cuu[3], WT = cuu[4], WB = wb;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T ucc = qu[d][1];
            T F, D, V;
            {   // x
                const T ue1 = su[b][d][wy + 2][lane + 3], ue2 = su[b][d][wy + 2][lane + 4];
                const T uw1 = su[b][d][wy + 2][lane + 1], uw2 = su[b][d][wy + 2][lane];
                const T ap = ue1 - ucc, bp = ue2 - uw1, am = ucc - uw1, bm = ue1 - uw2;
                F = UE * (T(27) * ap - bp) + UW * (T(27) * am - bm);
                D = aux * ((bp - T(3) * ap) - (bm - T(3) * am));
                V = ke * ap - kw * am;
            }
            {   // y
                const T un1 = su[b][d][wy + 3][lane + 2], un2 = su[b][d][wy + 4][lane + 2];
                const T us1 = su[b][d][wy + 1][lane + 2], us2 = su[b][d][wy][lane + 2];
                const T ap = un1 - ucc, bp = un2 - us1, am = ucc - us1, bm = un1 - us2;
                F += VN * (T(27) * ap - bp) + VS * (T(27) * am - bm);
                D += auy * ((bp - T(3) * ap) - (bm - T(3) * am));
                V += kn * ap - ks * am;
            }
            {   // z: the minus face is the previous plane's plus face
                const T ap = qu[d][2] - ucc, bp = qu[d][3] - qu[d][0];
                const T pp = T(27) * ap - bp, tp = bp - T(3) * ap;
                F += WT * pp + WB * pzm[d];
                D += auz * (tp - tzm[d]);
                V += kt_ * ap - kb * azm[d];
                pzm[d] = pp;
                tzm[d] = tp;
                azm[d] = ap;
            }
            const T adv = F * rj + D;
            if (act) st_at(ua.c[d], cb, ucc + dt * (V - adv + cff[d]));
        }
        {   // plane k+1 into the other LDS buffer; rotate the queues and stages
            lds_fill(b ^ 1, phv[0], qu[0][2], qu[1][2], qu[2][2], qn[2]);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                qu[c][0] = qu[c][1];
                qu[c][1] = qu[c][2];
                qu[c][2] = qu[c][3];
                qu[c][3] = pu[0][c];
            }
            qn[0] = qn[1];
            qn[1] = qn[2];
            qn[2] = pn[0];
            wb = WT;
#pragma unroll
            for (int d = 0; d < 5; d++) cuu[d] = puu[0][d];
#pragma unroll
            for (int d = 0; d < 3; d++) cff[d] = pff[0][d];
#pragma unroll
            for (int st = 0; st + 1 < PD; st++) {
#pragma unroll
                for (int c = 0; c < 3; c++) pu[st][c] = pu[st + 1][c];
                pn[st] = pn[st + 1];
#pragma unroll
                for (int d = 0; d < 5; d++) puu[st][d] = puu[st + 1][d];
#pragma unroll
                for (int d = 0; d < 3; d++) pff[st][d] = pff[st + 1][d];
#pragma unroll
                for (int q = 0; q < KM_HPT; q++) phv[st][q] = phv[st + 1][q];
            }
        }
        cb += szb;
        __syncthreads();
    }
}

// UtoCU + InterpolateCU fused (src/dev/FalmCFDDevCall.cu:303-349): the
// face flux of face (i+1/2) is the mean of the two cell-centre fluxes
// ja*kx*u, each evaluated exactly as the reference does.
template <class T, class M>
__global__ __launch_bounds__(256) void k_u_to_uu(Tiling tl, Geom g, Vec3<T> u, Vec3W<T> uu, M met) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const int nb[3] = {1, g.nx, g.nxy};
        const T jc = met.JA(cc, i, j, k);
    #pragma unroll
        for (int d = 0; d < 3; d++) {
            const int cn = cc + nb[d];
            const int in = i + (d == 0), jn = j + (d == 1), kn = k + (d == 2);
            const T ucc = jc * met.KX(d, cc, i, j, k) * u.c[d][cc];
            const T ucn = met.JA(cn, in, jn, kn) * met.KX(d, cn, in, jn, kn) * u.c[d][cn];
            uu.c[d][cc] = T(0.5) * (ucc + ucn);
        }
    });
}

// Divergence (src/dev/FalmCFDDevCall.cu:557-581), optionally followed by the
// ScaleMatrix pass of MACCalcPoissonRHS (src/FalmCFD.h:96-109): the launch
// covers every padded cell; interior cells get div*scale, the others are
// multiplied by scale as ScaleMatrix does over the whole array.
template <class T, class TO, bool SCALE, class M>
__global__ __launch_bounds__(256) void k_divergence(Tiling tl, Geom g, Vec3<T> uu, TO *__restrict__ div, M met,
                                                    double scale) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const bool inner = i >= g.gc && i < g.nx - g.gc && j >= g.gc && j < g.ny - g.gc && k >= g.gc &&
                           k < g.nz - g.gc;
        if (inner) {
            TO UE = uu.c[0][cc], UW = uu.c[0][cc - 1];
            TO VN = uu.c[1][cc], VS = uu.c[1][cc - g.nx];
            TO WT = uu.c[2][cc], WB = uu.c[2][cc - g.nxy];
            TO d = met.div_ja(UE - UW + VN - VS + WT - WB, cc, i, j, k);
            if (SCALE) d *= (TO)scale;
            div[cc] = d;
        } else if (SCALE) {
            div[cc] *= (TO)scale;
        }
    });
}

// ProjectPGrid + ProjectPFace (src/dev/FalmCFDDevCall.cu:351-407) in one
// launch over the face region (which contains the interior): cell update on
// interior cells, face update on every face-region cell.  u may alias ua and
// uu may alias uua.  The pressure may be kept in a wider type P.
template <class T, class P, class M>
__global__ __launch_bounds__(256) void k_project(Tiling tl, Geom g, P dt, T *u0, T *u1, T *u2, const T *ua0,
                                                 const T *ua1, const T *ua2, T *uu0, T *uu1, T *uu2,
                                                 const T *uua0, const T *uua1, const T *uua2,
                                                 const P *__restrict__ p, M met) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int sy = g.nx, sz = g.nxy;
        const int cc = g.id(i, j, k);
        const P pcc = p[cc];
        if (i >= g.gc && j >= g.gc && k >= g.gc) {  // face region upper bound == interior upper bound
            P dpx = P(0.5) * (P)met.KX(0, cc, i, j, k) * (p[cc + 1] - p[cc - 1]);
            P dpy = P(0.5) * (P)met.KX(1, cc, i, j, k) * (p[cc + sy] - p[cc - sy]);
            P dpz = P(0.5) * (P)met.KX(2, cc, i, j, k) * (p[cc + sz] - p[cc - sz]);
            T a0 = ua0[cc], a1 = ua1[cc], a2 = ua2[cc];
            u0[cc] = (T)((P)a0 - dt * dpx);
            u1[cc] = (T)((P)a1 - dt * dpy);
            u2[cc] = (T)((P)a2 - dt * dpz);
        }
        P dfx = P(0.5) * ((P)met.G(0, cc, i, j, k) + (P)met.G(0, cc + 1, i + 1, j, k)) * (p[cc + 1] - pcc);
        P dfy = P(0.5) * ((P)met.G(1, cc, i, j, k) + (P)met.G(1, cc + sy, i, j + 1, k)) * (p[cc + sy] - pcc);
        P dfz = P(0.5) * ((P)met.G(2, cc, i, j, k) + (P)met.G(2, cc + sz, i, j, k + 1)) * (p[cc + sz] - pcc);
        T b0 = uua0[cc], b1 = uua1[cc], b2 = uua2[cc];
        uu0[cc] = (T)((P)b0 - dt * dfx);
        uu1[cc] = (T)((P)b1 - dt * dfy);
        uu2[cc] = (T)((P)b2 - dt * dfz);
    });
}

// kernel_Cartesian_Smagorinsky, src/dev/FalmCFDDevCall.cu:409-476
template <class T, class M>
__global__ __launch_bounds__(256) void k_smagorinsky(Tiling tl, Geom g, T Cs, Vec3<T> u, T *__restrict__ nut,
                                                     M met) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int sy = g.nx, sz = g.nxy;
        const int cc = g.id(i, j, k);
        const T kxx = met.KX(0, cc, i, j, k), kxy = met.KX(1, cc, i, j, k), kxz = met.KX(2, cc, i, j, k);
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
        T De = cbrt(met.JA(cc, i, j, k));
        T lc = Cs * De;
        nut[cc] = lc * lc * Du;
    });
}

// ------------------------------------------- k-marching forms (tables)
// Projection on the interior cells: cell update and face update, the
// pressure walked as a Plane7, the metric factors from the per-axis tables
// (in the pressure's precision P):
//   u_d  -= dt kx_d (p(c+e) - p(c-e))/2                (ProjectPGrid)
//   uu_d -= dt (g_d(c) + g_d(c+e))/2 (p(c+e) - p(c))   (ProjectPFace)
// with kx_d = r_d and (g_d(c) + g_d(c+e))/2 = (product of the other two
// widths) x (r_d(c) + r_d(c+e))/2 on the tensor-product mesh
// (src/dev/FalmCFDDevCall.cu:351-407).
template <class T, class P>
__global__ __launch_bounds__(K7_NT) void k_project_km(K7Tiling kt, Geom g, P dt, Vec3W<T> u, Vec3W<T> uu,
                                                      const P *__restrict__ p, SepTab<P> tb, const int *gate) {
    if (gated_off(gate)) return;
    __shared__ K7Lds<P> lp;
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    const unsigned szt = (unsigned)sz * sizeof(T);
    Plane7<P> pp;
    pp.setup(p, &lp, Q, g);
    pp.prologue(Q, sz);
    __syncthreads();
    const int i = Q.act ? Q.i : kt.box.lo[0], j = Q.act ? Q.j : kt.box.lo[1];
    const P cdx = dt * P(0.5) * tb.t[SEP_R][0][i], cdy = dt * P(0.5) * tb.t[SEP_R][1][j];
    const P ffx = dt * tb.t[SEP_RF][0][i] * tb.t[SEP_H][1][j];
    const P ffy = dt * tb.t[SEP_RF][1][j] * tb.t[SEP_H][0][i];
    const P ffz = dt * tb.t[SEP_H][0][i] * tb.t[SEP_H][1][j];
    unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szt;
    int b = 0;
    for (int k = Q.k0; k < Q.k1; k++, cb += szt) {
        const bool more = k + 1 < Q.k1;
        if (more) pp.prefetch(Q, k, sz);
        const P hz = ld_uniform(tb.t[SEP_H][2], k);
        const P cdz = dt * P(0.5) * ld_uniform(tb.t[SEP_R][2], k), fz = ffz * ld_uniform(tb.t[SEP_RF][2], k);
        if (Q.act) {
            const P pcc = pp.c();
            const T a0 = ld_at((const T *)u.c[0], cb), a1 = ld_at((const T *)u.c[1], cb),
                    a2 = ld_at((const T *)u.c[2], cb);
            const T b0 = ld_at((const T *)uu.c[0], cb), b1 = ld_at((const T *)uu.c[1], cb),
                    b2 = ld_at((const T *)uu.c[2], cb);
            st_at(u.c[0], cb, (T)((P)a0 - cdx * (pp.xp(b) - pp.xm(b))));
            st_at(u.c[1], cb, (T)((P)a1 - cdy * (pp.yp(b) - pp.ym(b))));
            st_at(u.c[2], cb, (T)((P)a2 - cdz * (pp.zp() - pp.zm())));
            st_at(uu.c[0], cb, (T)((P)b0 - (ffx * hz) * (pp.xp(b) - pcc)));
            st_at(uu.c[1], cb, (T)((P)b1 - (ffy * hz) * (pp.yp(b) - pcc)));
            st_at(uu.c[2], cb, (T)((P)b2 - fz * (pp.zp() - pcc)));
        }
        pp.advance(b, more);
        b ^= 1;
        __syncthreads();
    }
}

// The face-region cells outside the interior: the layer at index gc-1 of
// each axis (three disjoint slabs), face update only.
__device__ inline bool face_layer_cell(const Geom &g, long long t, int &i, int &j, int &k) {
    const int gc = g.gc, lo = gc - 1;
    const int fx = g.nx - 2 * gc + 1, fy = g.ny - 2 * gc + 1, fz = g.nz - 2 * gc + 1;  // face box sizes
    const long long sx = (long long)fy * fz, sy = (long long)(fx - 1) * fz, sz = (long long)(fx - 1) * (fy - 1);
    if (t < sx) {  // i = gc-1, all j, k of the face box
        i = lo;
        j = lo + (int)(t % fy);
        k = lo + (int)(t / fy);
        return true;
    }
    t -= sx;
    if (t < sy) {  // j = gc-1, i >= gc
        i = gc + (int)(t % (fx - 1));
        j = lo;
        k = lo + (int)(t / (fx - 1));
        return true;
    }
    t -= sy;
    if (t < sz) {  // k = gc-1, i, j >= gc
        i = gc + (int)(t % (fx - 1));
        j = gc + (int)(t / (fx - 1));
        k = lo;
        return true;
    }
    return false;
}
inline long long face_layer_count(const Geom &g) {
    const int gc = g.gc;
    const long long fx = g.nx - 2 * gc + 1, fy = g.ny - 2 * gc + 1, fz = g.nz - 2 * gc + 1;
    return fy * fz + (fx - 1) * fz + (fx - 1) * (fy - 1);
}

template <class T, class P>
__global__ __launch_bounds__(256) void k_project_layer(Geom g, P dt, Vec3W<T> uu, const P *__restrict__ p,
                                                       SepTab<P> tb, const int *gate) {
    if (gated_off(gate)) return;
    int i, j, k;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!face_layer_cell(g, t, i, j, k)) return;
    const int sy = g.nx, sz = g.nxy;
    const int cc = g.id(i, j, k);
    const P pcc = p[cc];
    const P hx = tb.t[SEP_H][0][i], hy = tb.t[SEP_H][1][j], hz = tb.t[SEP_H][2][k];
    const P fx = dt * tb.t[SEP_RF][0][i] * hy * hz;
    const P fy = dt * tb.t[SEP_RF][1][j] * hx * hz;
    const P fz = dt * hx * hy * tb.t[SEP_RF][2][k];
    const T b0 = uu.c[0][cc], b1 = uu.c[1][cc], b2 = uu.c[2][cc];
    uu.c[0][cc] = (T)((P)b0 - fx * (p[cc + 1] - pcc));
    uu.c[1][cc] = (T)((P)b1 - fy * (p[cc + sy] - pcc));
    uu.c[2][cc] = (T)((P)b2 - fz * (p[cc + sz] - pcc));
}

// Face flux on the interior cells (UtoCU + InterpolateCU,
// src/dev/FalmCFDDevCall.cu:303-349): uu_d = (ja kx_d u_d)(c)/2 +
// (ja kx_d u_d)(c+e_d)/2, where ja kx_d is the product of the other two
// widths and so the same at c and c+e_d.  Each component needs one
// neighbour only: u_x(c+1) and u_y(c+sy) are loaded directly (the same or
// the next cache lines of the wavefront's row), u_z(c+sz) is the next
// plane's value of the column, carried in a register.  No LDS.
template <class T>
__global__ __launch_bounds__(K7_NT) void k_u_to_uu_km(K7Tiling kt, Geom g, Vec3<T> u, Vec3W<T> uu, SepTab<T> tb) {
    const K7Pos Q = k7_pos(kt, g);
    if (!Q.act) return;  // no LDS, no barriers: inactive threads just leave
    const int sz = g.nxy;
    const unsigned szb = (unsigned)sz * sizeof(T), syb = (unsigned)g.nx * sizeof(T);
    const T hx2 = T(0.5) * tb.t[SEP_H][0][Q.i], hy2 = T(0.5) * tb.t[SEP_H][1][Q.j];
    const T hxy2 = hx2 * tb.t[SEP_H][1][Q.j];
    unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szb;
    T wz = ld_at(u.c[2], cb);
    for (int k = Q.k0; k < Q.k1; k++, cb += szb) {
        const T ux = ld_at(u.c[0], cb), uxe = ld_at(u.c[0], cb + (unsigned)sizeof(T));
        const T vy = ld_at(u.c[1], cb), vyn = ld_at(u.c[1], cb + syb);
        const T wzt = ld_at(u.c[2], cb + szb);
        const T hz = ld_uniform(tb.t[SEP_H][2], k);
        st_at(uu.c[0], cb, (hy2 * hz) * (ux + uxe));
        st_at(uu.c[1], cb, (hx2 * hz) * (vy + vyn));
        st_at(uu.c[2], cb, hxy2 * (wz + wzt));
        wz = wzt;
    }
}

// k_u_to_uu_km followed by the Poisson right-hand side of the interior
// (MACCalcPoissonRHS, src/FalmCFD.h:96-109) in one pass.  Besides its own
// upper faces each cell recomputes its lower faces from the same operands,
// in the same order, as the cell (or the ghost-face layer) that stores them,
// so rhs equals the divergence of the stored uu bit for bit without reading
// uu back.  Used where no wall-flux correction runs between the two.
// Ghost cells are read from gu: the step driver passes the previous step's
// velocity, whose ghost cells are the ones the in-place update of the
// reference leaves in u until the boundary fill (its interior is never
// read here).
template <class T>
__global__ __launch_bounds__(K7_NT) void k_u_to_uu_rhs_km(K7Tiling kt, Geom g, Vec3<T> u, Vec3<T> gu, Vec3W<T> uu,
                                                          SepTab<T> tb, double *__restrict__ rhs, double scale) {
    const K7Pos Q = k7_pos(kt, g);
    if (!Q.act) return;
    const int sz = g.nxy;
    const unsigned szb = (unsigned)sz * sizeof(T), syb = (unsigned)g.nx * sizeof(T), xb = (unsigned)sizeof(T);
    const T hx2 = T(0.5) * tb.t[SEP_H][0][Q.i], hy2 = T(0.5) * tb.t[SEP_H][1][Q.j];
    const T hxy2 = hx2 * tb.t[SEP_H][1][Q.j];
    const T rxy = tb.t[SEP_R][0][Q.i] * tb.t[SEP_R][1][Q.j];
    // neighbours in the ghost layers (the interior box is at least one cell
    // from the array's edge)
    const bool gw = Q.i - 1 < g.gc, ge = Q.i + 1 >= g.nx - g.gc;
    const bool gs = Q.j - 1 < g.gc, gn = Q.j + 1 >= g.ny - g.gc;
    const int ktop = g.nz - g.gc;  // first ghost plane above
    unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szb;
    unsigned rb = ((unsigned)Q.col0 + (unsigned)Q.k0 * (unsigned)sz) * (unsigned)sizeof(double);
    const unsigned rzb = (unsigned)sz * (unsigned)sizeof(double);
    T wz = ld_at(u.c[2], cb);
    const T wbelow = Q.k0 - 1 < g.gc ? ld_at(gu.c[2], cb - szb) : ld_at(u.c[2], cb - szb);
    T fzb = hxy2 * (wbelow + wz);  // the face below plane k0
    for (int k = Q.k0; k < Q.k1; k++, cb += szb, rb += rzb) {
        const T ux = ld_at(u.c[0], cb), vy = ld_at(u.c[1], cb);
        T uxe = ld_at(u.c[0], cb + xb), uxw = ld_at(u.c[0], cb - xb);
        T vyn = ld_at(u.c[1], cb + syb), vys = ld_at(u.c[1], cb - syb);
        if (ge) uxe = ld_at(gu.c[0], cb + xb);
        if (gw) uxw = ld_at(gu.c[0], cb - xb);
        if (gn) vyn = ld_at(gu.c[1], cb + syb);
        if (gs) vys = ld_at(gu.c[1], cb - syb);
        const T wzt = k + 1 >= ktop ? ld_at(gu.c[2], cb + szb) : ld_at(u.c[2], cb + szb);
        const T hz = ld_uniform(tb.t[SEP_H][2], k);
        const T fxe = (hy2 * hz) * (ux + uxe), fxw = (hy2 * hz) * (uxw + ux);
        const T fyn = (hx2 * hz) * (vy + vyn), fys = (hx2 * hz) * (vys + vy);
        const T fzt = hxy2 * (wz + wzt);
        st_at(uu.c[0], cb, fxe);
        st_at(uu.c[1], cb, fyn);
        st_at(uu.c[2], cb, fzt);
        double d = (double)fxe - (double)fxw + (double)fyn - (double)fys + (double)fzt - (double)fzb;
        d = d * (double)(rxy * ld_uniform(tb.t[SEP_R][2], k));
        st_at(rhs, rb, d * scale);
        fzb = fzt;
        wz = wzt;
    }
}

// gu: where ghost cells are read (u itself, or the previous step's
// velocity as in k_u_to_uu_rhs_km)
template <class T>
__global__ __launch_bounds__(256) void k_u_to_uu_layer(Geom g, Vec3<T> u, Vec3<T> gu, Vec3W<T> uu, SepTab<T> tb) {
    int i, j, k;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!face_layer_cell(g, t, i, j, k)) return;
    const int cc = g.id(i, j, k);
    const T hx = tb.t[SEP_H][0][i], hy = tb.t[SEP_H][1][j], hz = tb.t[SEP_H][2][k];
    auto at = [&](int d, int ii, int jj, int kk, int c) {
        const bool ghost = ii < g.gc || ii >= g.nx - g.gc || jj < g.gc || jj >= g.ny - g.gc || kk < g.gc ||
                           kk >= g.nz - g.gc;
        return ghost ? gu.c[d][c] : u.c[d][c];
    };
    uu.c[0][cc] = (T(0.5) * hy * hz) * (at(0, i, j, k, cc) + at(0, i + 1, j, k, cc + 1));
    uu.c[1][cc] = (T(0.5) * hx * hz) * (at(1, i, j, k, cc) + at(1, i, j + 1, k, cc + g.nx));
    uu.c[2][cc] = (T(0.5) * hx * hy) * (at(2, i, j, k, cc) + at(2, i, j, k + 1, cc + g.nxy));
}

// Smagorinsky on the interior (kernel_Cartesian_Smagorinsky,
// src/dev/FalmCFDDevCall.cu:409-476): nut = (Cs ja^(1/3))^2 |S| with
// ja^(2/3) the product of the per-axis cbrt(h)^2 table entries.
template <class T>
__global__ __launch_bounds__(K7_NT) void k_smagorinsky_km(K7Tiling kt, Geom g, T Cs, Vec3<T> u, T *__restrict__ nut,
                                                          SepTab<T> tb) {
    __shared__ K7Lds<T> l0, l1, l2;
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    const unsigned szb = (unsigned)sz * sizeof(T);
    Plane7<T> pu[3];
    pu[0].setup(u.c[0], &l0, Q, g);
    pu[1].setup(u.c[1], &l1, Q, g);
    pu[2].setup(u.c[2], &l2, Q, g);
#pragma unroll
    for (int d = 0; d < 3; d++) pu[d].prologue(Q, sz);
    __syncthreads();
    const int i = Q.act ? Q.i : kt.box.lo[0], j = Q.act ? Q.j : kt.box.lo[1];
    const T hrx = T(0.5) * tb.t[SEP_R][0][i], hry = T(0.5) * tb.t[SEP_R][1][j];
    const T c2 = Cs * Cs * tb.t[SEP_H23][0][i] * tb.t[SEP_H23][1][j];
    unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szb;
    int b = 0;
    for (int k = Q.k0; k < Q.k1; k++, cb += szb) {
        const bool more = k + 1 < Q.k1;
        if (more)
#pragma unroll
            for (int d = 0; d < 3; d++) pu[d].prefetch(Q, k, sz);
        const T hrz = T(0.5) * ld_uniform(tb.t[SEP_R][2], k), h23z = ld_uniform(tb.t[SEP_H23][2], k);
        if (Q.act) {
            T gr[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gr[d][0] = hrx * (pu[d].xp(b) - pu[d].xm(b));
                gr[d][1] = hry * (pu[d].yp(b) - pu[d].ym(b));
                gr[d][2] = hrz * (pu[d].zp() - pu[d].zm());
            }
            const T dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
            const T dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
            const T dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
            T d1 = T(2) * dux * dux, d2 = T(2) * dvy * dvy, d3 = T(2) * dwz * dwz;
            T d4 = (duy + dvx) * (duy + dvx), d5 = (dvz + dwy) * (dvz + dwy), d6 = (duz + dwx) * (duz + dwx);
            const T Du = sqrt(d1 + d2 + d3 + d4 + d5 + d6);
            st_at(nut, cb, (c2 * h23z) * Du);
        }
#pragma unroll
        for (int d = 0; d < 3; d++) pu[d].advance(b, more);
        b ^= 1;
        __syncthreads();
    }
}

// |div|^2 of the step monitor fused into the divergence pass: the interior
// tiles are walked in the chunked order of the reduction kernels, so the
// partial sums equal those of a separate dot product over the stored field.
template <class T, class M>
__global__ __launch_bounds__(256) void k_div_norm(Tiling tl, Geom g, Vec3<T> uu, T *__restrict__ div, M met,
                                                  double *partials) {
    __shared__ double lds[256 / 64];
    int i, j, k, t0, t1;
    double v[1] = {0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        T UE = uu.c[0][cc], UW = uu.c[0][cc - 1];
        T VN = uu.c[1][cc], VS = uu.c[1][cc - g.nx];
        T WT = uu.c[2][cc], WB = uu.c[2][cc - g.nxy];
        T d = met.div_ja(UE - UW + VN - VS + WT - WB, cc, i, j, k);
        div[cc] = d;
        v[0] += (double)d * (double)d;
    }
    block_partials<256, 1>(v, partials, lds);
}

// The same on K7 column tiles (step driver): uu_x(c-1) and uu_y(c-sy) are
// loaded directly, uu_z(c-sz) is the previous plane's value of the column;
// one partial per workgroup.
template <class T>
__global__ __launch_bounds__(K7_NT) void k_div_norm_km(K7Tiling kt, Geom g, Vec3<T> uu, T *__restrict__ div,
                                                       SepTab<T> tb, double *partials, const int *gate) {
    if (gated_off(gate)) return;
    __shared__ double lds[K7_NT / 64];
    const K7Pos Q = k7_pos(kt, g);
    double v[1] = {0.0};
    if (Q.act) {
        const int sz = g.nxy;
        const unsigned szb = (unsigned)sz * sizeof(T), syb = (unsigned)g.nx * sizeof(T);
        const T rxy = tb.t[SEP_R][0][Q.i] * tb.t[SEP_R][1][Q.j];
        unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szb;
        T wb = ld_at(uu.c[2], cb - szb);
        for (int k = Q.k0; k < Q.k1; k++, cb += szb) {
            const T UE = ld_at(uu.c[0], cb), UW = ld_at(uu.c[0], cb - (unsigned)sizeof(T));
            const T VN = ld_at(uu.c[1], cb), VS = ld_at(uu.c[1], cb - syb);
            const T WT = ld_at(uu.c[2], cb);
            const T d = (UE - UW + VN - VS + WT - wb) * (rxy * ld_uniform(tb.t[SEP_R][2], k));
            st_at(div, cb, d);
            v[0] += (double)d * (double)d;
            wb = WT;
        }
    }
    block_partials<K7_NT, 1>(v, partials, lds);
}

// k_smagorinsky_km and k_div_norm_km in one pass over the same column
// tiles (the step's last two stencils read u and uu, both final by then):
// the same arithmetic for nut and div, the same partial per workgroup, and
// the face fluxes loaded a plane ahead like the velocity.
template <class T>
__global__ __launch_bounds__(K7_NT) void k_sgs_div_km(K7Tiling kt, Geom g, T Cs, Vec3<T> u, T *__restrict__ nut,
                                                      Vec3<T> uu, T *__restrict__ div, SepTab<T> tb,
                                                      double *partials, const int *gate) {
    if (gated_off(gate)) return;
    __shared__ K7Lds<T> l0, l1, l2;
    __shared__ double lds[K7_NT / 64];
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    const unsigned szb = (unsigned)sz * sizeof(T), syb = (unsigned)g.nx * sizeof(T);
    Plane7<T> pu[3];
    pu[0].setup(u.c[0], &l0, Q, g);
    pu[1].setup(u.c[1], &l1, Q, g);
    pu[2].setup(u.c[2], &l2, Q, g);
#pragma unroll
    for (int d = 0; d < 3; d++) pu[d].prologue(Q, sz);
    const int i = Q.act ? Q.i : kt.box.lo[0], j = Q.act ? Q.j : kt.box.lo[1];
    const T hrx = T(0.5) * tb.t[SEP_R][0][i], hry = T(0.5) * tb.t[SEP_R][1][j];
    const T c2 = Cs * Cs * tb.t[SEP_H23][0][i] * tb.t[SEP_H23][1][j];
    const T rxy = tb.t[SEP_R][0][i] * tb.t[SEP_R][1][j];
    unsigned cb = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szb;
    // face fluxes of plane k0 (east, west, north, south, top) and the face below
    T wb = ld_at(uu.c[2], cb - szb);
    T fe = ld_at(uu.c[0], cb), fw = ld_at(uu.c[0], cb - (unsigned)sizeof(T));
    T fn = ld_at(uu.c[1], cb), fs = ld_at(uu.c[1], cb - syb), ft = ld_at(uu.c[2], cb);
    double v[1] = {0.0};
    __syncthreads();
    int b = 0;
    for (int k = Q.k0; k < Q.k1; k++, cb += szb) {
        const bool more = k + 1 < Q.k1;
        T ne = T(0), nw = T(0), nn = T(0), ns = T(0), nt = T(0);
        if (more) {
#pragma unroll
            for (int d = 0; d < 3; d++) pu[d].prefetch(Q, k, sz);
            const unsigned c1 = cb + szb;
            ne = ld_at(uu.c[0], c1);
            nw = ld_at(uu.c[0], c1 - (unsigned)sizeof(T));
            nn = ld_at(uu.c[1], c1);
            ns = ld_at(uu.c[1], c1 - syb);
            nt = ld_at(uu.c[2], c1);
        }
        const T hrz = T(0.5) * ld_uniform(tb.t[SEP_R][2], k), h23z = ld_uniform(tb.t[SEP_H23][2], k);
        if (Q.act) {
            T gr[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gr[d][0] = hrx * (pu[d].xp(b) - pu[d].xm(b));
                gr[d][1] = hry * (pu[d].yp(b) - pu[d].ym(b));
                gr[d][2] = hrz * (pu[d].zp() - pu[d].zm());
            }
            const T dux = gr[0][0], duy = gr[0][1], duz = gr[0][2];
            const T dvx = gr[1][0], dvy = gr[1][1], dvz = gr[1][2];
            const T dwx = gr[2][0], dwy = gr[2][1], dwz = gr[2][2];
            T d1 = T(2) * dux * dux, d2 = T(2) * dvy * dvy, d3 = T(2) * dwz * dwz;
            T d4 = (duy + dvx) * (duy + dvx), d5 = (dvz + dwy) * (dvz + dwy), d6 = (duz + dwx) * (duz + dwx);
            const T Du = sqrt(d1 + d2 + d3 + d4 + d5 + d6);
            st_at(nut, cb, (c2 * h23z) * Du);
            const T dv = (fe - fw + fn - fs + ft - wb) * (rxy * ld_uniform(tb.t[SEP_R][2], k));
            st_at(div, cb, dv);
            v[0] += (double)dv * (double)dv;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) pu[d].advance(b, more);
        wb = ft;
        fe = ne;
        fw = nw;
        fn = nn;
        fs = ns;
        ft = nt;
        b ^= 1;
        __syncthreads();
    }
    block_partials<K7_NT, 1>(v, partials, lds);
}

// hout (may be null): a pinned host copy of the result
__global__ __launch_bounds__(FIN_NT) void k_fin_norm(const double *partials, int nb, double *out, const int *gate,
                                                     double *hout) {
    if (gated_off(gate)) return;
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) {
        out[0] = tot[0];
        if (hout) hout[0] = tot[0];
    }
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
template <class T>
static MetField<T> met_field(const void *kx, const void *gg, const void *ja, size_t n) {
    MetField<T> m;
    for (int d = 0; d < 3; d++) {
        m.kx[d] = kx ? col<T>(kx, n, d) : nullptr;
        m.g[d] = gg ? col<T>(gg, n, d) : nullptr;
    }
    m.ja = (const T *)ja;
    return m;
}
template <class T>
static SepTab<T> sep_tab(const SepMesh &sm);
template <>
SepTab<double> sep_tab<double>(const SepMesh &sm) {
    SepTab<double> t;
    for (int q = 0; q < SEP_NTAB; q++)
        for (int d = 0; d < 3; d++) t.t[q][d] = sm.td[q][d];
    return t;
}
template <>
SepTab<float> sep_tab<float>(const SepMesh &sm) {
    SepTab<float> t;
    for (int q = 0; q < SEP_NTAB; q++)
        for (int d = 0; d < 3; d++) t.t[q][d] = sm.tf[q][d];
    return t;
}

void sep_tables(const double *h, int len, std::vector<double> (&out)[SEP_NTAB]) {
    for (auto &v : out) v.assign(len, 0.0);
    for (int i = 0; i < len; i++) {
        const double r = 1.0 / h[i];
        const double rp = 1.0 / h[i + 1 < len ? i + 1 : i], rm = 1.0 / h[i > 0 ? i - 1 : i];
        const double c = cbrt(h[i]);
        out[SEP_R][i] = r;
        out[SEP_H][i] = h[i];
        out[SEP_CP][i] = 0.5 * (r + rp) * r;
        out[SEP_CM][i] = 0.5 * (r + rm) * r;
        out[SEP_RF][i] = 0.5 * (r + rp);
        out[SEP_H23][i] = c * c;
    }
}

template <class T>
static MetTab<T> met_tab(const SepMesh &sm) {
    return MetTab<T>{sep_tab<T>(sm)};
}

template <class T>
static MetSep<T> met_sep(const SepMesh &sm) {
    MetSep<T> m;
    for (int d = 0; d < 3; d++) {
        m.h[d] = sm.h[d];
        m.rh[d] = sm.rh[d];
    }
    return m;
}

static bool overlaps(const void *a, size_t abytes, const void *b, size_t bbytes) {
    const char *x = (const char *)a, *y = (const char *)b;
    return x < y + bbytes && y < x + abytes;
}

template <class T, class M>
static falm_status pseudo_u_t(const falm_domain *dom, int scheme, double ReI, double dt, const void *un,
                              const void *u, const void *uu, void *ua, const void *nut, const M &met,
                              const void *ff, const int *ffbox, double words, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_PSEUDO_U, s);
#define FALM_PU(S)                                                                                      \
    hipLaunchKernelGGL((k_pseudo_u<S, T, M>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (T)ReI, (T)dt, \
                       v3<T>(un, n), v3<T>(u, n), v3<T>(uu, n), v3w<T>(ua, n), (const T *)nut, met, v3<T>(ff, n), \
                       ffbox)
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
    tm.done((double)tl.box.size() * words * sizeof(T));
    return FALM_OK;
}

template <class T, class M>
static falm_status u_to_uu_t(const falm_domain *dom, const void *u, void *uu, const M &met, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(face_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_FACE_FLUX, s);
    hipLaunchKernelGGL((k_u_to_uu<T, M>), tiling_grid(tl), tiling_block(), 0, s, tl, g, v3<T>(u, n), v3w<T>(uu, n),
                       met);
    FALM_CHECK_LAUNCH("k_u_to_uu");
    // u read, uu written (+ kx, ja)
    tm.done((double)tl.box.size() * (6.0 + (M::per_cell ? 4.0 : 0.0)) * sizeof(T));
    return FALM_OK;
}

// whole_array: the ScaleMatrix pass of MACCalcPoissonRHS covers every padded
// cell; the step driver keeps ghost rhs at zero and passes false
template <class T, class TO, bool SCALE, class M>
static falm_status divergence_t(const falm_domain *dom, const void *uu, void *div, const M &met, double scale,
                                bool whole_array, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Box b = (SCALE && whole_array) ? Box{{0, 0, 0}, {g.nx, g.ny, g.nz}} : interior_box(g);
    Tiling tl = make_tiling(b);
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(SCALE ? FALM_TIME_RHS : FALM_TIME_DIVERGENCE, s);
    hipLaunchKernelGGL((k_divergence<T, TO, SCALE, M>), tiling_grid(tl), tiling_block(), 0, s, tl, g, v3<T>(uu, n),
                       (TO *)div, met, scale);
    FALM_CHECK_LAUNCH("k_divergence");
    // uu read, div written (+ ja)
    tm.done((double)interior_box(g).size() * ((3.0 + (M::per_cell ? 1.0 : 0.0)) * sizeof(T) + sizeof(TO)));
    return FALM_OK;
}

template <class T, class P, class M>
static falm_status project_t(const falm_domain *dom, double dt, void *u, const void *ua, void *uu,
                             const void *uua, const void *p, const M &met, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(face_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_PROJECT, s);
    hipLaunchKernelGGL((k_project<T, P, M>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (P)dt, col<T>(u, n, 0),
                       col<T>(u, n, 1), col<T>(u, n, 2), col<T>(ua, n, 0), col<T>(ua, n, 1), col<T>(ua, n, 2),
                       col<T>(uu, n, 0), col<T>(uu, n, 1), col<T>(uu, n, 2), col<T>(uua, n, 0), col<T>(uua, n, 1),
                       col<T>(uua, n, 2), (const P *)p, met);
    FALM_CHECK_LAUNCH("k_project");
    // ua, uua read, u, uu written, p read (+ kx, g)
    tm.done((double)tl.box.size() * ((12.0 + (M::per_cell ? 6.0 : 0.0)) * sizeof(T) + sizeof(P)));
    return FALM_OK;
}

template <class T, class M>
static falm_status smag_t(const falm_domain *dom, double Cs, const void *u, void *nut, const M &met,
                          hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_SGS, s);
    hipLaunchKernelGGL((k_smagorinsky<T, M>), tiling_grid(tl), tiling_block(), 0, s, tl, g, (T)Cs, v3<T>(u, n),
                       (T *)nut, met);
    FALM_CHECK_LAUNCH("k_smagorinsky");
    // u read, nut written (+ kx, ja)
    tm.done((double)tl.box.size() * (4.0 + (M::per_cell ? 4.0 : 0.0)) * sizeof(T));
    return FALM_OK;
}

template <class T>
static falm_status div_norm_t(const falm_domain *dom, const void *uu, void *div, const SepMesh &sm,
                              double *dev_out, double *partials, hipStream_t s, const int *gate, double *host_out) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    FALM_TRY(check_byte_offsets(n, sizeof(T)));
    TimedLaunch tm(FALM_TIME_DIVERGENCE, s);
    hipLaunchKernelGGL((k_div_norm_km<T>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, v3<T>(uu, n), (T *)div,
                       sep_tab<T>(sm), partials, gate);
    tm.done((double)interior_box(g).size() * 4.0 * sizeof(T));  // uu read, div written
    hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(FIN_NT), 0, s, (const double *)partials, kt.blocks, dev_out, gate,
                       host_out);
    FALM_CHECK_LAUNCH("k_div_norm");
    return FALM_OK;
}

// ---------------------------------------------- step-driver entry points
// Tensor-product mesh metrics (MetSep); CFD fields in prec, pressure and
// Poisson right-hand side in double.
template <class T>
static falm_status pseudo_u_km_t(const falm_domain *dom, int scheme, double ReI, double dt, const void *u,
                                 const void *uu, void *ua, const void *nut, const SepMesh &sm, const void *ff,
                                 const int *ffbox, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    KmTiling kt = make_km_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    FALM_TRY(check_byte_offsets(n, sizeof(T)));
    TimedLaunch tm(FALM_TIME_PSEUDO_U, s);
#define FALM_PUK(S)                                                                                        \
    hipLaunchKernelGGL((k_pseudo_u_km<S, T>), dim3(kt.blocks), dim3(KM_NT), 0, s, kt, g, (T)ReI, (T)dt,   \
                       v3<T>(u, n), v3<T>(uu, n), v3w<T>(ua, n), (const T *)nut, met_sep<T>(sm), v3<T>(ff, n), \
                       ffbox)
#define FALM_PU3(S)                                                                                        \
    hipLaunchKernelGGL((k_pseudo_u3_km<S, T>), dim3(kt.blocks), dim3(KM_NT), 0, s, kt, g, (T)ReI, (T)dt,  \
                       v3<T>(u, n), v3<T>(uu, n), v3w<T>(ua, n), (const T *)nut, sep_tab<T>(sm), v3<T>(ff, n), \
                       ffbox)
    switch (scheme) {
    case FALM_ADV_UPWIND1: FALM_PUK(FALM_ADV_UPWIND1); break;
    case FALM_ADV_UPWIND3: FALM_PU3(FALM_ADV_UPWIND3); break;
    case FALM_ADV_QUICK: FALM_PUK(FALM_ADV_QUICK); break;
    case FALM_ADV_UTOPIA: FALM_PU3(FALM_ADV_UTOPIA); break;
    case FALM_ADV_KK: FALM_PU3(FALM_ADV_KK); break;
    default: FALM_PUK(FALM_ADV_UPWIND1); break;
    }
#undef FALM_PUK
#undef FALM_PU3
    FALM_CHECK_LAUNCH("k_pseudo_u_km");
    // u, uu, nut read, ua written (ff only inside the rotor's force box)
    tm.done((double)interior_box(g).size() * 10.0 * sizeof(T));
    return FALM_OK;
}

falm_status sep_pseudo_u(const falm_domain *dom, int prec, int scheme, double ReI, double dt, const void *un,
                         const void *u, const void *uu, void *ua, const void *nut, const SepMesh &sm, const void *ff,
                         const int *ffbox, hipStream_t s) {
    if (un == u) {  // the step driver's form: ua = u + dt(...)
        if (prec == FALM_F32) return pseudo_u_km_t<float>(dom, scheme, ReI, dt, u, uu, ua, nut, sm, ff, ffbox, s);
        return pseudo_u_km_t<double>(dom, scheme, ReI, dt, u, uu, ua, nut, sm, ff, ffbox, s);
    }
    if (prec == FALM_F32)
        return pseudo_u_t<float>(dom, scheme, ReI, dt, un, u, uu, ua, nut, met_sep<float>(sm), ff, ffbox, 13.0, s);
    return pseudo_u_t<double>(dom, scheme, ReI, dt, un, u, uu, ua, nut, met_sep<double>(sm), ff, ffbox, 13.0, s);
}
template <class T>
static falm_status u_to_uu_km_t(const falm_domain *dom, const void *u, void *uu, const SepMesh &sm, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_FACE_FLUX, s);
    FALM_TRY(check_byte_offsets(n, sizeof(T)));
    hipLaunchKernelGGL((k_u_to_uu_km<T>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, v3<T>(u, n), v3w<T>(uu, n),
                       sep_tab<T>(sm));
    const long long nl = face_layer_count(g);
    hipLaunchKernelGGL((k_u_to_uu_layer<T>), ew_grid((size_t)nl), dim3(EW_BLOCK), 0, s, g, v3<T>(u, n),
                       v3<T>(u, n), v3w<T>(uu, n), sep_tab<T>(sm));
    FALM_CHECK_LAUNCH("k_u_to_uu_km");
    tm.done((double)make_tiling(face_box(g)).box.size() * 6.0 * sizeof(T));
    return FALM_OK;
}

template <class T>
static falm_status u_to_uu_rhs_km_t(const falm_domain *dom, const void *u, const void *gu, void *uu, double *rhs,
                                    const SepMesh &sm, double scale, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_FACE_FLUX, s);
    FALM_TRY(check_byte_offsets(n, sizeof(double)));
    hipLaunchKernelGGL((k_u_to_uu_rhs_km<T>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, v3<T>(u, n),
                       v3<T>(gu, n), v3w<T>(uu, n), sep_tab<T>(sm), rhs, scale);
    const long long nl = face_layer_count(g);
    hipLaunchKernelGGL((k_u_to_uu_layer<T>), ew_grid((size_t)nl), dim3(EW_BLOCK), 0, s, g, v3<T>(u, n),
                       v3<T>(gu, n), v3w<T>(uu, n), sep_tab<T>(sm));
    FALM_CHECK_LAUNCH("k_u_to_uu_rhs_km");
    // u read, uu written, rhs written
    tm.done((double)make_tiling(face_box(g)).box.size() * 6.0 * sizeof(T) +
            (double)interior_box(g).size() * sizeof(double));
    return FALM_OK;
}

falm_status sep_u_to_uu_rhs(const falm_domain *dom, int prec, const void *u, const void *gu, void *uu, double *rhs,
                            const SepMesh &sm, double scale, hipStream_t s) {
    if (prec == FALM_F32) return u_to_uu_rhs_km_t<float>(dom, u, gu, uu, rhs, sm, scale, s);
    return u_to_uu_rhs_km_t<double>(dom, u, gu, uu, rhs, sm, scale, s);
}

falm_status sep_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu, const SepMesh &sm,
                        hipStream_t s) {
    if (prec == FALM_F32) return u_to_uu_km_t<float>(dom, u, uu, sm, s);
    return u_to_uu_km_t<double>(dom, u, uu, sm, s);
}
falm_status sep_poisson_rhs(const falm_domain *dom, int prec, const void *uu, double *rhs, const SepMesh &sm,
                            double scale, hipStream_t s) {
    if (prec == FALM_F32) return divergence_t<float, double, true>(dom, uu, rhs, met_tab<float>(sm), scale, false, s);
    return divergence_t<double, double, true>(dom, uu, rhs, met_tab<double>(sm), scale, false, s);
}
template <class T>
static falm_status project_km_t(const falm_domain *dom, double dt, void *u, void *uu, const double *p,
                                const SepMesh &sm, hipStream_t s, const int *gate) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_PROJECT, s);
    FALM_TRY(check_byte_offsets(n, sizeof(double)));
    hipLaunchKernelGGL((k_project_km<T, double>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, dt, v3w<T>(u, n),
                       v3w<T>(uu, n), p, sep_tab<double>(sm), gate);
    const long long nl = face_layer_count(g);
    hipLaunchKernelGGL((k_project_layer<T, double>), ew_grid((size_t)nl), dim3(EW_BLOCK), 0, s, g, dt, v3w<T>(uu, n),
                       p, sep_tab<double>(sm), gate);
    FALM_CHECK_LAUNCH("k_project_km");
    // u, uu read and written, p read
    tm.done((double)make_tiling(face_box(g)).box.size() * (12.0 * sizeof(T) + sizeof(double)));
    return FALM_OK;
}

falm_status sep_project(const falm_domain *dom, int prec, double dt, void *u, void *uu, const double *p,
                        const SepMesh &sm, hipStream_t s, const int *gate) {
    if (prec == FALM_F32) return project_km_t<float>(dom, dt, u, uu, p, sm, s, gate);
    return project_km_t<double>(dom, dt, u, uu, p, sm, s, gate);
}
template <class T>
static falm_status smag_km_t(const falm_domain *dom, double Cs, const void *u, void *nut, const SepMesh &sm,
                             hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_SGS, s);
    FALM_TRY(check_byte_offsets(n, sizeof(T)));
    hipLaunchKernelGGL((k_smagorinsky_km<T>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, (T)Cs, v3<T>(u, n),
                       (T *)nut, sep_tab<T>(sm));
    FALM_CHECK_LAUNCH("k_smagorinsky_km");
    tm.done((double)interior_box(g).size() * 4.0 * sizeof(T));  // u read, nut written
    return FALM_OK;
}

falm_status sep_sgs(const falm_domain *dom, int prec, double Cs, const void *u, void *nut, const SepMesh &sm,
                    hipStream_t s) {
    if (prec == FALM_F32) return smag_km_t<float>(dom, Cs, u, nut, sm, s);
    return smag_km_t<double>(dom, Cs, u, nut, sm, s);
}
template <class T>
static falm_status sgs_div_norm_t(const falm_domain *dom, double Cs, const void *u, void *nut, const void *uu,
                                  void *div, const SepMesh &sm, double *dev_out, double *partials, hipStream_t s,
                                  const int *gate, double *host_out) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    K7Tiling kt = make_k7_tiling(interior_box(g));
    if (kt.blocks == 0) return FALM_OK;
    FALM_TRY(check_byte_offsets(n, sizeof(T)));
    TimedLaunch tm(FALM_TIME_SGS, s);
    hipLaunchKernelGGL((k_sgs_div_km<T>), dim3(kt.blocks), dim3(K7_NT), 0, s, kt, g, (T)Cs, v3<T>(u, n), (T *)nut,
                       v3<T>(uu, n), (T *)div, sep_tab<T>(sm), partials, gate);
    tm.done((double)interior_box(g).size() * 8.0 * sizeof(T));  // u, uu read; nut, div written
    hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(FIN_NT), 0, s, (const double *)partials, kt.blocks, dev_out, gate,
                       host_out);
    FALM_CHECK_LAUNCH("k_sgs_div_km");
    return FALM_OK;
}

falm_status sep_sgs_div_norm(const falm_domain *dom, int prec, double Cs, const void *u, void *nut, const void *uu,
                             void *div, const SepMesh &sm, double *dev_out, double *partials, hipStream_t s,
                             const int *gate, double *host_out) {
    if (prec == FALM_F32)
        return sgs_div_norm_t<float>(dom, Cs, u, nut, uu, div, sm, dev_out, partials, s, gate, host_out);
    return sgs_div_norm_t<double>(dom, Cs, u, nut, uu, div, sm, dev_out, partials, s, gate, host_out);
}

falm_status sep_div_norm(const falm_domain *dom, int prec, const void *uu, void *div, const SepMesh &sm,
                         double *dev_out, double *partials, hipStream_t s, const int *gate, double *host_out) {
    if (prec == FALM_F32) return div_norm_t<float>(dom, uu, div, sm, dev_out, partials, s, gate, host_out);
    return div_norm_t<double>(dom, uu, div, sm, dev_out, partials, s, gate, host_out);
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
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32)
        return pseudo_u_t<float>(dom, adv_scheme, ReI, dt, un, u, uu, ua, nut, met_field<float>(kx, g, ja, n), ff,
                                 nullptr, 23.0, s);
    return pseudo_u_t<double>(dom, adv_scheme, ReI, dt, un, u, uu, ua, nut, met_field<double>(kx, g, ja, n), ff,
                              nullptr, 23.0, s);
}

falm_status falm_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu, const void *kx,
                         const void *ja, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!u || !uu || !kx || !ja) {
        set_error("falm_u_to_uu: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32) return u_to_uu_t<float>(dom, u, uu, met_field<float>(kx, nullptr, ja, n), s);
    return u_to_uu_t<double>(dom, u, uu, met_field<double>(kx, nullptr, ja, n), s);
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
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32)
        return divergence_t<float, float, true>(dom, uu, rhs, met_field<float>(nullptr, nullptr, ja, n), scale, true, s);
    return divergence_t<double, double, true>(dom, uu, rhs, met_field<double>(nullptr, nullptr, ja, n), scale, true,
                                              s);
}

falm_status falm_project_p(const falm_domain *dom, int prec, double dt, void *u, const void *ua, void *uu,
                           const void *uua, const void *p, const void *kx, const void *g, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!u || !ua || !uu || !uua || !p || !kx || !g) {
        set_error("falm_project_p: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32) return project_t<float, float>(dom, dt, u, ua, uu, uua, p, met_field<float>(kx, g, nullptr, n), s);
    return project_t<double, double>(dom, dt, u, ua, uu, uua, p, met_field<double>(kx, g, nullptr, n), s);
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
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32) return smag_t<float>(dom, Cs, u, nut, met_field<float>(kx, nullptr, ja, n), s);
    return smag_t<double>(dom, Cs, u, nut, met_field<double>(kx, nullptr, ja, n), s);
}

falm_status falm_divergence(const falm_domain *dom, int prec, const void *uu, void *div, const void *ja,
                            void *stream) {
    FALM_TRY(check_domain(dom));
    if (!uu || !div || !ja) {
        set_error("falm_divergence: null field pointer");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t n = ncell_of(dom);
    if (prec == FALM_F32)
        return divergence_t<float, float, false>(dom, uu, div, met_field<float>(nullptr, nullptr, ja, n), 1.0, false, s);
    return divergence_t<double, double, false>(dom, uu, div, met_field<double>(nullptr, nullptr, ja, n), 1.0, false,
                                               s);
}

}  // extern "C"
