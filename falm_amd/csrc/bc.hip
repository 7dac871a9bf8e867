// bc.hip — ghost-cell boundary fill, metrics and Poisson matrix assembly.
//
// On one GPU the halo exchange of the reference (CPMComm, a no-op on one
// rank) disappears and the ghost cells are filled only by the apps' boundary
// kernels.  The reference launches one kernel per face on six streams
// (app/windtunnel/bc.h:6-68); the faces write disjoint cells and read only
// interior cells, so here all faces go in one launch.
#include <cmath>
#include <vector>

#include "linalg.hpp"

namespace falm {

struct FaceDesc {
    int type;          // FALM_BC_* / FALM_PBC_* / non-zero mask
    int d, ax, bx;     // normal axis and the two in-face axes
    int plus;          // 1 for the + face
    int na, nb;        // interior extent along ax, bx
    int start;         // first flattened index of this face
    double value[3];
};
struct FaceSet {
    FaceDesc f[6];
    int nface;
    int total;
};

static FaceSet make_faces(const Geom &g, const int type[6], const double (*value)[3]) {
    FaceSet fs;
    fs.nface = 0;
    fs.total = 0;
    const int shp[3] = {g.nx, g.ny, g.nz};
    for (int f = 0; f < 6; f++) {
        if (!type[f]) continue;
        FaceDesc &F = fs.f[fs.nface++];
        F.type = type[f];
        F.d = f / 2;
        F.ax = F.d == 0 ? 1 : 0;
        F.bx = F.d == 2 ? 1 : 2;
        F.plus = f & 1;
        F.na = shp[F.ax] - 2 * g.gc;
        F.nb = shp[F.bx] - 2 * g.gc;
        F.start = fs.total;
        for (int m = 0; m < 3; m++) F.value[m] = value ? value[f][m] : 0.0;
        fs.total += F.na * F.nb;
    }
    return fs;
}

// locate the face and the in-face cell of flattened index t
__device__ inline const FaceDesc *face_of(const FaceSet &fs, int t, int &ka, int &kb) {
    int q = 0;
    while (q + 1 < fs.nface && t >= fs.f[q + 1].start) q++;
    const FaceDesc &F = fs.f[q];
    int r = t - F.start;
    kb = r / F.na;
    ka = r - kb * F.na;
    return &F;
}

// cell index of the cell at normal position `pos` (counted from the face:
// 0 = first ghost o1, -1 = second ghost o2, 1 = first interior i1, 2 = i2)
__device__ inline int face_cell(const Geom &g, const FaceDesc &F, int ka, int kb, int pos) {
    // arithmetic rather than a select of the three shape fields, which the
    // compiler turns into an indexed load from a scratch copy of g
    const int n = (F.d == 0) * g.nx + (F.d == 1) * g.ny + (F.d == 2) * g.nz;
    const int a = ka + g.gc, b = kb + g.gc;
    // - face: o1 = gc-1, o2 = gc-2, i1 = gc, i2 = gc+1
    // + face: o1 = n-gc, o2 = n-gc+1, i1 = n-gc-1, i2 = n-gc-2
    const int cn = F.plus ? (n - g.gc - pos) : (g.gc - 1 + pos);
    // axis d gets cn, ax gets a, bx gets b (selects, not a dynamically
    // indexed array, which would live in scratch memory)
    const int c0 = F.d == 0 ? cn : F.ax == 0 ? a : b;
    const int c1 = F.d == 1 ? cn : F.ax == 1 ? a : b;
    const int c2 = F.d == 2 ? cn : F.ax == 2 ? a : b;
    return g.id(c0, c1, c2);
}

// Philox4x32-10 (Salmon et al., SC'11) keyed by the seed, counter
// (step, a, b); Box-Muller gives three N(0,1) deviates.
__device__ inline void inlet_normals(uint64_t seed, int64_t step, int a, int b, double out[3]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t c0 = (uint32_t)step, c1 = (uint32_t)((uint64_t)step >> 32), c2 = (uint32_t)a, c3 = (uint32_t)b;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    const double S = 2.3283064365386963e-10;
    double u1 = ((double)c0 + 1.0) * S, u2 = (double)c1 * S, u3 = ((double)c2 + 1.0) * S, u4 = (double)c3 * S;
    const double TWO_PI = 6.283185307179586;
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(TWO_PI * u2);
    out[1] = r1 * sin(TWO_PI * u2);
    out[2] = r2 * cos(TWO_PI * u4);
}

// ubc: app/windtunnel/bcdevcall.cu:8-115, app/lid3d/bcdev.cu:11-105
template <class T>
__global__ void k_bc_velocity(FaceSet fs, Geom g, T *u0, T *u1, T *u2, const T *p0, const T *p1, const T *p2,
                              const T *xc, T dt, falm_inlet_noise noise, const int *gate) {
    if (gated_off(gate)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fs.total) return;
    int ka, kb;
    const FaceDesc &F = *face_of(fs, t, ka, kb);
    T *u[3] = {u0, u1, u2};
    const T *up[3] = {p0, p1, p2};
    const int o1 = face_cell(g, F, ka, kb, 0), o2 = face_cell(g, F, ka, kb, -1);
    const int i1 = face_cell(g, F, ka, kb, 1), i2 = face_cell(g, F, ka, kb, 2);
    if (F.type == FALM_BC_INLET) {
        T v[3] = {(T)F.value[0], (T)F.value[1], (T)F.value[2]};
        if (noise.intensity != 0.0) {
            double z[3];
            inlet_normals(noise.seed, noise.step, ka + g.gc, kb + g.gc, z);
            for (int m = 0; m < 3; m++) v[m] = (T)(F.value[m] + noise.intensity * z[m]);
        }
        for (int m = 0; m < 3; m++) {
            u[m][o1] = v[m];
            u[m][o2] = v[m];
        }
    } else if (F.type == FALM_BC_CONVECTIVE) {
        const T dxi = T(1.0) / (xc[o1] - xc[i1]);
        for (int m = 0; m < 3; m++) {
            const T du = T(0.5) * dxi * (T(3) * up[m][o1] - T(4) * up[m][i1] + up[m][i2]);
            const T ub = up[m][o1] - dt * du;
            u[m][o1] = ub;
            u[m][o2] = T(2) * ub - u[m][i1];
        }
    } else if (F.type == FALM_BC_SLIP) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const T ui = u[m][i1];
            const T ub = m == F.d ? T(0.0) : ui;
            u[m][o1] = ub;
            u[m][o2] = T(2) * ub - ui;
        }
    } else if (F.type == FALM_BC_WALL0) {
        for (int m = 0; m < 3; m++) {
            const T ui = u[m][i1];
            u[m][o1] = T(0.0);
            u[m][o2] = T(2) * T(0.0) - ui;
        }
    } else if (F.type == FALM_BC_REFLECT) {
        for (int m = 0; m < 3; m++) {
            const T v = (T)F.value[m];
            const T a = u[m][i1], b = u[m][i2];
            u[m][o1] = T(2) * v - a;
            u[m][o2] = T(2) * v - b;
        }
    }
}

// pbc: app/windtunnel/bcdevcall.cu:153-211 (Neumann copy or p = 0)
template <class T>
__global__ void k_bc_pressure(FaceSet fs, Geom g, T *p, const int *gate) {
    if (gated_off(gate)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fs.total) return;
    int ka, kb;
    const FaceDesc &F = *face_of(fs, t, ka, kb);
    const int o1 = face_cell(g, F, ka, kb, 0), i1 = face_cell(g, F, ka, kb, 1);
    p[o1] = F.type == FALM_PBC_ZERO ? T(0.0) : p[i1];
}

// uubc: app/lid3d/bcdev.cu:251-321 (the wall face is the east/north/top
// face of the last ghost cell on the - side, of the last interior cell on
// the + side)
template <class T>
__global__ void k_bc_face_flux(FaceSet fs, Geom g, T *uu0, T *uu1, T *uu2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fs.total) return;
    int ka, kb;
    const FaceDesc &F = *face_of(fs, t, ka, kb);
    const int c = face_cell(g, F, ka, kb, F.plus ? 1 : 0);
    T *uu = F.d == 0 ? uu0 : F.d == 1 ? uu1 : uu2;
    uu[c] = T(0);
}

// copy_z5: app/ld2d/bcdevcall.cu:240-256
template <class T>
__global__ void k_copy_z5(Geom g, T *v, int ncol, size_t ncell, const int *gate) {
    if (gated_off(gate)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nplane = g.nx * g.ny;
    if (t >= nplane * ncol) return;
    const int n = t / nplane, ij = t - n * nplane;
    T *q = v + (size_t)n * ncell;
    const T val = q[ij + g.nxy * g.gc];
    q[ij + g.nxy * (g.gc - 2)] = val;
    q[ij + g.nxy * (g.gc + 2)] = val;
    q[ij + g.nxy * (g.gc - 1)] = val;
    q[ij + g.nxy * (g.gc + 1)] = val;
}

// makePoissonCoefficientMatrix (app/windtunnel/main.cpp:20-92 and the
// lid3d / ld2d variants)
template <class T>
__global__ __launch_bounds__(256) void k_poisson_matrix(Tiling tl, Geom g, int kind, int ox, int oy, int oz, int gnx,
                                                        int gny, int gnz, const T *__restrict__ gx,
                                                        const T *__restrict__ gy, const T *__restrict__ gz,
                                                        const T *__restrict__ ja, T *a0, T *a1, T *a2, T *a3, T *a4,
                                                        T *a5, T *a6) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const int gi = ox + i, gj = oy + j, gk = oz + k, gc = g.gc;
        const int xe_lim = kind == FALM_POISSON_OUTLET_XPLUS ? gnx - gc : gnx - gc - 1;
        T ac = 0, ae = 0, aw = 0, an = 0, as = 0, at = 0, ab = 0;
        const T jacob = ja[cc];
        T coef;
        coef = T(0.5) * (gx[cc] + gx[cc + 1]) / jacob;
        if (gi < xe_lim) { ac -= coef; ae = coef; }
        coef = T(0.5) * (gx[cc] + gx[cc - 1]) / jacob;
        if (gi > gc) { ac -= coef; aw = coef; }
        coef = T(0.5) * (gy[cc] + gy[cc + g.nx]) / jacob;
        if (gj < gny - gc - 1) { ac -= coef; an = coef; }
        coef = T(0.5) * (gy[cc] + gy[cc - g.nx]) / jacob;
        if (gj > gc) { ac -= coef; as = coef; }
        if (kind != FALM_POISSON_2D) {
            coef = T(0.5) * (gz[cc] + gz[cc + g.nxy]) / jacob;
            if (gk < gnz - gc - 1) { ac -= coef; at = coef; }
            coef = T(0.5) * (gz[cc] + gz[cc - g.nxy]) / jacob;
            if (gk > gc) { ac -= coef; ab = coef; }
        }
        a0[cc] = ac; a1[cc] = aw; a2[cc] = ae; a3[cc] = as; a4[cc] = an; a5[cc] = ab; a6[cc] = at;
    });
}

template <class T>
__global__ void k_scale_all(T *a, size_t n, T s) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        a[t] *= s;
}

// -------------------------------------------------------------- host
template <class T>
static falm_status bc_velocity_t(const falm_domain *dom, const falm_face_bc faces[6], void *u, const void *uprev,
                                 const void *xyz, double dt, const falm_inlet_noise *noise, hipStream_t s,
                                 const int *gate) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    int type[6];
    double val[6][3];
    bool need_prev = false;
    for (int f = 0; f < 6; f++) {
        type[f] = faces[f].type;
        for (int m = 0; m < 3; m++) val[f][m] = faces[f].value[m];
        if (type[f] < FALM_BC_NONE || type[f] > FALM_BC_REFLECT) {
            set_error("falm_bc_velocity: bad face type %d", type[f]);
            return FALM_ERR_INVALID;
        }
        if (type[f] == FALM_BC_CONVECTIVE) {
            if (f != 1) {
                set_error("falm_bc_velocity: the convective outflow is the x+ face only");
                return FALM_ERR_UNSUPPORTED;
            }
            need_prev = true;
        }
    }
    if (need_prev && (!uprev || !xyz)) {
        set_error("falm_bc_velocity: convective outflow needs uprev and xyz");
        return FALM_ERR_INVALID;
    }
    FaceSet fs = make_faces(g, type, val);
    if (fs.total == 0) return FALM_OK;
    falm_inlet_noise nz{0.0, 0, 0};
    if (noise) nz = *noise;
    const T *pv = (const T *)(uprev ? uprev : u);
    const T *xv = (const T *)(xyz ? xyz : u);
    hipLaunchKernelGGL((k_bc_velocity<T>), dim3((fs.total + 255) / 256), dim3(256), 0, s, fs, g, col<T>(u, n, 0),
                       col<T>(u, n, 1), col<T>(u, n, 2), pv, pv + n, pv + 2 * n, xv, (T)dt, nz, gate);
    FALM_CHECK_LAUNCH("k_bc_velocity");
    return FALM_OK;
}

falm_status launch_bc_velocity(const falm_domain *dom, int prec, const falm_face_bc faces[6], void *u,
                               const void *uprev, const void *xyz, double dt, const falm_inlet_noise *noise,
                               hipStream_t s, const int *gate) {
    return prec == FALM_F32 ? bc_velocity_t<float>(dom, faces, u, uprev, xyz, dt, noise, s, gate)
                            : bc_velocity_t<double>(dom, faces, u, uprev, xyz, dt, noise, s, gate);
}

falm_status launch_bc_pressure(const falm_domain *dom, int prec, const int faces[6], void *p, hipStream_t s,
                               const int *gate) {
    Geom g = make_geom(dom);
    FaceSet fs = make_faces(g, faces, nullptr);
    if (fs.total == 0) return FALM_OK;
    if (prec == FALM_F32)
        hipLaunchKernelGGL((k_bc_pressure<float>), dim3((fs.total + 255) / 256), dim3(256), 0, s, fs, g, (float *)p,
                           gate);
    else
        hipLaunchKernelGGL((k_bc_pressure<double>), dim3((fs.total + 255) / 256), dim3(256), 0, s, fs, g, (double *)p,
                           gate);
    FALM_CHECK_LAUNCH("k_bc_pressure");
    return FALM_OK;
}

falm_status launch_bc_face_flux(const falm_domain *dom, int prec, const int faces[6], void *uu, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    FaceSet fs = make_faces(g, faces, nullptr);
    if (fs.total == 0) return FALM_OK;
    if (prec == FALM_F32)
        hipLaunchKernelGGL((k_bc_face_flux<float>), dim3((fs.total + 255) / 256), dim3(256), 0, s, fs, g,
                           col<float>(uu, n, 0), col<float>(uu, n, 1), col<float>(uu, n, 2));
    else
        hipLaunchKernelGGL((k_bc_face_flux<double>), dim3((fs.total + 255) / 256), dim3(256), 0, s, fs, g,
                           col<double>(uu, n, 0), col<double>(uu, n, 1), col<double>(uu, n, 2));
    FALM_CHECK_LAUNCH("k_bc_face_flux");
    return FALM_OK;
}

falm_status launch_copy_z5(const falm_domain *dom, int prec, void *v, int ncol, hipStream_t s, const int *gate) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    int total = g.nx * g.ny * ncol;
    if (prec == FALM_F32)
        hipLaunchKernelGGL((k_copy_z5<float>), dim3((total + 255) / 256), dim3(256), 0, s, g, (float *)v, ncol, n,
                           gate);
    else
        hipLaunchKernelGGL((k_copy_z5<double>), dim3((total + 255) / 256), dim3(256), 0, s, g, (double *)v, ncol, n,
                           gate);
    FALM_CHECK_LAUNCH("k_copy_z5");
    return FALM_OK;
}

template <class T>
static falm_status metrics_t(const falm_domain *dom, const double *x, const double *y, const double *z,
                             const double *hx, const double *hy, const double *hz, void *xyz, void *kx, void *gg,
                             void *ja, hipStream_t s) {
    const int nx = dom->shape[0], ny = dom->shape[1], nz = dom->shape[2];
    const size_t n = ncell_of(dom);
    std::vector<T> hxyz(3 * n), hkx(3 * n), hg(3 * n), hja(n);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                hxyz[c] = (T)x[i];
                hxyz[c + n] = (T)y[j];
                hxyz[c + 2 * n] = (T)z[k];
                const double pitch[3] = {hx[i], hy[j], hz[k]};
                const double volume = pitch[0] * pitch[1] * pitch[2];
                hja[c] = (T)volume;
                for (int d = 0; d < 3; d++) {
                    const double dk = 1.0 / pitch[d];
                    hg[c + d * n] = (T)(volume * dk * dk);
                    hkx[c + d * n] = (T)dk;
                }
            }
    FALM_HIP(hipMemcpyAsync(xyz, hxyz.data(), 3 * n * sizeof(T), hipMemcpyHostToDevice, s));
    FALM_HIP(hipMemcpyAsync(kx, hkx.data(), 3 * n * sizeof(T), hipMemcpyHostToDevice, s));
    FALM_HIP(hipMemcpyAsync(gg, hg.data(), 3 * n * sizeof(T), hipMemcpyHostToDevice, s));
    FALM_HIP(hipMemcpyAsync(ja, hja.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

template <class T>
static falm_status poisson_matrix_t(const falm_domain *dom, int kind, const void *gg, const void *ja, void *a,
                                    double *maxdiag, hipStream_t s) {
    Geom g = make_geom(dom);
    size_t n = ncell_of(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_ERR_INVALID;
    hipLaunchKernelGGL((k_poisson_matrix<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, kind, dom->offset[0],
                       dom->offset[1], dom->offset[2], dom->gshape[0], dom->gshape[1], dom->gshape[2],
                       col<T>(gg, n, 0), col<T>(gg, n, 1), col<T>(gg, n, 2), (const T *)ja, col<T>(a, n, 0),
                       col<T>(a, n, 1), col<T>(a, n, 2), col<T>(a, n, 3), col<T>(a, n, 4), col<T>(a, n, 5),
                       col<T>(a, n, 6));
    FALM_CHECK_LAUNCH("k_poisson_matrix");
    double md = 0.0;
    FALM_TRY(host_minmax(dom, PrecOf<T>::value, FALM_RED_COL_ABSMAX, a, 7, 0, &md, s));
    if (!(md > 0.0)) {
        set_error("falm_poisson_matrix: max |a_c| is %g", md);
        return FALM_ERR_INVALID;
    }
    hipLaunchKernelGGL((k_scale_all<T>), ew_grid(7 * n), dim3(EW_BLOCK), 0, s, (T *)a, 7 * n, (T)(1.0 / md));
    FALM_CHECK_LAUNCH("k_scale_all");
    if (maxdiag) *maxdiag = md;
    return FALM_OK;
}

}  // namespace falm

using namespace falm;

extern "C" {

falm_status falm_bc_velocity(const falm_domain *dom, int prec, const falm_face_bc faces[6], void *u,
                             const void *uprev, const void *xyz, double dt, const falm_inlet_noise *noise,
                             void *stream) {
    FALM_TRY(check_domain(dom));
    if (!faces || !u) return FALM_ERR_INVALID;
    return launch_bc_velocity(dom, prec, faces, u, uprev, xyz, dt, noise, (hipStream_t)stream, nullptr);
}

falm_status falm_bc_pressure(const falm_domain *dom, int prec, const int faces[6], void *p, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!faces || !p) return FALM_ERR_INVALID;
    for (int f = 0; f < 6; f++)
        if (faces[f] < FALM_PBC_NONE || faces[f] > FALM_PBC_ZERO) return FALM_ERR_INVALID;
    return launch_bc_pressure(dom, prec, faces, p, (hipStream_t)stream, nullptr);
}

falm_status falm_bc_face_flux(const falm_domain *dom, int prec, const int faces[6], void *uu, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!faces || !uu) return FALM_ERR_INVALID;
    return launch_bc_face_flux(dom, prec, faces, uu, (hipStream_t)stream);
}

falm_status falm_copy_z5(const falm_domain *dom, int prec, void *v, int ncol, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!v || ncol < 1) return FALM_ERR_INVALID;
    if (dom->shape[2] != 2 * dom->gc + 1) {
        set_error("falm_copy_z5: needs a z-thin domain (one interior layer)");
        return FALM_ERR_INVALID;
    }
    return launch_copy_z5(dom, prec, v, ncol, (hipStream_t)stream, nullptr);
}

falm_status falm_metrics(const falm_domain *dom, int prec, const double *x, const double *y, const double *z,
                         const double *hx, const double *hy, const double *hz, void *xyz, void *kx, void *g,
                         void *ja, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!x || !y || !z || !hx || !hy || !hz || !xyz || !kx || !g || !ja) return FALM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? metrics_t<float>(dom, x, y, z, hx, hy, hz, xyz, kx, g, ja, s)
                            : metrics_t<double>(dom, x, y, z, hx, hy, hz, xyz, kx, g, ja, s);
}

falm_status falm_poisson_matrix(const falm_domain *dom, int prec, int kind, const void *g, const void *ja, void *a,
                                double *maxdiag, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!g || !ja || !a) return FALM_ERR_INVALID;
    if (kind < FALM_POISSON_NEUMANN || kind > FALM_POISSON_2D) return FALM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? poisson_matrix_t<float>(dom, kind, g, ja, a, maxdiag, s)
                            : poisson_matrix_t<double>(dom, kind, g, ja, a, maxdiag, s);
}

}  // extern "C"
