// linalg.hip — FalmMV / FalmEq building blocks on gfx950: 7-point
// matrix-vector product, residual, dot / squared norm (deterministic grid
// reduction), max/min reductions, scale/add/axpby, red-black SOR and Jacobi
// sweeps.  Stencil coefficients are the 7 columns [c,w,e,s,n,b,t] of the
// reference's Poisson matrix (src/dev/MVDevCall.cu:25-31).
#include <mutex>

#include "linalg.hpp"

namespace falm {

// ------------------------------------------------------------ kernels
template <class TA, class TX>
__global__ __launch_bounds__(256) void k_mv(Tiling tl, Geom g, Mat7<TA> a, const TX *__restrict__ x,
                                            TX *__restrict__ ax) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        ax[cc] = apply7<TA, TX, TX>(a, x, cc, g.nx, g.nxy);
    });
}

template <class TA, class TX>
__global__ __launch_bounds__(256) void k_residual(Tiling tl, Geom g, Mat7<TA> a, const TX *__restrict__ x,
                                                  const TX *__restrict__ b, TX *__restrict__ r) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        r[cc] = b[cc] - apply7<TA, TX, TX>(a, x, cc, g.nx, g.nxy);
    });
}

// kernel_DotProduct / kernel_EuclideanNormSq (src/dev/MVDevCall.cu:57-179):
// products accumulated in double, result in out[0].
template <class T, bool NORM>
__global__ __launch_bounds__(256) void k_dot(Tiling tl, Geom g, const T *__restrict__ a, const T *__restrict__ b,
                                             int ncol, size_t ncell, double *partials) {
    __shared__ double lds[4];
    int i, j, k, t0, t1;
    double v[1] = {0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        if (NORM) {
            for (int d = 0; d < ncol; d++) {
                double x = (double)a[cc + d * ncell];
                v[0] += x * x;
            }
        } else {
            v[0] += (double)a[cc] * (double)b[cc];
        }
    }
    block_partials<256, 1>(v, partials, lds);
}

__global__ __launch_bounds__(FIN_NT) void k_fin_sum(const double *partials, int nb, double *out) {
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) out[0] = tot[0];
}

// per-block max/min partials (kernels src/dev/MVDevCall.cu:216-598)
template <class T>
__global__ __launch_bounds__(256) void k_minmax(Tiling tl, Geom g, const T *__restrict__ a, int op, int ncol,
                                                int col, size_t ncell, double *partials) {
    __shared__ double lds[256];
    __shared__ int any[256];
    int i, j, k;
    double v = 0.0;
    int have = 0;
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        if (op == FALM_RED_VEC_MAX || op == FALM_RED_VEC_MIN) {
            double t = 0.0;
            for (int d = 0; d < ncol; d++) {
                double x = (double)a[cc + d * ncell];
                t += x * x;
            }
            v = sqrt(t);
        } else {
            v = (double)a[cc + (size_t)col * ncell];
            if (op == FALM_RED_COL_ABSMAX || op == FALM_RED_COL_ABSMIN) v = fabs(v);
        }
        have = 1;
    }
    const bool mx = (op == FALM_RED_COL_MAX || op == FALM_RED_COL_ABSMAX || op == FALM_RED_VEC_MAX);
    lds[threadIdx.x] = v;
    any[threadIdx.x] = have;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            int o = threadIdx.x + s;
            if (any[o]) {
                if (!any[threadIdx.x] || (mx ? lds[o] > lds[threadIdx.x] : lds[o] < lds[threadIdx.x])) {
                    lds[threadIdx.x] = lds[o];
                    any[threadIdx.x] = 1;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = lds[0];
        partials[2 * blockIdx.x + 1] = any[0] ? 1.0 : 0.0;
    }
}

template <class T>
__global__ void k_scale(T *a, size_t n, T s) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        a[t] *= s;
}
template <class T>
__global__ void k_add(T *a, const T *b, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        a[t] += b[t];
}
template <class T>
__global__ void k_axpby(T a, const T *x, T b, const T *y, T *r, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        r[t] = a * x[t] + b * y[t];
}

// kernel_Struct3d7p_SOR (src/dev/FalmEqDevCall.cu:175-211): the reference
// rewrites cells of the other colour with their own value; here only the
// cells of `color` are written.  Optional device `done` flag: the sweep is a
// no-op once a solver has finished.
template <class TA, class TX>
__global__ __launch_bounds__(256) void k_sor(Tiling tl, Geom g, Mat7<TA> a, TX *x, const TX *__restrict__ b,
                                             TX omega, int color, const int *done) {
    if (done && *done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        if (((i + j + k + g.osum) & 1) != color) return;
        const int cc = g.id(i, j, k);
        const TX xc = x[cc];
        const TX s = apply7<TA, TX, TX>(a, x, cc, g.nx, g.nxy);
        const TX cv = (b[cc] - s) / (TX)a.c[0][cc];
        x[cc] = xc + omega * cv;
    });
}

// kernel_Struct3d7p_Jacobi (src/dev/FalmEqDevCall.cu:101-132)
template <class TA, class TX>
__global__ __launch_bounds__(256) void k_jacobi(Tiling tl, Geom g, Mat7<TA> a, TX *__restrict__ x,
                                                const TX *__restrict__ xp, const TX *__restrict__ b,
                                                const int *done) {
    if (done && *done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const TX xc = xp[cc];
        x[cc] = xc + (b[cc] - apply7<TA, TX, TX>(a, xp, cc, g.nx, g.nxy)) / (TX)a.c[0][cc];
    });
}

// ------------------------------------------------------------ launches
template <class T>
falm_status launch_sor(const falm_domain *dom, const void *a, void *x, const void *b, double omega, int color,
                       const int *done, hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_SOR, s);
    hipLaunchKernelGGL((k_sor<T, T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, mat7<T>(a, ncell_of(dom)),
                       (T *)x, (const T *)b, (T)omega, color, done);
    FALM_CHECK_LAUNCH("k_sor");
    tm.done((double)tl.box.size() * 10.0 * sizeof(T));
    return FALM_OK;
}
template falm_status launch_sor<double>(const falm_domain *, const void *, void *, const void *, double, int,
                                        const int *, hipStream_t);
template falm_status launch_sor<float>(const falm_domain *, const void *, void *, const void *, double, int,
                                       const int *, hipStream_t);

template <class T>
falm_status launch_jacobi(const falm_domain *dom, const void *a, void *x, const void *xp, const void *b,
                          const int *done, hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_jacobi<T, T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, mat7<T>(a, ncell_of(dom)),
                       (T *)x, (const T *)xp, (const T *)b, done);
    FALM_CHECK_LAUNCH("k_jacobi");
    return FALM_OK;
}
template falm_status launch_jacobi<double>(const falm_domain *, const void *, void *, const void *, const void *,
                                           const int *, hipStream_t);
template falm_status launch_jacobi<float>(const falm_domain *, const void *, void *, const void *, const void *,
                                          const int *, hipStream_t);

template <class T>
static falm_status mv_t(const falm_domain *dom, const void *a, const void *x, void *ax, hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    TimedLaunch tm(FALM_TIME_MV, s);
    hipLaunchKernelGGL((k_mv<T, T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, mat7<T>(a, ncell_of(dom)),
                       (const T *)x, (T *)ax);
    FALM_CHECK_LAUNCH("k_mv");
    tm.done((double)tl.box.size() * 9.0 * sizeof(T));
    return FALM_OK;
}

template <class T>
falm_status launch_residual(const falm_domain *dom, const void *a, const void *x, const void *b, void *r,
                            hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) return FALM_OK;
    hipLaunchKernelGGL((k_residual<T, T>), tiling_grid(tl), tiling_block(), 0, s, tl, g,
                       mat7<T>(a, ncell_of(dom)), (const T *)x, (const T *)b, (T *)r);
    FALM_CHECK_LAUNCH("k_residual");
    return FALM_OK;
}
template falm_status launch_residual<double>(const falm_domain *, const void *, const void *, const void *, void *,
                                             hipStream_t);
template falm_status launch_residual<float>(const falm_domain *, const void *, const void *, const void *, void *,
                                            hipStream_t);

// dot or squared norm into a device scalar (async) and optionally a host value
falm_status launch_dot(const falm_domain *dom, int prec, const void *a, const void *b, int ncol, bool norm,
                       double *dev_out, double *partials, hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) {  // an empty interior sums to zero (a kernel, not a memset node under capture)
        hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(FIN_NT), 0, s, (const double *)partials, 0, dev_out);
        FALM_CHECK_LAUNCH("k_fin_sum");
        return FALM_OK;
    }
    size_t n = ncell_of(dom);
#define FALM_DOT(T, N)                                                                                          \
    hipLaunchKernelGGL((k_dot<T, N>), red_grid(tl), tiling_block(), 0, s, tl, g, (const T *)a, (const T *)b, ncol, \
                       n, partials)
    if (prec == FALM_F32) {
        if (norm) FALM_DOT(float, true); else FALM_DOT(float, false);
    } else {
        if (norm) FALM_DOT(double, true); else FALM_DOT(double, false);
    }
#undef FALM_DOT
    FALM_CHECK_LAUNCH("k_dot");
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(FIN_NT), 0, s, (const double *)partials, red_blocks(tl), dev_out);
    FALM_CHECK_LAUNCH("k_fin_sum");
    return FALM_OK;
}

static falm_status host_scalar(const falm_domain *dom, int prec, const void *a, const void *b, int ncol,
                               bool norm, double *out, hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    ReduceScratch *rs;
    FALM_TRY(reduce_scratch((size_t)tl.tiles, &rs));
    FALM_TRY(launch_dot(dom, prec, a, b, ncol, norm, rs->result, rs->partials, s));
    FALM_HIP(hipMemcpyAsync(rs->host, rs->result, sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    *out = rs->host[0];
    return FALM_OK;
}

falm_status host_minmax(const falm_domain *dom, int prec, int op, const void *a, int ncol, int col, double *out,
                        hipStream_t s) {
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (tl.tiles == 0) {
        *out = 0.0;
        return FALM_OK;
    }
    ReduceScratch *rs;
    FALM_TRY(reduce_scratch((size_t)tl.tiles, &rs));
    size_t n = ncell_of(dom);
    if (prec == FALM_F32)
        hipLaunchKernelGGL((k_minmax<float>), one_tile_grid(tl), tiling_block(), 0, s, tl, g, (const float *)a, op,
                           ncol, col, n, rs->partials);
    else
        hipLaunchKernelGGL((k_minmax<double>), one_tile_grid(tl), tiling_block(), 0, s, tl, g, (const double *)a, op,
                           ncol, col, n, rs->partials);
    FALM_CHECK_LAUNCH("k_minmax");
    std::vector<double> h(2 * (size_t)tl.tiles);
    FALM_HIP(hipMemcpyAsync(h.data(), rs->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    const bool mx = (op == FALM_RED_COL_MAX || op == FALM_RED_COL_ABSMAX || op == FALM_RED_VEC_MAX);
    bool first = true;
    double r = 0.0;
    for (int b = 0; b < tl.tiles; b++) {
        if (h[2 * b + 1] == 0.0) continue;
        double v = h[2 * b];
        if (first || (mx ? v > r : v < r)) r = v;
        first = false;
    }
    *out = r;
    return FALM_OK;
}

// ------------------------------------------------- allocation-free dot
// Workspace of falm_dot_noalloc: the first pass's per-workgroup partial sums.
// The reduction grid never exceeds RED_BLOCKS workgroups, so this fixed array
// (32 KiB per device, held until the process ends) serves every domain.
constexpr int NOALLOC_PARTIALS = RED_BLOCKS;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

// Device address of `scratch` per device.  In host code the array's name is
// only a stand-in; kernels get this address, taken once per device with that
// device current.
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_noalloc_mx;

falm_status noalloc_setup(double **out) {
    int dev = 0;
    FALM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("device %d beyond the %d devices the allocation-free workspace serves", dev,
                  NOALLOC_MAX_DEVICES);
        return FALM_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mx);
    if (!g_noalloc_ptr[dev]) {
        double *p = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_status(e == hipSuccess ? hipErrorInvalidSymbol : e,
                                                     "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[dev] = p;
    }
    if (out) *out = g_noalloc_ptr[dev];
    return FALM_OK;
}

}  // namespace falm

using namespace falm;

extern "C" {

falm_status falm_mv(const falm_domain *dom, int prec, const void *a, const void *x, void *ax, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !x || !ax) {
        set_error("falm_mv: null pointer");
        return FALM_ERR_INVALID;
    }
    if (ax == x) {
        set_error("falm_mv: ax aliases x");
        return FALM_ERR_ALIAS;
    }
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? mv_t<float>(dom, a, x, ax, s) : mv_t<double>(dom, a, x, ax, s);
}

falm_status falm_residual(const falm_domain *dom, int prec, const void *a, const void *x, const void *b, void *r,
                          void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !x || !b || !r) {
        set_error("falm_residual: null pointer");
        return FALM_ERR_INVALID;
    }
    if (r == x) {
        set_error("falm_residual: r aliases x");
        return FALM_ERR_ALIAS;
    }
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? launch_residual<float>(dom, a, x, b, r, s) : launch_residual<double>(dom, a, x, b, r, s);
}

falm_status falm_dot(const falm_domain *dom, int prec, const void *a, const void *b, double *out, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !b || !out) {
        set_error("falm_dot: null pointer");
        return FALM_ERR_INVALID;
    }
    return host_scalar(dom, prec, a, b, 1, false, out, (hipStream_t)stream);
}

falm_status falm_dot_noalloc(const falm_domain *dom, int prec, const void *a, const void *b, double *out,
                             void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !b || !out) {
        set_error("falm_dot_noalloc: null pointer");
        return FALM_ERR_INVALID;
    }
    Geom g = make_geom(dom);
    Tiling tl = make_tiling(interior_box(g));
    if (red_blocks(tl) > NOALLOC_PARTIALS) {  // cannot happen while the grid is capped at RED_BLOCKS
        set_error("falm_dot_noalloc: %d partial sums exceed the workspace of %d", red_blocks(tl), NOALLOC_PARTIALS);
        return FALM_ERR_UNSUPPORTED;
    }
    double *partials = nullptr;
    FALM_TRY(noalloc_setup(&partials));
    return launch_dot(dom, prec, a, b, 1, false, out, partials, (hipStream_t)stream);
}

falm_status falm_norm2(const falm_domain *dom, int prec, const void *a, int ncol, double *out, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !out || ncol < 1) {
        set_error("falm_norm2: bad argument");
        return FALM_ERR_INVALID;
    }
    return host_scalar(dom, prec, a, a, ncol, true, out, (hipStream_t)stream);
}

falm_status falm_reduce(const falm_domain *dom, int prec, int op, const void *a, int ncol, int col, double *out,
                        void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !out || op < 0 || op > FALM_RED_VEC_MIN || col < 0 || (op < FALM_RED_VEC_MAX && col >= ncol)) {
        set_error("falm_reduce: bad argument");
        return FALM_ERR_INVALID;
    }
    return host_minmax(dom, prec, op, a, ncol, col, out, (hipStream_t)stream);
}

falm_status falm_scale(int prec, void *a, size_t n, double scale, void *stream) {
    if (!a && n) return FALM_ERR_INVALID;
    if (!n) return FALM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32)
        hipLaunchKernelGGL(k_scale<float>, ew_grid(n), dim3(EW_BLOCK), 0, s, (float *)a, n, (float)scale);
    else
        hipLaunchKernelGGL(k_scale<double>, ew_grid(n), dim3(EW_BLOCK), 0, s, (double *)a, n, scale);
    FALM_CHECK_LAUNCH("k_scale");
    return FALM_OK;
}

falm_status falm_matrix_add(int prec, void *a, const void *b, size_t n, void *stream) {
    if ((!a || !b) && n) return FALM_ERR_INVALID;
    if (!n) return FALM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32)
        hipLaunchKernelGGL(k_add<float>, ew_grid(n), dim3(EW_BLOCK), 0, s, (float *)a, (const float *)b, n);
    else
        hipLaunchKernelGGL(k_add<double>, ew_grid(n), dim3(EW_BLOCK), 0, s, (double *)a, (const double *)b, n);
    FALM_CHECK_LAUNCH("k_add");
    return FALM_OK;
}

falm_status falm_vecaxby(int prec, double a, const void *x, double b, const void *y, void *result, size_t n,
                         void *stream) {
    if ((!x || !y || !result) && n) return FALM_ERR_INVALID;
    if (!n) return FALM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (prec == FALM_F32)
        hipLaunchKernelGGL(k_axpby<float>, ew_grid(n), dim3(EW_BLOCK), 0, s, (float)a, (const float *)x, (float)b,
                           (const float *)y, (float *)result, n);
    else
        hipLaunchKernelGGL(k_axpby<double>, ew_grid(n), dim3(EW_BLOCK), 0, s, a, (const double *)x, b,
                           (const double *)y, (double *)result, n);
    FALM_CHECK_LAUNCH("k_axpby");
    return FALM_OK;
}

falm_status falm_sor_sweep(const falm_domain *dom, int prec, const void *a, void *x, const void *b, double omega,
                           int color, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !x || !b || (color != 0 && color != 1)) {
        set_error("falm_sor_sweep: bad argument");
        return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? launch_sor<float>(dom, a, x, b, omega, color, nullptr, s)
                            : launch_sor<double>(dom, a, x, b, omega, color, nullptr, s);
}

falm_status falm_jacobi_sweep(const falm_domain *dom, int prec, const void *a, void *x, const void *xp,
                              const void *b, void *stream) {
    FALM_TRY(check_domain(dom));
    if (!a || !x || !xp || !b) {
        set_error("falm_jacobi_sweep: null pointer");
        return FALM_ERR_INVALID;
    }
    if (x == xp) {
        set_error("falm_jacobi_sweep: x aliases xp");
        return FALM_ERR_ALIAS;
    }
    hipStream_t s = (hipStream_t)stream;
    return prec == FALM_F32 ? launch_jacobi<float>(dom, a, x, xp, b, nullptr, s)
                            : launch_jacobi<double>(dom, a, x, xp, b, nullptr, s);
}

}  // extern "C"
