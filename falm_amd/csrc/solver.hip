// solver.hip — FalmEq on gfx950 (src/FalmEq.h:14-104, src/FalmEq.cpp:45-410).
//
// The whole Krylov loop is device resident: every scalar of BiCGStab
// (rho, alpha, omega, beta, err, it) lives in a DevState in device memory,
// reductions finish in the last workgroup of the kernel that produces them
// (fixed summation order, so results do not depend on scheduling), and
// every solver kernel returns immediately once `done` is set.  The host only
// enqueues iterations and polls the state every `check_every` iterations
// instead of copying partial sums back after every dot product.
//
// Fusions relative to the reference kernels:
//   residual + |r|^2 + rr = r copy + rho                      (init)
//   p = r + beta(p - omega q) + first preconditioner sweep from zero
//   q = A p^ + (rr, q) -> alpha
//   s = r - alpha q + first preconditioner sweep from zero
//   t = A s^ + (t, s), (t, t) -> omega
//   x += alpha p^ + omega s^, r = s - omega t, |r|^2, (r, rr) -> err, it, beta
//
// Mixed precision (FALM_SOLVE_MIXED): the Krylov vectors, the matrix copy and
// the preconditioner sweeps are float; dots and norms accumulate in double;
// x (the pressure) is updated in double; each pass restarts from the true
// residual b - A x computed in double with the double matrix, and the solve
// ends when that true residual meets the tolerance.
#include <cfloat>
#include <vector>

#include "linalg.hpp"

namespace falm {

struct DevState {
    double rho, rrho, alpha, omega, beta;
    double err;        // FalmEq::err of the current pass
    double tol_inner;  // the pass stops when err <= tol_inner
    double gN;         // global interior cell count
    int it_pass;       // iterations of the current pass
    int maxit_pass;
    int done;
    int status;
    int first;         // next p update is p = r
    int mixed;
    int pad[2];
};

constexpr int NT = 256;

template <class T>
__device__ inline T pc_first(int pc_type, T omega_pc, T b, T ac, bool mine) {
    if (pc_type == FALM_LS_SOR) {
        // kernel_Struct3d7p_SOR from x = 0: cc = (b - 0)/ac on the colour,
        // x = 0 + omega*cc (src/dev/FalmEqDevCall.cu:204-209)
        return mine ? T(0) + omega_pc * (b / ac) : T(0);
    } else if (pc_type == FALM_LS_JACOBI) {
        // kernel_Struct3d7p_Jacobi from xp = 0 (src/dev/FalmEqDevCall.cu:130)
        return T(0) + b / ac;
    }
    return b;  // identity (see falm_solver_create)
}

// residual in double + float/double Krylov copies + |r|^2 and (r, rr)
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_init(Tiling tl, Geom g, Mat7<double> a, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *r64, T *r, T *rr,
                                                  DevState *st, double *partials, unsigned *ticket) {
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k;
    double v[2] = {0.0, 0.0};
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        const double rv = b[cc] - apply7<double, double, double>(a, x, cc, g.nx, g.nxy);
        r64[cc] = rv;
        const T rt = (T)rv;
        if ((void *)r != (void *)r64) r[cc] = rt;
        rr[cc] = rt;
        v[0] = rv * rv;
        v[1] = (double)rt * (double)rt;
    }
    double tot[2];
    if (grid_sum_last<NT, 2>(v, partials, ticket, tot, lds) && threadIdx.x == 0) {
        st->err = sqrt(tot[0]) / st->gN;
        st->rho = tot[1];
        st->rrho = 1.0;
        st->alpha = 0.0;
        st->omega = 1.0;
        st->beta = 0.0;
        st->it_pass = 0;
        st->first = 1;
        st->done = 0;
        st->status = FALM_OK;
        if (fabs(st->rho) < (double)FLT_MIN) {  // src/FalmEq.cpp:345-351
            if (!st->mixed) st->err = st->rho;
            st->status = FALM_ERR_BREAKDOWN;
            st->done = 1;
        }
    }
}

// p = r (first) or r + beta(p - omega q) (kernel_PBiCGStab_1), then the
// first preconditioner sweep pp = M0(p)
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_p(Tiling tl, Geom g, const T *__restrict__ ac, const T *__restrict__ r,
                                               T *__restrict__ p, const T *__restrict__ q, T *__restrict__ pp,
                                               const DevState *st, int pc_type, T omega_pc) {
    if (st->done) return;
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int cc = g.id(i, j, k);
    T pv;
    if (st->first) {
        pv = r[cc];
    } else {
        const T beta = (T)st->beta, omega = (T)st->omega;
        pv = r[cc] + beta * (p[cc] - omega * q[cc]);
    }
    p[cc] = pv;
    pp[cc] = pc_first<T>(pc_type, omega_pc, pv, ac[cc], ((i + j + k + g.osum) & 1) == FALM_COLOR_BLACK);
}

// s = r - alpha q (kernel_PBiCGStab_2), then ss = M0(s)
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_s(Tiling tl, Geom g, const T *__restrict__ ac, const T *__restrict__ r,
                                               const T *__restrict__ q, T *__restrict__ s, T *__restrict__ ss,
                                               const DevState *st, int pc_type, T omega_pc) {
    if (st->done) return;
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int cc = g.id(i, j, k);
    const T alpha = (T)st->alpha;
    const T sv = r[cc] - alpha * q[cc];
    s[cc] = sv;
    ss[cc] = pc_first<T>(pc_type, omega_pc, sv, ac[cc], ((i + j + k + g.osum) & 1) == FALM_COLOR_BLACK);
}

// q = A pp ; alpha = rho / (rr, q)
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_mvdot(Tiling tl, Geom g, Mat7<T> a, const T *__restrict__ pp,
                                                   T *__restrict__ q, const T *__restrict__ rr, DevState *st,
                                                   double *partials, unsigned *ticket) {
    if (st->done) return;
    __shared__ double lds[NT / 64];
    int i, j, k;
    double v[1] = {0.0};
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        const T qv = apply7<T, T, T>(a, pp, cc, g.nx, g.nxy);
        q[cc] = qv;
        v[0] = (double)rr[cc] * (double)qv;
    }
    double tot[1];
    if (grid_sum_last<NT, 1>(v, partials, ticket, tot, lds) && threadIdx.x == 0) st->alpha = st->rho / tot[0];
}

// t = A ss ; omega = (t, s) / (t, t)
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_mvdot2(Tiling tl, Geom g, Mat7<T> a, const T *__restrict__ ss,
                                                    T *__restrict__ t, const T *__restrict__ s, DevState *st,
                                                    double *partials, unsigned *ticket) {
    if (st->done) return;
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k;
    double v[2] = {0.0, 0.0};
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        const T tv = apply7<T, T, T>(a, ss, cc, g.nx, g.nxy);
        t[cc] = tv;
        v[0] = (double)tv * (double)s[cc];
        v[1] = (double)tv * (double)tv;
    }
    double tot[2];
    if (grid_sum_last<NT, 2>(v, partials, ticket, tot, lds) && threadIdx.x == 0) st->omega = tot[0] / tot[1];
}

// x += alpha pp + omega ss (kernel_PBiCGStab_3, x in double);
// r = s - omega t (kernel_PBiCGStab_4); err, it, next rho and beta
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_update(Tiling tl, Geom g, double *__restrict__ x,
                                                    const T *__restrict__ pp, const T *__restrict__ ss,
                                                    const T *__restrict__ s, const T *__restrict__ t,
                                                    T *__restrict__ r, const T *__restrict__ rr, DevState *st,
                                                    double *partials, unsigned *ticket) {
    if (st->done) return;
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k;
    double v[2] = {0.0, 0.0};
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        const double alpha = st->alpha, omega = st->omega;
        x[cc] += alpha * (double)pp[cc] + omega * (double)ss[cc];
        const T rv = s[cc] - (T)omega * t[cc];
        r[cc] = rv;
        v[0] = (double)rv * (double)rv;
        v[1] = (double)rv * (double)rr[cc];
    }
    double tot[2];
    if (grid_sum_last<NT, 2>(v, partials, ticket, tot, lds) && threadIdx.x == 0) {
        st->rrho = st->rho;
        st->err = sqrt(tot[0]) / st->gN;
        st->it_pass += 1;
        st->first = 0;
        if (!(st->it_pass < st->maxit_pass && st->err > st->tol_inner)) {
            st->done = 1;
        } else {
            st->rho = tot[1];
            if (fabs(st->rho) < (double)FLT_MIN) {
                if (!st->mixed) st->err = st->rho;
                st->status = FALM_ERR_BREAKDOWN;
                st->done = 1;
            } else {
                st->beta = (st->rho * st->alpha) / (st->rrho * st->omega);
            }
        }
    }
}

// plain SOR / Jacobi solvers: residual norm -> err, it, done
__global__ __launch_bounds__(NT) void k_res_norm(Tiling tl, Geom g, Mat7<double> a, const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ r, DevState *st,
                                                 double *partials, unsigned *ticket) {
    if (st->done) return;
    __shared__ double lds[NT / 64];
    int i, j, k;
    double v[1] = {0.0};
    if (tile_cell(tl, i, j, k)) {
        const int cc = g.id(i, j, k);
        const double rv = b[cc] - apply7<double, double, double>(a, x, cc, g.nx, g.nxy);
        r[cc] = rv;
        v[0] = rv * rv;
    }
    double tot[1];
    if (grid_sum_last<NT, 1>(v, partials, ticket, tot, lds) && threadIdx.x == 0) {
        st->err = sqrt(tot[0]) / st->gN;
        st->it_pass += 1;
        if (!(st->it_pass < st->maxit_pass && st->err > st->tol_inner)) st->done = 1;
    }
}

template <class T>
__global__ __launch_bounds__(NT) void k_copy_interior(Tiling tl, Geom g, const T *__restrict__ src,
                                                      T *__restrict__ dst, const DevState *st) {
    if (st && st->done) return;
    int i, j, k;
    if (!tile_cell(tl, i, j, k)) return;
    const int cc = g.id(i, j, k);
    dst[cc] = src[cc];
}

template <class TS, class TD>
__global__ void k_convert(const TS *__restrict__ s, TD *__restrict__ d, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        d[t] = (TD)s[t];
}

// limits of the next pass (the init kernel leaves them alone)
__global__ void k_state_arm(DevState *st, double tol, int maxit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->tol_inner = tol;
        st->maxit_pass = maxit;
    }
}

__global__ void k_state_reset(DevState *st, double gN, int mixed, double tol, int maxit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->gN = gN;
        st->mixed = mixed;
        st->done = 0;
        st->status = FALM_OK;
        st->it_pass = 0;
        st->err = 0.0;
        st->tol_inner = tol;
        st->maxit_pass = maxit;
    }
}

}  // namespace falm

using namespace falm;

struct falm_solver {
    falm_domain dom;
    falm_solver_config cfg;
    size_t n;
    Geom g;
    Tiling tl;
    DevState *st = nullptr;
    DevState *hst = nullptr;  // pinned mirror
    double *partials = nullptr;
    unsigned *ticket = nullptr;
    // Krylov vectors: double (F64) or float (MIXED)
    void *rr = nullptr, *p = nullptr, *q = nullptr, *s = nullptr, *pp = nullptr, *ss = nullptr, *t = nullptr,
         *w = nullptr, *r32 = nullptr;
    float *a32 = nullptr;
    const void *a32_src = nullptr;
    bool a32_pinned = false;
};

namespace falm {

static falm_status solver_alloc(falm_solver *S) {
    const size_t vb = S->n * (S->cfg.precision == FALM_SOLVE_MIXED ? sizeof(float) : sizeof(double));
    void **vecs[] = {&S->rr, &S->p, &S->q, &S->s, &S->pp, &S->ss, &S->t, &S->w};
    for (void **v : vecs) {
        FALM_HIP(hipMalloc(v, vb));
        FALM_HIP(hipMemset(*v, 0, vb));
    }
    if (S->cfg.precision == FALM_SOLVE_MIXED) {
        FALM_HIP(hipMalloc(&S->r32, vb));
        FALM_HIP(hipMemset(S->r32, 0, vb));
        FALM_HIP(hipMalloc(&S->a32, 7 * S->n * sizeof(float)));
        FALM_HIP(hipMemset(S->a32, 0, 7 * S->n * sizeof(float)));
    }
    FALM_HIP(hipMalloc(&S->st, sizeof(DevState)));
    FALM_HIP(hipMemset(S->st, 0, sizeof(DevState)));
    FALM_HIP(hipHostMalloc(&S->hst, sizeof(DevState), hipHostMallocDefault));
    size_t nb = (size_t)(S->tl.tiles > 0 ? S->tl.tiles : 1);
    FALM_HIP(hipMalloc(&S->partials, 2 * nb * sizeof(double)));
    FALM_HIP(hipMalloc(&S->ticket, 64));
    FALM_HIP(hipMemset(S->ticket, 0, 64));
    return FALM_OK;
}

static falm_status read_state(falm_solver *S, hipStream_t s) {
    FALM_HIP(hipMemcpyAsync(S->hst, S->st, sizeof(DevState), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

static int check_period(falm_solver *S) {
    if (S->cfg.check_every > 0) return S->cfg.check_every;
    double n = (double)S->tl.box.size();
    int c = (int)(4.0e6 / (n > 1 ? n : 1));
    return c < 1 ? 1 : (c > 32 ? 32 : c);
}

// preconditioner sweeps after the fused first one (SORPC: pc_maxit x
// {black, red}; JacobiPC: pc_maxit sweeps; both run at least once,
// src/FalmEq.cpp:147-170,258-302)
template <class T>
static falm_status pc_rest(falm_solver *S, const void *a, void *x, const void *b, hipStream_t s) {
    const falm_solver_config &c = S->cfg;
    const int m = c.pc_maxit < 1 ? 1 : c.pc_maxit;
    const int *done = &S->st->done;
    if (c.pc_type == FALM_LS_SOR) {
        FALM_TRY(launch_sor<T>(&S->dom, a, x, b, c.pc_relax, FALM_COLOR_RED, done, s));
        for (int it = 1; it < m; it++) {
            FALM_TRY(launch_sor<T>(&S->dom, a, x, b, c.pc_relax, FALM_COLOR_BLACK, done, s));
            FALM_TRY(launch_sor<T>(&S->dom, a, x, b, c.pc_relax, FALM_COLOR_RED, done, s));
        }
    } else if (c.pc_type == FALM_LS_JACOBI) {
        for (int it = 1; it < m; it++) {
            hipLaunchKernelGGL((k_copy_interior<T>), tiling_grid(S->tl), tiling_block(), 0, s, S->tl, S->g,
                               (const T *)x, (T *)S->w, (const DevState *)S->st);
            FALM_CHECK_LAUNCH("k_copy_interior");
            FALM_TRY(launch_jacobi<T>(&S->dom, a, x, S->w, b, done, s));
        }
    }
    return FALM_OK;
}

// one BiCGStab iteration (src/FalmEq.cpp:345-407)
template <class T>
static falm_status bicg_iteration(falm_solver *S, const void *aT, double *x, T *r, hipStream_t s) {
    const falm_solver_config &c = S->cfg;
    const Tiling &tl = S->tl;
    const Geom &g = S->g;
    const T *ac = (const T *)aT;
    T *rr = (T *)S->rr, *p = (T *)S->p, *q = (T *)S->q, *sv = (T *)S->s, *pp = (T *)S->pp, *ss = (T *)S->ss,
      *t = (T *)S->t;
    const T wpc = (T)c.pc_relax;
    hipLaunchKernelGGL((k_bicg_p<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, ac, r, p, q, pp,
                       (const DevState *)S->st, c.pc_type, wpc);
    FALM_CHECK_LAUNCH("k_bicg_p");
    FALM_TRY(pc_rest<T>(S, aT, pp, p, s));
    TimedLaunch tm1(FALM_TIME_MV, s);
    hipLaunchKernelGGL((k_bicg_mvdot<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, mat7<T>(aT, S->n),
                       (const T *)pp, q, (const T *)rr, S->st, S->partials, S->ticket);
    FALM_CHECK_LAUNCH("k_bicg_mvdot");
    tm1.done((double)tl.box.size() * 10.0 * sizeof(T));
    hipLaunchKernelGGL((k_bicg_s<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, ac, (const T *)r,
                       (const T *)q, sv, ss, (const DevState *)S->st, c.pc_type, wpc);
    FALM_CHECK_LAUNCH("k_bicg_s");
    FALM_TRY(pc_rest<T>(S, aT, ss, sv, s));
    hipLaunchKernelGGL((k_bicg_mvdot2<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, mat7<T>(aT, S->n),
                       (const T *)ss, t, (const T *)sv, S->st, S->partials, S->ticket);
    FALM_CHECK_LAUNCH("k_bicg_mvdot2");
    TimedLaunch tm2(FALM_TIME_SOLVER_UPDATE, s);
    hipLaunchKernelGGL((k_bicg_update<T>), tiling_grid(tl), tiling_block(), 0, s, tl, g, x, (const T *)pp,
                       (const T *)ss, (const T *)sv, (const T *)t, r, (const T *)rr, S->st, S->partials, S->ticket);
    FALM_CHECK_LAUNCH("k_bicg_update");
    tm2.done((double)tl.box.size() * (2.0 * sizeof(double) + 6.0 * sizeof(T)));
    return FALM_OK;
}

// run iterations until the device says done; returns iterations of the pass
template <class T>
static falm_status bicg_pass(falm_solver *S, const void *aT, double *x, T *r, int maxit_pass, hipStream_t s) {
    const int period = check_period(S);
    int enqueued = 0;
    while (true) {
        for (int c = 0; c < period && enqueued < maxit_pass; c++, enqueued++)
            FALM_TRY(bicg_iteration<T>(S, aT, x, r, s));
        FALM_TRY(read_state(S, s));
        if (S->hst->done || enqueued >= maxit_pass) break;
    }
    return FALM_OK;
}

static falm_status launch_init(falm_solver *S, const void *a, const double *x, const double *b, double *r64,
                               void *rT, bool f32, hipStream_t s) {
    if (f32)
        hipLaunchKernelGGL((k_bicg_init<float>), tiling_grid(S->tl), tiling_block(), 0, s, S->tl, S->g,
                           mat7<double>(a, S->n), x, b, r64, (float *)rT, (float *)S->rr, S->st, S->partials,
                           S->ticket);
    else
        hipLaunchKernelGGL((k_bicg_init<double>), tiling_grid(S->tl), tiling_block(), 0, s, S->tl, S->g,
                           mat7<double>(a, S->n), x, b, r64, (double *)rT, (double *)S->rr, S->st, S->partials,
                           S->ticket);
    FALM_CHECK_LAUNCH("k_bicg_init");
    return FALM_OK;
}

static double interior_count(const falm_solver *S) {
    return (double)(S->dom.gshape[0] - 2 * S->dom.gc) * (double)(S->dom.gshape[1] - 2 * S->dom.gc) *
           (double)(S->dom.gshape[2] - 2 * S->dom.gc);
}

static falm_status solve_bicgstab_f64(falm_solver *S, const void *a, double *x, const double *b, double *r,
                                      hipStream_t s, falm_solve_info *info) {
    const falm_solver_config &c = S->cfg;
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, interior_count(S), 0, c.tol, c.maxit);
    FALM_CHECK_LAUNCH("k_state_reset");
    FALM_TRY(launch_init(S, a, x, b, r, r, false, s));
    FALM_TRY(bicg_pass<double>(S, a, x, r, c.maxit, s));
    info->it = S->hst->it_pass;
    info->err = S->hst->err;
    info->outer_it = 1;
    info->status = S->hst->status;
    if (info->status == FALM_OK && !(S->hst->err <= c.tol)) info->status = FALM_ERR_NOT_CONVERGED;
    return FALM_OK;
}

static falm_status solve_bicgstab_mixed(falm_solver *S, const void *a, double *x, const double *b, double *r,
                                        hipStream_t s, falm_solve_info *info) {
    const falm_solver_config &c = S->cfg;
    const double gN = interior_count(S);
    const int inner_maxit = c.inner_maxit > 0 ? c.inner_maxit : 50;
    const double inner_rtol = c.inner_rtol > 0 ? c.inner_rtol : 1e-4;
    if (!S->a32_pinned || S->a32_src != a) {
        hipLaunchKernelGGL((k_convert<double, float>), ew_grid(7 * S->n), dim3(EW_BLOCK), 0, s, (const double *)a,
                           S->a32, 7 * S->n);
        FALM_CHECK_LAUNCH("k_convert");
    }
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, gN, 1, c.tol, c.maxit);
    FALM_CHECK_LAUNCH("k_state_reset");
    int total = 0, pass = 0;
    double err = 0.0;
    int status = FALM_OK;
    while (true) {
        // true residual in double; the inner target follows from it
        FALM_TRY(launch_init(S, a, x, b, r, S->r32, true, s));
        FALM_TRY(read_state(S, s));
        err = S->hst->err;
        if ((pass > 0 && err <= c.tol) || total >= c.maxit) break;
        if (S->hst->done) {  // r32 vanished: nothing more a float pass can do
            status = (err <= c.tol) ? FALM_OK : FALM_ERR_BREAKDOWN;
            break;
        }
        double target = inner_rtol * err;
        if (target < 0.5 * c.tol) target = 0.5 * c.tol;
        int maxit_pass = c.maxit - total < inner_maxit ? c.maxit - total : inner_maxit;
        hipLaunchKernelGGL(k_state_arm, dim3(1), dim3(64), 0, s, S->st, target, maxit_pass);
        FALM_CHECK_LAUNCH("k_state_arm");
        FALM_TRY(bicg_pass<float>(S, S->a32, x, (float *)S->r32, maxit_pass, s));
        total += S->hst->it_pass;
        pass++;
        if (S->hst->it_pass == 0) {  // breakdown at the start of the pass
            status = FALM_ERR_BREAKDOWN;
            break;
        }
    }
    info->it = total;
    info->err = err;
    info->outer_it = pass;
    info->status = status != FALM_OK ? status : (err <= c.tol ? FALM_OK : FALM_ERR_NOT_CONVERGED);
    return FALM_OK;
}

// FalmEq::SOR (src/FalmEq.cpp:173-240) and FalmEq::Jacobi (:81-127)
static falm_status solve_stationary(falm_solver *S, const void *a, double *x, const double *b, double *r,
                                    hipStream_t s, falm_solve_info *info) {
    const falm_solver_config &c = S->cfg;
    const double gN = interior_count(S);
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, gN, 0, c.tol, c.maxit);
    FALM_CHECK_LAUNCH("k_state_reset");
    const int period = check_period(S);
    const int *done = &S->st->done;
    int enqueued = 0;
    while (true) {
        for (int k = 0; k < period && enqueued < c.maxit; k++, enqueued++) {
            if (c.type == FALM_LS_SOR) {
                FALM_TRY(launch_sor<double>(&S->dom, a, x, b, c.relax, FALM_COLOR_BLACK, done, s));
                FALM_TRY(launch_sor<double>(&S->dom, a, x, b, c.relax, FALM_COLOR_RED, done, s));
            } else {
                hipLaunchKernelGGL((k_copy_interior<double>), tiling_grid(S->tl), tiling_block(), 0, s, S->tl, S->g,
                                   (const double *)x, (double *)S->w, (const DevState *)S->st);
                FALM_TRY(launch_jacobi<double>(&S->dom, a, x, S->w, b, done, s));
            }
            hipLaunchKernelGGL(k_res_norm, tiling_grid(S->tl), tiling_block(), 0, s, S->tl, S->g,
                               mat7<double>(a, S->n), (const double *)x, b, r, S->st, S->partials, S->ticket);
            FALM_CHECK_LAUNCH("k_res_norm");
        }
        FALM_TRY(read_state(S, s));
        if (S->hst->done || enqueued >= c.maxit) break;
    }
    info->it = S->hst->it_pass;
    info->err = S->hst->err;
    info->outer_it = 1;
    info->status = S->hst->err <= c.tol ? FALM_OK : FALM_ERR_NOT_CONVERGED;
    return FALM_OK;
}

}  // namespace falm

extern "C" {

falm_status falm_solver_create(const falm_domain *dom, const falm_solver_config *cfg, falm_solver **out) {
    FALM_TRY(check_domain(dom));
    if (!cfg || !out) return FALM_ERR_INVALID;
    *out = nullptr;
    if (cfg->type != FALM_LS_PBICGSTAB && cfg->type != FALM_LS_SOR && cfg->type != FALM_LS_JACOBI) {
        set_error("falm_solver_create: solver type %d not supported", cfg->type);
        return FALM_ERR_UNSUPPORTED;
    }
    if (cfg->precision != FALM_SOLVE_F64 && cfg->precision != FALM_SOLVE_MIXED) {
        set_error("falm_solver_create: precision %d not supported", cfg->precision);
        return FALM_ERR_UNSUPPORTED;
    }
    if (cfg->precision == FALM_SOLVE_MIXED && cfg->type != FALM_LS_PBICGSTAB) {
        set_error("falm_solver_create: mixed precision needs the BiCGStab solver");
        return FALM_ERR_UNSUPPORTED;
    }
    if (cfg->maxit < 1) {
        set_error("falm_solver_create: maxit must be >= 1");
        return FALM_ERR_INVALID;
    }
    falm_solver *S = new falm_solver();
    S->dom = *dom;
    S->cfg = *cfg;
    S->n = ncell_of(dom);
    S->g = make_geom(dom);
    S->tl = make_tiling(interior_box(S->g));
    falm_status st = solver_alloc(S);
    if (st != FALM_OK) {
        falm_solver_destroy(S);
        return st;
    }
    *out = S;
    return FALM_OK;
}

falm_status falm_solver_set_matrix(falm_solver *S, const void *a, void *stream) {
    if (!S || !a) return FALM_ERR_INVALID;
    if (S->cfg.precision != FALM_SOLVE_MIXED) return FALM_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL((k_convert<double, float>), ew_grid(7 * S->n), dim3(EW_BLOCK), 0, s, (const double *)a, S->a32,
                       7 * S->n);
    FALM_CHECK_LAUNCH("k_convert");
    S->a32_src = a;
    S->a32_pinned = true;
    return FALM_OK;
}

falm_status falm_solver_solve(falm_solver *S, const void *a, void *x, const void *b, void *r, void *stream,
                              falm_solve_info *info) {
    if (!S || !a || !x || !b || !r) {
        set_error("falm_solver_solve: null argument");
        return FALM_ERR_INVALID;
    }
    if (x == r || x == b || r == b) {
        set_error("falm_solver_solve: x, b and r must be distinct");
        return FALM_ERR_ALIAS;
    }
    falm_solve_info tmp;
    if (!info) info = &tmp;
    hipStream_t s = (hipStream_t)stream;
    if (S->cfg.type == FALM_LS_PBICGSTAB) {
        if (S->cfg.precision == FALM_SOLVE_MIXED)
            return solve_bicgstab_mixed(S, a, (double *)x, (const double *)b, (double *)r, s, info);
        return solve_bicgstab_f64(S, a, (double *)x, (const double *)b, (double *)r, s, info);
    }
    return solve_stationary(S, a, (double *)x, (const double *)b, (double *)r, s, info);
}

falm_status falm_solver_destroy(falm_solver *S) {
    if (!S) return FALM_OK;
    void *bufs[] = {S->rr, S->p, S->q, S->s, S->pp, S->ss, S->t, S->w, S->r32, S->a32,
                    S->st, S->partials, S->ticket};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (S->hst) (void)hipHostFree(S->hst);
    delete S;
    return FALM_OK;
}

}  // extern "C"
