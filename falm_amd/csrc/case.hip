// case.hip — the per-step orchestration of the reference apps
// (app/nedo/main.cpp:150-204, app/windtunnel/main.cpp:94-165,
// app/lid3d/main.cpp:94-175, app/ld2d/main.cpp:70-136) with all state on
// the device.  Pseudo-velocity is computed out of place from the u_previous
// snapshot the apps already take (the apps pass ua == u, a data race in the
// reference kernel).
//
// Precision: FALM_CASE_F64 keeps everything in double.  FALM_CASE_MIXED keeps
// velocity, face flux, eddy viscosity, force, divergence and metrics in
// float (all CFD stencils run in float) while the pressure, the Poisson
// right-hand side, the Poisson matrix and the residual stay double and the
// Poisson solve runs as float Krylov passes inside a double
// defect-correction loop.
//
// The stencils take their metrics from the 1-D cell widths of the mesh
// (bit-identical to the per-cell metric fields, which are still kept for
// falm_case_get_field), and the velocity snapshot is a buffer swap plus a
// copy of the ghost shell instead of a full copy.
#include <cmath>
#include <utility>
#include <vector>

#include "alm.hpp"
#include "case.hpp"
#include "cfd.hpp"
#include "linalg.hpp"

namespace falm {
falm_status launch_bc_velocity(const falm_domain *dom, int prec, const falm_face_bc faces[6], void *u,
                               const void *uprev, const void *xyz, double dt, const falm_inlet_noise *noise,
                               hipStream_t s, const int *gate);
falm_status launch_bc_pressure(const falm_domain *dom, int prec, const int faces[6], void *p, hipStream_t s,
                               const int *gate);
falm_status launch_bc_face_flux(const falm_domain *dom, int prec, const int faces[6], void *uu, hipStream_t s);
falm_status launch_copy_z5(const falm_domain *dom, int prec, void *v, int ncol, hipStream_t s, const int *gate);
}  // namespace falm

using namespace falm;

namespace falm {

falm_status case_dalloc(falm_case *C, void **p, size_t bytes) {
    FALM_HIP(hipMalloc(p, bytes));
    C->bufs.push_back(*p);
    FALM_HIP(hipMemset(*p, 0, bytes));
    return FALM_OK;
}

static void case_faces(falm_case *C) {
    const falm_case_config &c = C->cfg;
    for (int f = 0; f < 6; f++) {
        C->ubc[f].type = FALM_BC_NONE;
        C->ubc[f].value[0] = C->ubc[f].value[1] = C->ubc[f].value[2] = 0.0;
        C->pbc[f] = FALM_PBC_NONE;
        C->uubc[f] = 0;
    }
    if (c.kind == FALM_CASE_WINDTUNNEL) {
        // app/windtunnel/bc.h:6-68, bcdevcall.cu:8-211 (z- slip on request)
        C->ubc[0].type = FALM_BC_INLET;
        C->ubc[0].value[0] = c.u_inflow;
        C->ubc[1].type = FALM_BC_CONVECTIVE;
        C->ubc[2].type = FALM_BC_SLIP;
        C->ubc[3].type = FALM_BC_SLIP;
        C->ubc[4].type = c.zminus_slip ? FALM_BC_SLIP : FALM_BC_WALL0;
        C->ubc[5].type = FALM_BC_SLIP;
        for (int f = 0; f < 6; f++) C->pbc[f] = FALM_PBC_NEUMANN;
        C->pbc[1] = FALM_PBC_ZERO;
    } else {
        // app/lid3d/bcdev.cu:11-357 and app/ld2d/bcdevcall.cu: no-slip
        // reflection, lid on y+ moving in +x, Neumann p, zero wall flux
        const int nf = c.kind == FALM_CASE_LID2D ? 4 : 6;
        for (int f = 0; f < nf; f++) {
            C->ubc[f].type = FALM_BC_REFLECT;
            C->pbc[f] = FALM_PBC_NEUMANN;
            C->uubc[f] = 1;
        }
        C->ubc[3].value[0] = c.lid_velocity;
    }
}

template <class T>
__global__ void k_fill3(T *u, size_t n, T a, T b, T c) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        u[t] = a;
        u[t + n] = b;
        u[t + 2 * n] = c;
    }
}

// Copy the ghost shell (every padded cell outside the interior) of an
// ncol-column field.  Cells on shell edges belong to two slabs and are
// copied twice with the same value.
template <class T>
__global__ void k_copy_shell(Geom g, const T *__restrict__ src, T *__restrict__ dst, size_t n, int ncol) {
    const int G = 2 * g.gc;
    const long long cx = (long long)G * g.ny * g.nz, cy = (long long)G * g.nx * g.nz;
    const long long cz = (long long)G * g.nx * g.ny;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < cx + cy + cz;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        long long r = t;
        if (r < cx) {
            const int q = (int)(r % G);
            const long long rest = r / G;
            j = (int)(rest % g.ny);
            k = (int)(rest / g.ny);
            i = q < g.gc ? q : g.nx - G + q;
        } else if ((r -= cx) < cy) {
            i = (int)(r % g.nx);
            const long long rest = r / g.nx;
            const int q = (int)(rest % G);
            k = (int)(rest / G);
            j = q < g.gc ? q : g.ny - G + q;
        } else {
            r -= cy;
            i = (int)(r % g.nx);
            const long long rest = r / g.nx;
            j = (int)(rest % g.ny);
            const int q = (int)(rest / g.ny);
            k = q < g.gc ? q : g.nz - G + q;
        }
        const size_t c = (size_t)g.id(i, j, k);
        for (int m = 0; m < ncol; m++) dst[c + m * n] = src[c + m * n];
    }
}

static falm_status copy_shell(falm_case *C, const void *src, void *dst, hipStream_t s) {
    const Geom g = make_geom(&C->dom);
    const int G = 2 * g.gc;
    const size_t total = (size_t)G * ((size_t)g.ny * g.nz + (size_t)g.nx * g.nz + (size_t)g.nx * g.ny);
    if (C->prec == FALM_F32)
        hipLaunchKernelGGL((k_copy_shell<float>), ew_grid(total), dim3(EW_BLOCK), 0, s, g, (const float *)src,
                           (float *)dst, C->n, 3);
    else
        hipLaunchKernelGGL((k_copy_shell<double>), ew_grid(total), dim3(EW_BLOCK), 0, s, g, (const double *)src,
                           (double *)dst, C->n, 3);
    FALM_CHECK_LAUNCH("k_copy_shell");
    return FALM_OK;
}

static falm_status case_init(falm_case *C) {
    const falm_case_config &c = C->cfg;
    hipStream_t s = C->stream;
    const int gc = 2;
    for (int m = 0; m < 3; m++) {
        if (c.n[m] < 1 || (!c.cv[2 * m] && !(c.h[m] > 0.0)) || (!c.cv[2 * m]) != (!c.cv[2 * m + 1])) {
            set_error("falm_case_create: bad grid");
            return FALM_ERR_INVALID;
        }
        C->dom.shape[m] = c.n[m] + 2 * gc;
        C->dom.offset[m] = 0;
        C->dom.gshape[m] = C->dom.shape[m];
    }
    C->dom.gc = gc;
    if (c.kind == FALM_CASE_LID2D && c.n[2] != 1) {
        set_error("falm_case_create: the 2-D cavity needs n[2] == 1");
        return FALM_ERR_INVALID;
    }
    FALM_TRY(check_domain(&C->dom));
    C->n = ncell_of(&C->dom);
    C->prec = c.precision == FALM_CASE_MIXED ? FALM_F32 : FALM_F64;
    C->tb = prec_bytes(C->prec);
    const size_t n = C->n, tb = C->tb;
    FALM_TRY(case_dalloc(C, &C->u, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->uprev, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->uu, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->nut, n * tb));
    FALM_TRY(case_dalloc(C, &C->ff, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->div, n * tb));
    FALM_TRY(case_dalloc(C, &C->xyz, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->kx, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->g, 3 * n * tb));
    FALM_TRY(case_dalloc(C, &C->ja, n * tb));
    FALM_TRY(case_dalloc(C, (void **)&C->p, n * sizeof(double)));
    FALM_TRY(case_dalloc(C, (void **)&C->rhs, n * sizeof(double)));
    FALM_TRY(case_dalloc(C, (void **)&C->a, 7 * n * sizeof(double)));
    FALM_TRY(case_dalloc(C, (void **)&C->dnorm, 8 * sizeof(double)));
    Tiling tl = make_tiling(interior_box(make_geom(&C->dom)));
    FALM_TRY(case_dalloc(C, (void **)&C->partials, (size_t)(tl.tiles + 1) * sizeof(double)));
    // the zero fills ran on the null stream, which the case's non-blocking
    // stream does not wait for
    FALM_HIP(hipDeviceSynchronize());
    // coherent: the norm's finalize kernel writes it directly
    FALM_HIP(hipHostMalloc(&C->hnorm, sizeof(double), hipHostMallocCoherent));
    {
        const int empty[6] = {1, 0, 1, 0, 1, 0};
        FALM_TRY(case_dalloc(C, (void **)&C->nobox, sizeof(empty)));
        FALM_HIP(hipMemcpy(C->nobox, empty, sizeof(empty), hipMemcpyHostToDevice));
    }
    // uniform control-volume mesh: cell centres and widths of every padded
    // cell (what Mesher::build_mesh + readControlVolumeFile produce for a
    // uniform node list, src/mesher/mesher.hpp, src/FalmIO.h:99-116)
    std::vector<double> *cen[3] = {&C->x, &C->y, &C->z}, *wid[3] = {&C->hx, &C->hy, &C->hz};
    for (int m = 0; m < 3; m++) {
        cen[m]->resize(C->dom.shape[m]);
        wid[m]->resize(C->dom.shape[m]);
        for (int i = 0; i < C->dom.shape[m]; i++) {
            (*cen[m])[i] = c.cv[2 * m] ? c.cv[2 * m][i] : c.origin[m] + (i - gc + 0.5) * c.h[m];
            (*wid[m])[i] = c.cv[2 * m] ? c.cv[2 * m + 1][i] : c.h[m];
            if (!((*wid[m])[i] > 0.0)) {
                set_error("falm_case_create: control-volume widths must be positive");
                return FALM_ERR_INVALID;
            }
        }
    }
    FALM_TRY(falm_metrics(&C->dom, C->prec, C->x.data(), C->y.data(), C->z.data(), C->hx.data(), C->hy.data(),
                          C->hz.data(), C->xyz, C->kx, C->g, C->ja, s));
    for (int m = 0; m < 3; m++) {
        const int L = C->dom.shape[m];
        std::vector<double> r(L);
        for (int i = 0; i < L; i++) r[i] = 1.0 / (*wid[m])[i];  // Real(1)/pitch, src/falm.h:395
        double *dh = nullptr, *dr = nullptr;
        FALM_TRY(case_dalloc(C, (void **)&dh, L * sizeof(double)));
        FALM_TRY(case_dalloc(C, (void **)&dr, L * sizeof(double)));
        FALM_HIP(hipMemcpy(dh, wid[m]->data(), L * sizeof(double), hipMemcpyHostToDevice));
        FALM_HIP(hipMemcpy(dr, r.data(), L * sizeof(double), hipMemcpyHostToDevice));
        C->sm.h[m] = dh;
        C->sm.rh[m] = dr;
        std::vector<double> tab[SEP_NTAB];
        sep_tables(wid[m]->data(), L, tab);
        for (int q = 0; q < SEP_NTAB; q++) {
            std::vector<float> tf(tab[q].begin(), tab[q].end());
            double *td = nullptr;
            float *tff = nullptr;
            FALM_TRY(case_dalloc(C, (void **)&td, L * sizeof(double)));
            FALM_TRY(case_dalloc(C, (void **)&tff, L * sizeof(float)));
            FALM_HIP(hipMemcpy(td, tab[q].data(), L * sizeof(double), hipMemcpyHostToDevice));
            FALM_HIP(hipMemcpy(tff, tf.data(), L * sizeof(float), hipMemcpyHostToDevice));
            C->sm.td[q][m] = td;
            C->sm.tf[q][m] = tff;
        }
    }
    FALM_HIP(hipDeviceSynchronize());
    // Poisson matrix from double metrics
    int kind = c.kind == FALM_CASE_WINDTUNNEL ? FALM_POISSON_OUTLET_XPLUS
             : c.kind == FALM_CASE_LID2D     ? FALM_POISSON_2D
                                             : FALM_POISSON_NEUMANN;
    if (C->prec == FALM_F64) {
        FALM_TRY(falm_poisson_matrix(&C->dom, FALM_F64, kind, C->g, C->ja, C->a, &C->maxdiag, s));
    } else {
        void *x64, *k64, *g64, *j64;
        FALM_HIP(hipMalloc(&x64, 3 * n * 8));
        FALM_HIP(hipMalloc(&k64, 3 * n * 8));
        FALM_HIP(hipMalloc(&g64, 3 * n * 8));
        FALM_HIP(hipMalloc(&j64, n * 8));
        falm_status st = falm_metrics(&C->dom, FALM_F64, C->x.data(), C->y.data(), C->z.data(), C->hx.data(),
                                      C->hy.data(), C->hz.data(), x64, k64, g64, j64, s);
        if (st == FALM_OK) st = falm_poisson_matrix(&C->dom, FALM_F64, kind, g64, j64, C->a, &C->maxdiag, s);
        FALM_HIP(hipStreamSynchronize(s));
        (void)hipFree(x64);
        (void)hipFree(k64);
        (void)hipFree(g64);
        (void)hipFree(j64);
        FALM_TRY(st);
    }
    // initial field u = (u_inflow, 0, 0) in every cell, then UtoUU and SGS
    // (app/windtunnel/main.cpp:120-129)
    if (C->prec == FALM_F32)
        hipLaunchKernelGGL((k_fill3<float>), ew_grid(n), dim3(EW_BLOCK), 0, s, (float *)C->u, n, (float)c.u_inflow,
                           0.f, 0.f);
    else
        hipLaunchKernelGGL((k_fill3<double>), ew_grid(n), dim3(EW_BLOCK), 0, s, (double *)C->u, n, c.u_inflow, 0.0,
                           0.0);
    FALM_CHECK_LAUNCH("k_fill3");
    FALM_HIP(hipMemcpyAsync(C->uprev, C->u, 3 * n * tb, hipMemcpyDeviceToDevice, s));
    FALM_TRY(sep_u_to_uu(&C->dom, C->prec, C->u, C->uu, C->sm, s));
    if (c.sgs_model == FALM_SGS_SMAGORINSKY) FALM_TRY(sep_sgs(&C->dom, C->prec, c.Cs, C->u, C->nut, C->sm, s));
    else if (c.sgs_model != FALM_SGS_EMPTY) {
        set_error("falm_case_create: only the Smagorinsky model is implemented");
        return FALM_ERR_UNSUPPORTED;
    }
    FALM_TRY(falm_solver_create(&C->dom, &c.solver, &C->solver));
    FALM_TRY(falm_solver_set_matrix(C->solver, C->a, s));
    if (c.alm) FALM_TRY(falm_alm_create(&C->dom, c.alm, C->x.data(), C->y.data(), C->z.data(), &C->alm));
    case_faces(C);
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

// The step after the Poisson solve: pressure boundary fill, projection,
// velocity boundary fill, Smagorinsky and the divergence monitor.  With a
// gate, every kernel does nothing unless *gate is nonzero.
static falm_status post_solve(falm_case *C, const int *gate) {
    const falm_case_config &c = C->cfg;
    hipStream_t s = C->stream;
    const falm_domain *d = &C->dom;
    const int pr = C->prec;
    const bool is2d = c.kind == FALM_CASE_LID2D;
    FALM_TRY(launch_bc_pressure(d, FALM_F64, C->pbc, C->p, s, gate));
    if (is2d) FALM_TRY(launch_copy_z5(d, FALM_F64, C->p, 1, s, gate));
    FALM_TRY(sep_project(d, pr, c.dt, C->u, C->uu, C->p, C->sm, s, gate));
    falm_inlet_noise nz{c.inlet_intensity, c.seed, C->it};
    FALM_TRY(launch_bc_velocity(d, pr, C->ubc, C->u, C->uprev, C->xyz, c.dt, &nz, s, gate));
    if (is2d) FALM_TRY(launch_copy_z5(d, pr, C->u, 3, s, gate));
    // SGS and the divergence monitor in one pass (both read final fields)
    if (c.sgs_model == FALM_SGS_SMAGORINSKY) {
        FALM_TRY(sep_sgs_div_norm(d, pr, c.Cs, C->u, C->nut, C->uu, C->div, C->sm, C->dnorm, C->partials, s, gate,
                                  C->hnorm));
        if (is2d) FALM_TRY(launch_copy_z5(d, pr, C->nut, 1, s, gate));
    } else {
        FALM_TRY(sep_div_norm(d, pr, C->uu, C->div, C->sm, C->dnorm, C->partials, s, gate, C->hnorm));
    }
    return FALM_OK;
}

falm_status case_step_power(falm_case *C, falm_solve_info *si, double *power) {
    const falm_case_config &c = C->cfg;
    hipStream_t s = C->stream;
    const falm_domain *d = &C->dom;
    const int pr = C->prec;
    const bool is2d = c.kind == FALM_CASE_LID2D;
    C->it++;
    const double t = c.dt * (double)C->it;
    // u_previous <- u (app/windtunnel/main.cpp:136): swap the buffers.  The
    // new u's ghost cells are stale until the boundary fill rewrites them;
    // the wind tunnel's face-flux pass reads them from u_previous (equal to
    // what the reference's in-place update leaves there), the cavities copy
    // the ghost shell back first.  Edge and corner ghost cells are never
    // written nor read after the initial fill, so both buffers keep them.
    std::swap(C->u, C->uprev);
    if (c.kind != FALM_CASE_WINDTUNNEL) FALM_TRY(copy_shell(C, C->uprev, C->u, s));
    if (C->alm) {
        FALM_TRY(alm_step(C->alm, pr, C->uprev, C->ff, t, s, true));
        if (power) FALM_TRY(alm_accumulate_power(C->alm, power, s));
    }
    FALM_TRY(sep_pseudo_u(d, pr, c.adv_scheme, 1.0 / c.Re, c.dt, C->uprev, C->uprev, C->uu, C->u, C->nut, C->sm,
                          C->ff, C->alm ? alm_force_box(C->alm) : C->nobox, s));
    if (c.kind == FALM_CASE_WINDTUNNEL) {
        FALM_TRY(sep_u_to_uu_rhs(d, pr, C->u, C->uprev, C->uu, C->rhs, C->sm, 1.0 / (c.dt * C->maxdiag), s));
    } else {  // the cavities zero the wall fluxes before the right-hand side
        FALM_TRY(sep_u_to_uu(d, pr, C->u, C->uu, C->sm, s));
        FALM_TRY(launch_bc_face_flux(d, pr, C->uubc, C->uu, s));
        FALM_TRY(sep_poisson_rhs(d, pr, C->uu, C->rhs, C->sm, 1.0 / (c.dt * C->maxdiag), s));
    }
    // The solve's first batch of iterations (as many as the previous step
    // needed), then the rest of the step gated on the solver's "over" flag,
    // then the one wait of the step: if the first batch did not finish the
    // solve, the gated kernels did nothing and run again after it.  A solver
    // whose last solve took more than its first batch runs whole instead.
    // No residual output: the apps never read poi_res after the solve
    // (app/windtunnel/main.cpp:147), and the solver iterates on its own.
    double *const res = nullptr;
    const int *gate = solver_gate(C->solver);
    bool rerun = true;
    if (gate) {
        FALM_TRY(solver_solve_mode(C->solver, C->a, C->p, C->rhs, res, s, si, SOLVE_BEGIN, nullptr));
        FALM_TRY(post_solve(C, gate));
        FALM_TRY(solver_solve_mode(C->solver, C->a, C->p, C->rhs, res, s, si, SOLVE_RESUME, &rerun));
    } else {
        FALM_TRY(solver_solve_mode(C->solver, C->a, C->p, C->rhs, res, s, si, SOLVE_FULL, nullptr));
    }
    C->norm_pending = rerun;
    if (rerun) FALM_TRY(post_solve(C, nullptr));
    return FALM_OK;
}

static void *case_field(falm_case *C, int field, int &ncol, int &prec) {
    ncol = 1;
    prec = C->prec;
    switch (field) {
    case FALM_FIELD_U: ncol = 3; return C->u;
    case FALM_FIELD_UU: ncol = 3; return C->uu;
    case FALM_FIELD_P: prec = FALM_F64; return C->p;
    case FALM_FIELD_NUT: return C->nut;
    case FALM_FIELD_FF: ncol = 3; return C->ff;
    case FALM_FIELD_DIV: return C->div;
    case FALM_FIELD_RHS: prec = FALM_F64; return C->rhs;
    case FALM_FIELD_POI_A: ncol = 7; prec = FALM_F64; return C->a;
    case FALM_FIELD_XYZ: ncol = 3; return C->xyz;
    case FALM_FIELD_KX: ncol = 3; return C->kx;
    case FALM_FIELD_G: ncol = 3; return C->g;
    case FALM_FIELD_JA: return C->ja;
    default: return nullptr;
    }
}

}  // namespace falm

extern "C" {

falm_status falm_case_destroy(falm_case *C) {
    if (!C) return FALM_OK;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    if (C->solver) falm_solver_destroy(C->solver);
    if (C->alm) falm_alm_destroy(C->alm);
    for (void *p : C->bufs) (void)hipFree(p);
    if (C->hnorm) (void)hipHostFree(C->hnorm);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
    return FALM_OK;
}

falm_status falm_case_create(const falm_case_config *cfg, falm_case **out) {
    if (!cfg || !out) return FALM_ERR_INVALID;
    *out = nullptr;
    if (cfg->kind < FALM_CASE_WINDTUNNEL || cfg->kind > FALM_CASE_LID2D) return FALM_ERR_INVALID;
    if (cfg->precision != FALM_CASE_F64 && cfg->precision != FALM_CASE_MIXED) return FALM_ERR_UNSUPPORTED;
    {   // the pressure, Poisson right-hand side and the solver's vectors are
        // double in both precisions: their columns must stay below 4 GiB
        size_t cells = 1;
        for (int m = 0; m < 3; m++) cells *= (size_t)(cfg->n[m] > 0 ? cfg->n[m] + 4 : 1);
        FALM_TRY(check_byte_offsets(cells, sizeof(double)));
    }
    falm_case *C = new falm_case();
    C->cfg = *cfg;
    C->cfg.alm = cfg->alm;  // read during create only
    hipError_t e = hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        falm_status st = hip_status(e, "hipStreamCreate");
        delete C;
        return st;
    }
    falm_status st = case_init(C);
    C->cfg.alm = nullptr;
    for (int q = 0; q < 6; q++) C->cfg.cv[q] = nullptr;
    if (st != FALM_OK) {
        falm_case_destroy(C);
        return st;
    }
    *out = C;
    return FALM_OK;
}

falm_status falm_case_step(falm_case *C, int nsteps, falm_step_info *info) {
    if (!C || nsteps < 0) return FALM_ERR_INVALID;
    falm_solve_info si{0, 0.0, 0, FALM_OK};
    for (int k = 0; k < nsteps; k++) FALM_TRY(case_step_power(C, &si, nullptr));
    if (nsteps == 0) FALM_HIP(hipMemcpyAsync(C->hnorm, C->dnorm, sizeof(double), hipMemcpyDeviceToHost, C->stream));
    if (nsteps == 0 || C->norm_pending) FALM_HIP(hipStreamSynchronize(C->stream));
    if (info) {
        const falm_domain &d = C->dom;
        const double nint = (double)(d.shape[0] - 2 * d.gc) * (double)(d.shape[1] - 2 * d.gc) *
                            (double)(d.shape[2] - 2 * d.gc);
        info->step = C->it;
        info->t = C->cfg.dt * (double)C->it;
        info->div = sqrt(C->hnorm[0]) / nint;  // app/windtunnel/main.cpp:185
        info->solver_it = si.it;
        info->solver_err = si.err;
        info->outer_it = si.outer_it;
        info->status = si.status;
        if (!std::isfinite(info->div)) info->status = FALM_ERR_NONFINITE;
    }
    return FALM_OK;
}

falm_status falm_case_get_field(falm_case *C, int field, double *host_out, size_t n) {
    if (!C || !host_out) return FALM_ERR_INVALID;
    int ncol, prec;
    void *src = case_field(C, field, ncol, prec);
    if (!src || n != (size_t)ncol * C->n) {
        set_error("falm_case_get_field: bad field or size");
        return FALM_ERR_INVALID;
    }
    FALM_HIP(hipStreamSynchronize(C->stream));
    if (prec == FALM_F64) {
        FALM_HIP(hipMemcpy(host_out, src, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        FALM_HIP(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++) host_out[q] = (double)tmp[q];
    }
    return FALM_OK;
}

falm_status falm_case_set_field(falm_case *C, int field, const double *host_in, size_t n) {
    if (!C || !host_in) return FALM_ERR_INVALID;
    int ncol, prec;
    void *dst = case_field(C, field, ncol, prec);
    if (!dst || n != (size_t)ncol * C->n) {
        set_error("falm_case_set_field: bad field or size");
        return FALM_ERR_INVALID;
    }
    FALM_HIP(hipStreamSynchronize(C->stream));
    if (prec == FALM_F64) {
        FALM_HIP(hipMemcpy(dst, host_in, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(n);
        for (size_t q = 0; q < n; q++) tmp[q] = (float)host_in[q];
        FALM_HIP(hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (field == FALM_FIELD_POI_A) FALM_TRY(falm_solver_set_matrix(C->solver, C->a, C->stream));
    if (field == FALM_FIELD_FF && C->alm) alm_ff_reset(C->alm);
    FALM_HIP(hipStreamSynchronize(C->stream));
    return FALM_OK;
}

falm_status falm_case_set_tolerance(falm_case *C, double tol, int maxit) {
    if (!C || !(tol > 0.0) || maxit < 1) {
        set_error("falm_case_set_tolerance: bad tolerance or iteration limit");
        return FALM_ERR_INVALID;
    }
    FALM_HIP(hipStreamSynchronize(C->stream));
    FALM_TRY(solver_set_tol(C->solver, tol));
    FALM_TRY(solver_set_maxit(C->solver, maxit));
    C->cfg.solver.tol = tol;
    C->cfg.solver.maxit = maxit;
    return FALM_OK;
}

falm_status falm_case_domain(falm_case *C, falm_domain *dom) {
    if (!C || !dom) return FALM_ERR_INVALID;
    *dom = C->dom;
    return FALM_OK;
}

falm_status falm_case_alm(falm_case *C, falm_alm **alm) {
    if (!C || !alm) return FALM_ERR_INVALID;
    *alm = C->alm;
    return FALM_OK;
}

}  // extern "C"
