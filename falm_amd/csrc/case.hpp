// case.hpp — the device-resident step driver's state, shared by the step
// (case.hip) and the adjoint (adjoint.hip).
#pragma once

#include <vector>

#include "cfd.hpp"

struct falm_case {
    falm_case_config cfg;
    falm_domain dom;
    size_t n;
    int prec;                 // FALM_F64 or FALM_F32 (CFD fields)
    size_t tb;                // bytes per CFD element
    hipStream_t stream = nullptr;
    // CFD fields (prec)
    void *u = nullptr, *uprev = nullptr, *uu = nullptr, *nut = nullptr, *ff = nullptr, *div = nullptr;
    void *xyz = nullptr, *kx = nullptr, *g = nullptr, *ja = nullptr;
    // Poisson fields (double)
    double *p = nullptr, *rhs = nullptr, *a = nullptr;
    double maxdiag = 1.0;
    falm_solver *solver = nullptr;
    falm_alm *alm = nullptr;
    std::vector<double> x, y, z, hx, hy, hz;
    falm::SepMesh sm{};       // device 1-D widths and reciprocals
    int64_t it = 0;
    falm_face_bc ubc[6];
    int pbc[6], uubc[6];
    // device scalar for the divergence norm + its reduction scratch
    double *dnorm = nullptr, *partials = nullptr, *hnorm = nullptr;
    bool norm_pending = false;  // the last step's norm (written into hnorm by its kernel) is not waited for
    int *nobox = nullptr;     // empty cell box: no rotor, ff stays zero
    std::vector<void *> bufs;
    // the last falm_case_adjoint's attitude gradient: per turbine
    // (d/d yaw_deg, d/d angle_offset_deg)
    std::vector<double> g_att;
};


namespace falm {
// one time step on the case's stream; power, when given, accumulates the
// step's rotor power on the device.  It waits for the stream once, after the
// step's last kernel when the first batch of solver iterations finished the
// solve (else the step's end, whose last kernel writes the divergence norm
// into the pinned hnorm, is still in flight: norm_pending).
falm_status case_step_power(falm_case *C, falm_solve_info *si, double *power);
falm_status case_dalloc(falm_case *C, void **p, size_t bytes);
// stopping tolerance of a solver's next solves (|r|/N), replacing its config's
falm_status solver_set_tol(falm_solver *S, double tol);
falm_status solver_set_maxit(falm_solver *S, int maxit);
// falm_solver_solve split around work enqueued between a SOLVE_BEGIN call
// (the first batch of iterations, no wait) and a SOLVE_RESUME call (polls and
// finishes the solve; *rerun: the first batch had not finished it, so work
// gated on solver_gate() did nothing)
enum SolveMode { SOLVE_FULL = 0, SOLVE_BEGIN = 1, SOLVE_RESUME = 2 };
falm_status solver_solve_mode(falm_solver *S, const void *a, void *x, const void *b, void *r, hipStream_t s,
                              falm_solve_info *info, int mode, bool *rerun);
const int *solver_gate(const falm_solver *S);
}  // namespace falm
