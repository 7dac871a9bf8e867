/*
 * falm_amd.h — C ABI of the MI355X (gfx950) fractional-step LES step.
 *
 * Every entry point here replaces one operation of the reference's C++ class
 * API for the single-GPU time step (ki2098/falm).  The citation above each
 * declaration names the reference interface it stands in for.  Conventions:
 *
 *   - Fields follow the reference Matrix layout (src/matrix.h:20): element
 *     (cell, col) of an N-column field lives at ptr[cell + col*ncell], with
 *     cell = i + j*shape[0] + k*shape[0]*shape[1] (src/util.h:29-31) over the
 *     padded local shape (ghost/guide cells included) and
 *     ncell = shape[0]*shape[1]*shape[2].
 *   - Field pointers are device pointers.  Their element type is given by the
 *     `prec` argument (FALM_F64 = double, FALM_F32 = float).  Scalars are double.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Kernels are stream-ordered; entry points that return host scalars
 *     (dot, norms, reductions, solve) synchronise that stream before returning,
 *     which is the reference's "op returns => result visible" contract
 *     (falmWaitStream() at the end of every op, e.g. src/FalmCFD.cpp:66).
 *   - Every entry point returns a falm_status; nothing throws across the ABI.
 *     falm_last_error() gives a readable message for the last failure on the
 *     calling thread.
 */
#ifndef FALM_AMD_H
#define FALM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
/* Codes 0-5 keep the values of the reference's FalmErr (src/error.h:9-16);
 * HIP errors are reported as FALM_ERR_HIP + hipError_t, like the reference's
 * cuErrMask*100 convention (src/dev/cuhandler.h:39). */
typedef int falm_status;
enum {
    FALM_OK                = 0,
    FALM_ERR_MALLOC        = 1,
    FALM_ERR_SETUP         = 2,
    FALM_ERR_INVALID       = 10, /* bad argument, shape or pointer          */
    FALM_ERR_ALIAS         = 11, /* output aliases an input it must not      */
    FALM_ERR_UNSUPPORTED   = 12, /* option outside the implemented scope     */
    FALM_ERR_NONFINITE     = 13, /* NaN/Inf detected in a monitored result   */
    FALM_ERR_NOT_CONVERGED = 14, /* solver hit maxit with err > tol          */
    FALM_ERR_BREAKDOWN     = 15, /* BiCGStab |rho| < FLT_MIN                 */
    FALM_ERR_NO_DEVICE     = 16,
    FALM_ERR_HIP           = 100
};

typedef enum { FALM_F64 = 0, FALM_F32 = 1 } falm_prec;

/* Local sub-domain descriptor: the reference's Region pdm (src/region.h:8-31)
 * plus CPM::global and CPM::gc (src/CPMBase.h:25-32).  On one GPU
 * offset = {0,0,0} and gshape = shape. */
typedef struct falm_domain {
    int shape[3];   /* padded local shape, pdm.shape                   */
    int offset[3];  /* pdm.offset                                      */
    int gshape[3];  /* padded global shape, cpm.global.shape           */
    int gc;         /* guide cells, Falm::GuideCell = 2 (src/typedef.h:41) */
} falm_domain;

/* Reference flag values (src/FalmCFDDevCall.h:10-24, src/FalmEqDevCall.h:35-41,
 * src/flag.h:25-29). */
enum { FALM_ADV_UPWIND1 = 0, FALM_ADV_UPWIND3 = 1, FALM_ADV_QUICK = 2,
       FALM_ADV_UTOPIA = 3, FALM_ADV_KK = 4 };
enum { FALM_SGS_EMPTY = 0, FALM_SGS_SMAGORINSKY = 1, FALM_SGS_CSM = 2 };
enum { FALM_LS_EMPTY = 0, FALM_LS_JACOBI = 1, FALM_LS_SOR = 2, FALM_LS_PBICGSTAB = 4 };
enum { FALM_COLOR_BLACK = 0, FALM_COLOR_RED = 1 };

/* ------------------------------------------------------------- runtime */
/* Select the device used by the calling thread (src/falm.h:88-91). */
falm_status falm_init(int device);
const char *falm_last_error(void);
const char *falm_status_string(falm_status s);
/* falmMallocDevice / falmFreeDevice / falmMemcpy / falmMemsetDevice /
 * falmWaitStream (src/dev/cuhandler.h:12-118).  kind: 1 H2H, 2 H2D, 4 D2H,
 * 8 D2D (MCP flags, src/flag.h:8-15). */
falm_status falm_malloc(void **ptr, size_t bytes);
falm_status falm_free(void *ptr);
falm_status falm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
falm_status falm_memset(void *ptr, int value, size_t bytes, void *stream);
falm_status falm_stream_sync(void *stream);

/* ------------------------------------------------------------- FalmCFD */
/* FalmCFD::FSPseudoU (src/FalmCFD.h:32-47, src/FalmCFD.cpp:5-67; kernel
 * src/dev/FalmCFDDevCall.cu:10-301, schemes src/FalmCFDScheme.h:8-191).
 * ua = un + dt*(-adv(u) + vis(u) + ff) on interior cells.  ua must not alias
 * u or un (the reference apps pass ua == u, a data race; this entry point
 * returns FALM_ERR_ALIAS instead).  ReI = 1/Re. */
falm_status falm_pseudo_u(const falm_domain *dom, int prec, int adv_scheme,
                          double ReI, double dt,
                          const void *un, const void *u, const void *uu, void *ua,
                          const void *nut, const void *kx, const void *g,
                          const void *ja, const void *ff, void *stream);

/* FalmCFD::UtoUU (src/FalmCFD.cpp:69-129; kernels UtoCU + InterpolateCU,
 * src/dev/FalmCFDDevCall.cu:303-349), fused: no cell-centre flux buffer. */
falm_status falm_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu,
                         const void *kx, const void *ja, void *stream);

/* FalmCFD::MACCalcPoissonRHS (src/FalmCFD.h:96-109): divergence on interior
 * cells then ScaleMatrix by 1/(dt*maxdiag) over the whole array (ghosts too). */
falm_status falm_poisson_rhs(const falm_domain *dom, int prec, const void *uu, void *rhs,
                             const void *ja, double dt, double maxdiag, void *stream);

/* FalmCFD::ProjectP (src/FalmCFD.cpp:131-185; kernels
 * src/dev/FalmCFDDevCall.cu:351-407).  u may alias ua and uu may alias uua
 * (point-wise update), as the reference apps call it. */
falm_status falm_project_p(const falm_domain *dom, int prec, double dt,
                           void *u, const void *ua, void *uu, const void *uua,
                           const void *p, const void *kx, const void *g, void *stream);

/* FalmCFD::SGS (src/FalmCFD.cpp:187-238; Smagorinsky kernel
 * src/dev/FalmCFDDevCall.cu:409-476).  model FALM_SGS_EMPTY does nothing
 * (as the reference dispatch, :774-797); FALM_SGS_CSM is unsupported. */
falm_status falm_sgs(const falm_domain *dom, int prec, int model, double Cs,
                     const void *u, void *nut, const void *kx, const void *ja, void *stream);

/* FalmCFD::Divergence (src/FalmCFD.h:84-94; kernel :557-581). */
falm_status falm_divergence(const falm_domain *dom, int prec, const void *uu, void *div,
                            const void *ja, void *stream);

/* ------------------------------------------------------------- FalmMV */
/* FalmMV::MV (src/MV.h:13-45; kernel src/dev/MVDevCall.cu:10-41).
 * a has 7 columns in the order [c,w,e,s,n,b,t]. */
falm_status falm_mv(const falm_domain *dom, int prec, const void *a, const void *x,
                    void *ax, void *stream);
/* FalmMV::DotProduct / EuclideanNormSq (src/MV.h:47-65; kernels
 * src/dev/MVDevCall.cu:57-179).  Interior cells; result accumulated in
 * double; ncol columns summed for the norm. */
falm_status falm_dot(const falm_domain *dom, int prec, const void *a, const void *b,
                     double *out, void *stream);
falm_status falm_norm2(const falm_domain *dom, int prec, const void *a, int ncol,
                       double *out, void *stream);
/* falm_dot without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * the same result bits as falm_dot, but `out` is DEVICE memory (one double)
 * and the call only enqueues work on `stream` (no synchronisation).  The
 * per-workgroup partial sums live in a fixed 32 KiB device array (4096
 * doubles) per GPU, enough for every domain falm_dot accepts; the array is
 * shared by all calls on one GPU, so calls that overlap on one GPU race. */
falm_status falm_dot_noalloc(const falm_domain *dom, int prec, const void *a, const void *b,
                             double *out, void *stream);
/* FalmMV::MatColMax/MatColMin/MatColAbsMax/MatColAbsMin/VecMax/VecMin
 * (src/MV.h:67-136; kernels src/dev/MVDevCall.cu:216-598).  op selects one of
 * FALM_RED_*.  MaxDiag == MatColAbsMax(a, 0). */
enum { FALM_RED_COL_MAX = 0, FALM_RED_COL_MIN = 1, FALM_RED_COL_ABSMAX = 2,
       FALM_RED_COL_ABSMIN = 3, FALM_RED_VEC_MAX = 4, FALM_RED_VEC_MIN = 5 };
falm_status falm_reduce(const falm_domain *dom, int prec, int op, const void *a,
                        int ncol, int col, double *out, void *stream);
/* FalmMV::ScaleMatrix / MatrixAdd / Vecaxby (src/MVDevCall.h:61-65;
 * kernels src/dev/MVDevCall.cu:506-522,660-695), over n elements. */
falm_status falm_scale(int prec, void *a, size_t n, double scale, void *stream);
falm_status falm_matrix_add(int prec, void *a, const void *b, size_t n, void *stream);
falm_status falm_vecaxby(int prec, double a, const void *x, double b, const void *y,
                         void *result, size_t n, void *stream);

/* ------------------------------------------------------------- FalmEq */
/* FalmEq::Res (src/FalmEq.cpp:45-78; kernel src/dev/FalmEqDevCall.cu:54-85). */
falm_status falm_residual(const falm_domain *dom, int prec, const void *a, const void *x,
                          const void *b, void *r, void *stream);
/* One red-black SOR half sweep (FalmEqDevCall::SORSweep,
 * src/dev/FalmEqDevCall.cu:175-221): cells with
 * (i+j+k+sum(offset)) % 2 == color get x += omega*(b - A x)_c / a_c. */
falm_status falm_sor_sweep(const falm_domain *dom, int prec, const void *a, void *x,
                           const void *b, double omega, int color, void *stream);
/* One Jacobi sweep x = xp + (b - A xp)/a_c (src/dev/FalmEqDevCall.cu:101-142). */
falm_status falm_jacobi_sweep(const falm_domain *dom, int prec, const void *a, void *x,
                              const void *xp, const void *b, void *stream);

/* Linear solver object: FalmEq (src/FalmEq.h:14-104).  The library owns all
 * work vectors (FalmEq::alloc, src/FalmEq.h:27-40) and reduction buffers; the
 * solve loop allocates nothing.  precision FALM_SOLVE_MIXED runs the
 * preconditioned Krylov iterations on float copies inside a double
 * defect-correction loop (r = b - A x, dots, norms and x kept in double). */
enum { FALM_SOLVE_F64 = 0, FALM_SOLVE_F32 = 1, FALM_SOLVE_MIXED = 2 };
typedef struct falm_solver_config {
    int    type;         /* FALM_LS_*                                   */
    int    maxit;
    double tol;          /* on sqrt(|r|^2)/N_interior (src/FalmEq.cpp:330) */
    double relax;        /* relax_factor (SOR solver)                   */
    int    pc_type;      /* FALM_LS_EMPTY / FALM_LS_JACOBI / FALM_LS_SOR */
    int    pc_maxit;
    double pc_relax;
    int    precision;    /* FALM_SOLVE_*                                */
    int    inner_maxit;  /* mixed: Krylov iterations per correction (0 -> 50) */
    double inner_rtol;   /* mixed: inner residual reduction (0 -> 1e-4)  */
    int    check_every;  /* iterations between host polls (0: the previous
                            solve's count first, then max(1, 4e6/N) up to 32) */
} falm_solver_config;

typedef struct falm_solve_info {
    int    it;           /* FalmEq::it  (Krylov iterations, all corrections) */
    double err;          /* FalmEq::err                                   */
    int    outer_it;     /* mixed: defect-correction passes (else 1)      */
    int    status;       /* FALM_OK / FALM_ERR_BREAKDOWN / NOT_CONVERGED   */
} falm_solve_info;

typedef struct falm_solver falm_solver;
falm_status falm_solver_create(const falm_domain *dom, const falm_solver_config *cfg,
                               falm_solver **out);
/* FalmEq::Solve (src/FalmEq.h:56-64).  a, x, b, r are double fields; x is the
 * warm-started solution (pressure), r receives the final residual (the
 * recursive one of BiCGStab in double, b - A x in mixed precision).  r may be
 * NULL when the caller does not need it. */
falm_status falm_solver_solve(falm_solver *s, const void *a, void *x, const void *b,
                              void *r, void *stream, falm_solve_info *info);
/* Declare the matrix once: the solver derives its float copy (mixed
 * precision) and, when every off-diagonal column depends on one index only,
 * the one-index coefficient arrays and the diagonal's per-axis classes (each
 * checked bit for bit against a).  Solves with the same pointer use them, so
 * call this again after changing the matrix's values.  A solve with another
 * pointer reads that matrix's seven columns as given (in mixed precision it
 * first declares that matrix itself). */
falm_status falm_solver_set_matrix(falm_solver *s, const void *a, void *stream);
falm_status falm_solver_destroy(falm_solver *s);

/* ---------------------------------------------- ghost-cell boundary fill */
/* Per-face velocity boundary kinds, one per reference BC kernel:
 *   INLET      o1=o2=value (+ optional inlet noise)  app/windtunnel/bcdevcall.cu:8-22
 *   CONVECTIVE convective outflow from uprev          app/windtunnel/bcdevcall.cu:25-43
 *   SLIP       tangential copy, zero normal           app/windtunnel/bcdevcall.cu:46-79,100-115
 *   WALL0      o1 = 0, o2 = -u_i1                     app/windtunnel/bcdevcall.cu:82-97
 *   REFLECT    o1 = 2v-u_i1, o2 = 2v-u_i2             app/lid3d/bcdev.cu:11-105
 * Face order is CPM's: x-, x+, y-, y+, z-, z+ (src/CPMBase.h:14-19). */
enum { FALM_BC_NONE = 0, FALM_BC_INLET = 1, FALM_BC_CONVECTIVE = 2, FALM_BC_SLIP = 3,
       FALM_BC_WALL0 = 4, FALM_BC_REFLECT = 5 };
typedef struct falm_face_bc {
    int    type;
    double value[3];
} falm_face_bc;
/* Seeded Gaussian perturbation of INLET ghost values (no reference
 * counterpart): value[d] + intensity*N(0,1), with the normal deviate drawn
 * from Philox4x32-10 keyed by (seed, step, j, k, d). */
typedef struct falm_inlet_noise {
    double   intensity;
    uint64_t seed;
    int64_t  step;
} falm_inlet_noise;
/* ubc (app/windtunnel/bc.h:6-33, app/lid3d/bc.h, app/ld2d/bc.h).  uprev and
 * xyz are needed only by CONVECTIVE faces; noise may be NULL. */
falm_status falm_bc_velocity(const falm_domain *dom, int prec, const falm_face_bc faces[6],
                             void *u, const void *uprev, const void *xyz, double dt,
                             const falm_inlet_noise *noise, void *stream);
/* pbc (app/windtunnel/bcdevcall.cu:153-211): per face 0 = untouched,
 * 1 = Neumann (copy first interior cell), 2 = Dirichlet p = 0. */
enum { FALM_PBC_NONE = 0, FALM_PBC_NEUMANN = 1, FALM_PBC_ZERO = 2 };
falm_status falm_bc_pressure(const falm_domain *dom, int prec, const int faces[6], void *p,
                             void *stream);
/* uubc (app/lid3d/bcdev.cu:251-321): zero the wall-normal face flux on the
 * faces whose mask entry is non-zero. */
falm_status falm_bc_face_flux(const falm_domain *dom, int prec, const int faces[6], void *uu,
                              void *stream);
/* copy_z5 (app/ld2d/bcdevcall.cu:240-256): copy the k = gc slice of every
 * column into k = gc-2 .. gc+2 (2-D cases run as z-thin 3-D domains). */
falm_status falm_copy_z5(const falm_domain *dom, int prec, void *v, int ncol, void *stream);

/* ------------------------------------------------------- metrics / matrix */
/* FalmCore::computation_init metric loop (src/falm.h:383-404): from 1-D cell
 * centres x,y,z and widths hx,hy,hz (host arrays over the padded local
 * shape) fill xyz (3), kx (3), g (3), ja (1). */
falm_status falm_metrics(const falm_domain *dom, int prec,
                         const double *x, const double *y, const double *z,
                         const double *hx, const double *hy, const double *hz,
                         void *xyz, void *kx, void *g, void *ja, void *stream);
/* Poisson coefficient assembly + MaxDiag + ScaleMatrix
 * (app/windtunnel/main.cpp:20-92, app/lid3d/main.cpp:22-92,
 * app/ld2d/main.cpp:14-68, src/MV.h:67-76).  kind selects the app's rules. */
enum { FALM_POISSON_NEUMANN = 0,        /* lid3d: all walls Neumann          */
       FALM_POISSON_OUTLET_XPLUS = 1,   /* windtunnel: x+ coupling kept (p=0)*/
       FALM_POISSON_2D = 2 };           /* ld2d: 5-point, no z couplings     */
falm_status falm_poisson_matrix(const falm_domain *dom, int prec, int kind, const void *g,
                                const void *ja, void *a, double *maxdiag, void *stream);

/* ------------------------------------------------------ actuator line (ALM) */
/* One turbine (src/alm/turbineHandler.h:13-205).  angle_type: 0 none,
 * 1 roll, 2 pitch, 3 yaw (src/falmath.h:33); motion = {amplitude [rad],
 * angular frequency, phase} of that one Euler angle
 * (src/alm/almDevCall.cu:43-61).  surge, pitch_deg, yaw_deg and
 * angle_offset_deg have no reference counterpart: surge = {amplitude,
 * angular frequency, phase} of a base translation along x; pitch_deg adds a
 * per-blade pitch to the AP twist; yaw_deg turns the rotor about the vertical
 * axis through the base inside the platform motion (rotor frame = platform
 * frame turned by yaw_deg about z); angle_offset_deg is a static part of the
 * platform's Euler angle (angle = offset + A sin(w t + phase); used only when
 * angle_type is not 0). */
typedef struct falm_turbine {
    double base[3];
    double hub[3];        /* (-overhang, 0, tower)                     */
    double tip_rate;      /* rotor angular velocity (rad / time unit)  */
    int    angle_type;
    double motion[3];
    double surge[3];
    double pitch_deg[8];
    double yaw_deg;
    double angle_offset_deg;
} falm_turbine;
enum { FALM_ALM_FIX_DRAG_TANGENT = 1 };  /* use cos(phi) for the drag share of
                                            the tangential force instead of the
                                            reference's sin(phi)
                                            (src/alm/almDevCall.cu:239)       */
/* A culling radius at which exp(-r^2/eps^2) < 2^-53: every dropped term is
 * below double rounding of the point's peak contribution (exp(-6.1^2) =
 * 6.9e-17).  cutoff = 0 keeps every term that is non-zero in double, which
 * equals the reference's all-pairs sum (src/alm/almDevCall.cu:255-294). */
#define FALM_ALM_CUTOFF_ROUNDING 6.1
typedef struct falm_alm_config {
    int    n_turbine, n_blade, n_ap_per_blade, n_attack;
    double radius, hub_radius;
    double eps;                   /* Gaussian projection width (projectionWidth) */
    const falm_turbine *turbines; /* [n_turbine]                               */
    const double *ap_r;           /* [n_ap]  (ap.json "r")                     */
    const double *ap_chord;       /* [n_ap]                                    */
    const double *ap_twist;       /* [n_ap]  degrees                           */
    const double *attack;         /* [n_attack] degrees, ascending             */
    const double *cl;             /* [n_attack*n_ap], cl[ap + a*n_ap]          */
    const double *cd;
    const double *twist_offset;   /* [n_ap_per_blade] degrees, or NULL         */
    double cutoff;                /* cell culling radius / eps; 0 = exact      */
    int    flags;
} falm_alm_config;
typedef struct falm_alm falm_alm;
/* AlmHandler::init (src/alm/alm.h:20-57).  x, y, z are the host 1-D cell
 * centre arrays of the padded local shape (FalmBaseMesh, src/FalmMeshBase.h). */
falm_status falm_alm_create(const falm_domain *dom, const falm_alm_config *cfg,
                            const double *x, const double *y, const double *z,
                            falm_alm **out);
/* AlmHandler::Alm (src/alm/alm.h:66-77): turbine angles, AP positions,
 * AP forces from velocity u (3 cols), Gaussian projection into ff (3 cols,
 * interior cells), torque and thrust per turbine. */
falm_status falm_alm_step(falm_alm *alm, int prec, const void *u, void *ff, double t,
                          void *stream);
/* Host copies (synchronising): AP positions [n_ap*3], AP forces [n_ap*3],
 * per-turbine torque and thrust [n_turbine].  Any pointer may be NULL. */
falm_status falm_alm_get(falm_alm *alm, double *ap_xyz, double *ap_force,
                         double *torque, double *thrust, void *stream);
falm_status falm_alm_set_twist_offset(falm_alm *alm, const double *twist_offset_deg);
falm_status falm_alm_destroy(falm_alm *alm);
/* BladeHandler::buildAP (src/alm/bladeHandler.h:16-104) without the JSON
 * round trip: interpolate a blade table (n_station stations at radius rr,
 * chord, twist, cl/cd[station*n_attack + a]) to n_ap_per_blade points per
 * blade; outputs sized [n_ap] and [n_attack*n_ap].  Host-only. */
falm_status falm_build_ap(int n_station, int n_attack, const double *rr,
                          const double *chord, const double *twist,
                          const double *cl_table, const double *cd_table,
                          int n_turbine, int n_blade, int n_ap_per_blade,
                          double radius, double hub_radius,
                          double *ap_r, double *ap_chord, double *ap_twist,
                          double *ap_cl, double *ap_cd);

/* ------------------------------------------------------- input formats */
/* Mesher::build_mesh (src/mesher/mesher.hpp:12-260): convert the node list
 * in meshpath (mesh.json with "coordinatePoint": "node", x.txt, y.txt,
 * z.txt) into the control-volume file outputfile ("imax jmax kmax gc", then
 * "\t%.15e\t%.15e" centre/width lines for x, y, z; src/FalmIO.h:118-133).
 * cvcenter "innerNode": nodes are cell centres, the outermost ones the
 * first ghost layer; "cell": nodes are cell faces.  Host-only. */
falm_status falm_mesher_build(const char *cvcenter, const char *meshpath, const char *outputfile, int gc);

/* ------------------------------------------------------------ step driver */
/* The per-step orchestration of the reference apps (app/nedo/main.cpp:150-204,
 * app/windtunnel/main.cpp:135-165, app/lid3d/main.cpp:122-175,
 * app/ld2d/main.cpp:112-136) with device-resident state. */
enum { FALM_CASE_WINDTUNNEL = 0, FALM_CASE_LID3D = 1, FALM_CASE_LID2D = 2 };
enum { FALM_CASE_F64 = 0, FALM_CASE_MIXED = 1 };
typedef struct falm_case_config {
    int    kind;
    int    n[3];             /* interior cells                           */
    double h[3];             /* uniform cell widths                      */
    double origin[3];        /* low corner of the interior               */
    double Re, dt, Cs;
    int    adv_scheme, sgs_model;
    double u_inflow;         /* initial field / inlet velocity           */
    double lid_velocity;     /* cavity lid speed (y+ face, +x direction) */
    double inlet_intensity;  /* windtunnel inlet noise, 0 = steady       */
    uint64_t seed;
    int    zminus_slip;      /* windtunnel z- face: 0 wall (reference), 1 slip */
    int    precision;        /* FALM_CASE_F64 / FALM_CASE_MIXED          */
    falm_solver_config solver;
    const falm_alm_config *alm;  /* NULL = no rotor                      */
    /* Optional non-uniform control volumes, the centres and widths of the
     * padded cells of each axis as a control-volume file holds them
     * (src/FalmIO.h:99-116): cv = {x, hx, y, hy, z, hz}, n[m]+4 values
     * each.  NULL = uniform cells of width h[m] from origin[m]. */
    const double *cv[6];
} falm_case_config;
typedef struct falm_step_info {
    int64_t step;            /* FalmCore::it after the step              */
    double  t;               /* it*dt (src/falm.h:81)                    */
    double  div;             /* sqrt(sum div^2)/N_interior (app/windtunnel/main.cpp:185) */
    int     solver_it;
    double  solver_err;
    int     outer_it;
    int     status;
} falm_step_info;
enum { FALM_FIELD_U = 0, FALM_FIELD_UU = 1, FALM_FIELD_P = 2, FALM_FIELD_NUT = 3,
       FALM_FIELD_FF = 4, FALM_FIELD_DIV = 5, FALM_FIELD_RHS = 6, FALM_FIELD_POI_A = 7,
       FALM_FIELD_XYZ = 8, FALM_FIELD_KX = 9, FALM_FIELD_G = 10, FALM_FIELD_JA = 11 };
typedef struct falm_case falm_case;
falm_status falm_case_create(const falm_case_config *cfg, falm_case **out);
/* Advance nsteps steps; info (may be NULL) receives the last step's record. */
falm_status falm_case_step(falm_case *c, int nsteps, falm_step_info *info);
/* Host copy of a field as double (ncol*ncell values), shape and column count. */
falm_status falm_case_get_field(falm_case *c, int field, double *host_out, size_t n);
falm_status falm_case_set_field(falm_case *c, int field, const double *host_in, size_t n);
falm_status falm_case_domain(falm_case *c, falm_domain *dom);
falm_status falm_case_alm(falm_case *c, falm_alm **alm);
falm_status falm_case_destroy(falm_case *c);
/* The step driver's Poisson tolerance and iteration limit from the next step
 * on (FalmEq's public tol and maxit, which the reference reads on every
 * solve, src/FalmEq.h:56-64). */
falm_status falm_case_set_tolerance(falm_case *c, double tol, int maxit);

/* ------------------------------------------------- rotor-power adjoint
 * No reference counterpart (the reference has no adjoint).  The objective is
 * J = mean over the next nsteps steps of the rotor power
 * P = sum_t Omega_t * torque_t, the quantity AlmHandler::writePowerThrust
 * reports (src/alm/alm.h:126-138).  Both calls leave the case's state as it
 * was (fields and step counter).
 *
 * falm_case_window: J by running the window forward (any case precision). */
falm_status falm_case_window(falm_case *c, int nsteps, double *J);
/* falm_case_adjoint: J and its gradient by the discrete adjoint of the step:
 * g_pitch[n_turbine*8] = dJ/d(pitch_deg[b]) and g_twist[n_ap_per_blade] =
 * dJ/d(twist_offset_deg[s]), both per degree.  The forward window keeps a
 * checkpoint every ckpt steps and recomputes each segment in the reverse
 * sweep; the transposed Poisson solves stop at |r|/N <= adj_rtol*|rhs|/N.
 * bytes (may be NULL) receives the device memory the adjoint allocated.
 * Wind-tunnel cases in FP64 with Upwind3-family advection and Smagorinsky or
 * no SGS model; FALM_ERR_UNSUPPORTED otherwise. */
falm_status falm_case_adjoint(falm_case *c, int nsteps, int ckpt, double adj_rtol, double *J,
                              double *g_pitch, double *g_twist, size_t *bytes);
/* Per-blade pitch of one turbine (8 values, degrees; NULL = 0), the
 * falm_turbine::pitch_deg of falm_alm_create. */
falm_status falm_alm_set_pitch(falm_alm *alm, int turbine, const double *pitch_deg);
/* The rotor's static yaw and the static part of the platform angle of one
 * turbine (degrees), the falm_turbine::yaw_deg / angle_offset_deg of
 * falm_alm_create. */
falm_status falm_alm_set_attitude(falm_alm *alm, int turbine, double yaw_deg, double angle_offset_deg);
/* The attitude gradient of the last falm_case_adjoint call on this case:
 * g_yaw[n_turbine] = dJ/d(yaw_deg) and g_offset[n_turbine] =
 * dJ/d(angle_offset_deg), per degree (the offset's entry is 0 for a turbine
 * with angle_type 0).  Each step's point kinematics are differentiated in
 * forward mode with the velocity held (positions, sampling weights, frame
 * turns, polar segment) and contracted with that step's adjoints of the force
 * field and of the power; the flow's response comes through those adjoints.
 * FALM_ERR_SETUP before any falm_case_adjoint call. */
falm_status falm_case_adjoint_attitude(falm_case *c, double *g_yaw, double *g_offset);

/* ------------------------------------------- per-stage adjoints (VJPs)
 * No reference counterpart.  Vector-Jacobian products of the FP64 forward
 * stages above, the transposes the rotor-power adjoint chains: given the
 * adjoint of a stage's output, each adds the adjoint of its inputs to the
 * *bar outputs (+=, so several contributions accumulate), except where
 * noted.  FP64 device arrays in the Matrix layout; atomics make the sums
 * order-independent to rounding (not bitwise reproducible from run to run).
 * The *bar outputs (and falm_alm_vjp's g_pitch, g_twist) must be device
 * memory from hipMalloc: the kernels accumulate into them with hardware FP64
 * atomic adds, which can drop updates on host-pinned, managed or
 * fine-grained memory.  Host and managed pointers are refused with
 * FALM_ERR_INVALID.
 *
 * falm_pseudo_u: adjoint w.r.t. u in both roles (un = u, as the step drivers
 * call it), uu, nut and ff; Upwind3, UTOPIA and KK only (else
 * FALM_ERR_UNSUPPORTED); ffbar is overwritten on interior cells (= dt uabar). */
falm_status falm_pseudo_u_vjp(const falm_domain *dom, int adv_scheme, double ReI, double dt,
                              const void *u, const void *uu, const void *nut, const void *kx,
                              const void *g, const void *ja, const void *uabar, void *ubar,
                              void *uubar, void *nutbar, void *ffbar, void *stream);
/* falm_u_to_uu: uubar (face region) -> ubar */
falm_status falm_u_to_uu_vjp(const falm_domain *dom, const void *kx, const void *ja,
                             const void *uubar, void *ubar, void *stream);
/* falm_poisson_rhs: rhsbar (interior) -> uubar */
falm_status falm_poisson_rhs_vjp(const falm_domain *dom, const void *ja, double dt, double maxdiag,
                                 const void *rhsbar, void *uubar, void *stream);
/* falm_project_p: (ubar, uubar) of the outputs -> uabar, uuabar, pbar */
falm_status falm_project_p_vjp(const falm_domain *dom, double dt, const void *kx, const void *g,
                               const void *ubar, const void *uubar, void *uabar, void *uuabar,
                               void *pbar, void *stream);
/* falm_sgs (Smagorinsky): nutbar -> ubar; a zero subgradient where |S| = 0 */
falm_status falm_sgs_vjp(const falm_domain *dom, double Cs, const void *u, const void *kx,
                         const void *ja, const void *nutbar, void *ubar, void *stream);
/* falm_bc_velocity (INLET, CONVECTIVE, SLIP, WALL0 faces; no noise): the
 * adjoint unp1bar of the filled u -> usbar (the interior u it was filled
 * from; its interior is overwritten with unp1bar's) and unbar (uprev, read
 * by convective faces). */
falm_status falm_bc_velocity_vjp(const falm_domain *dom, const falm_face_bc faces[6],
                                 const void *xyz, double dt, const void *unp1bar, void *usbar,
                                 void *unbar, void *stream);
/* falm_bc_pressure: in place on pbar (ghost adjoints moved to the interior
 * cells they were copied from, then zeroed) */
falm_status falm_bc_pressure_vjp(const falm_domain *dom, const int faces[6], void *pbar,
                                 void *stream);
/* falm_alm_step (FP64) at time t: the adjoint ffbar of the projected force
 * field and the weight wP of the rotor power sum_t Omega_t torque_t ->
 * ubar (3 columns) and the adjoints of falm_turbine::pitch_deg
 * (g_pitch[n_turbine*8]) and of the twist offsets (g_twist[n_ap_per_blade]),
 * per degree, all device arrays, accumulated.  Refreshes the handle's point
 * positions and forces at t (as a step would), not ff or the force box. */
falm_status falm_alm_vjp(falm_alm *alm, const void *u, double t, const void *ffbar, double wP,
                         void *ubar, void *g_pitch, void *g_twist, void *stream);
/* Base position and base velocity of one turbine from the next step on
 * (velocity NULL = 0); the surge motion of falm_turbine adds to both.
 * Drives base motion read from a file: the reference parses a file name
 * given for a baseLocation coordinate but never reads it
 * (src/alm/turbineHandler.h:144-150).  Synchronises the stream first. */
falm_status falm_alm_set_base(falm_alm *alm, int turbine, const double *base, const double *velocity,
                              void *stream);

/* Device-time accounting of one kernel family (HIP events recorded around
 * each launch of that family on the stream it is launched on).  Used for the
 * roofline figure; off (FALM_TIME_NONE) by default.  bytes is the sum of the
 * algorithmic bytes of the timed launches. */
enum { FALM_TIME_NONE = 0, FALM_TIME_SOR = 1, FALM_TIME_MV = 2, FALM_TIME_PSEUDO_U = 3,
       FALM_TIME_ALM = 4, FALM_TIME_SOLVER_UPDATE = 5, FALM_TIME_PROJECT = 6,
       FALM_TIME_FACE_FLUX = 7, FALM_TIME_SOLVER_INIT = 8, FALM_TIME_SGS = 9,
       FALM_TIME_RHS = 10, FALM_TIME_DIVERGENCE = 11,
       FALM_TIME_SOR_S = 12,   /* reserved (no kernel books it)             */
       FALM_TIME_MV2 = 13,     /* the second operator product per iteration  */
       FALM_TIME_PCMV = 14,    /* r, p, its preconditioning and A in one pass */
       FALM_TIME_PCMV_S = 15,  /* s, its preconditioning and A in one pass    */
       FALM_TIME_PC_APPLY = 16 }; /* end of a pass: x += M(y + alpha p + omega s) */
falm_status falm_timing_enable(int family);
falm_status falm_timing_read(int *launches, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* FALM_AMD_H */
