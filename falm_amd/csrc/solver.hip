// solver.hip — FalmEq on gfx950 (src/FalmEq.h:14-104, src/FalmEq.cpp:45-410).
//
// The whole Krylov loop is device resident: every scalar of BiCGStab
// (rho, alpha, omega, beta, err, it) lives in a DevState in device memory.
// Each reduction is two-phase: the producing kernel writes one partial per
// workgroup, a one-workgroup finalize kernel sums them in a fixed order
// (results do not depend on scheduling) and updates the scalars.  Every
// solver kernel returns at once when `done` is set, so the host only
// enqueues iterations and polls the state every `check_every` iterations
// instead of copying partial sums back after each dot product.
//
// Fusions relative to the reference kernels (any preconditioner):
//   residual + |r|^2 + rr = r + rho                              (init)
//   p = r + beta(p - omega q) + first preconditioner sweep from zero
//   q = A p^ + (rr, q)                                           -> alpha
//   s = r - alpha q + first preconditioner sweep from zero
//   t = A s^ + (t, s), (t, t)                                    -> omega
//   x += alpha p^ + omega s^, r = s - omega t, |r|^2, (r, rr)    -> err, it, beta
// With the one-pass SOR preconditioner (pc_maxit <= 1) an iteration is two
// k-marching kernels (k_bicg_pcmv_km: the vector update, the preconditioner
// sweeps and the operator product in one pass, for p and for s) and one
// finalize (k_fin_fold); the preconditioner is applied to x once per pass
// (k_bicg_close_km, M being linear).
//
// Operator storage: falm_solver_set_matrix checks whether each off-diagonal
// column of A depends on one index only (bitwise); if so the solver keeps the
// diagonal column plus six 1-D arrays of the stored values (8 B/cell of
// matrix traffic instead of 56) — the same coefficients, so the same
// results.  Otherwise the seven columns are read as given.
//
// Mixed precision (FALM_SOLVE_MIXED): the Krylov vectors, a float copy of the
// operator and the preconditioner sweeps are float; dots and norms
// accumulate in double; x (the pressure) is updated in double; each pass
// starts from the true residual b - A x computed in double with the double
// operator.
#include <cfloat>
#include <cstdint>
#include <climits>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "case.hpp"
#include "km.hpp"
#include "linalg.hpp"

namespace falm {

struct DevState {
    double rho, rrho, alpha, omega, beta;
    double err;         // FalmEq::err: recursive residual of the pass; in
                        // mixed mode the init kernels store the true one
    double tol_inner;   // the pass stops when err <= tol_inner
    double gN;          // global interior cell count
    double tol, inner_rtol;
    int it_pass;        // iterations of the current pass
    int maxit_pass;
    int done;           // the pass is over: iteration kernels do nothing
    int final;          // the solve is over: init kernels do nothing
    int status;
    int first;          // next p update is p = r
    int mixed;
    int total;          // iterations over all passes
    int passes;         // passes started
    int maxit, inner_maxit;
    int pending;        // deferred preconditioning: the pass's sum is not yet in x
    int plast;          // folded update: the buffer the last p went to
};

constexpr int NT = 256;

// A mixed-precision float pass stops once its recursive residual is below
// this fraction of the tolerance (or below inner_rtol times the pass's
// starting residual); the true residual in double then decides, and a
// further pass runs only if it is still above the tolerance.  At the full
// tolerance the float pass takes the FP64 solve's iteration count (7 per
// step at 512 x 128 x 128, tol 1e-8); a margin of 0.5 cost 12.3 iterations.
#ifndef FALM_MIXED_TOL_MARGIN
#define FALM_MIXED_TOL_MARGIN 1.0
#endif

template <class T>
__device__ inline T pc_first(int pc_type, T omega_pc, T b, T ac, bool black) {
    if (pc_type == FALM_LS_SOR) {
        // kernel_Struct3d7p_SOR from x = 0: cc = (b - 0)/ac on the black
        // colour, x = 0 + omega*cc (src/dev/FalmEqDevCall.cu:204-209)
        return black ? T(0) + omega_pc * (b / ac) : T(0);
    } else if (pc_type == FALM_LS_JACOBI) {
        // kernel_Struct3d7p_Jacobi from xp = 0 (src/dev/FalmEqDevCall.cu:130)
        return T(0) + b / ac;
    }
    return b;  // identity (see falm_solver_create)
}

// ------------------------------------------------------------ kernels
// r64 = b - A x in double; float/double Krylov copies r, rr; partials of
// |r64|^2 and (r, rr)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_bicg_init(Tiling tl, Geom g, OP a, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *r64, T *r, T *rr,
                                                  const DevState *st, double *partials) {
    if (st->final || !st->done) return;  // a pass is still running, or the solve is over
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k, t0, t1;
    double v[2] = {0.0, 0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        const double rv = b[cc] - apply_op<OP, double, double, double>(a, x, cc, i, j, k, g.nx, g.nxy);
        if (r64) r64[cc] = rv;
        const T rt = (T)rv;
        if ((void *)r != (void *)r64) r[cc] = rt;
        rr[cc] = rt;
        v[0] += rv * rv;
        v[1] += (double)rt * (double)rt;
    }
    block_partials<NT, 2>(v, partials, lds);
}

// Start of a pass (src/FalmEq.cpp:318-351).  In mixed mode this is also the
// outer loop's test: the true residual of the accumulated x ends the solve
// once a pass has run, and otherwise sets the float pass's target.
// k-marching forms of the solver's stencil kernels (x walked as a Plane7):
// same arithmetic per cell, partial sums per workgroup in the K7 order.
// Eight waves per SIMD (64 VGPRs at most): four workgroups per CU, so the
// 1024 tiles of a 512 x 128 x 128 interior run in one round instead of
// three workgroups per CU and a partly empty second round.
// CHECK: only |b - A x|^2 and, if r64 is given, the residual (the mixed
// solve's test after a pass; a pass that must follow starts with the full
// kernel).
template <class T, class OP, bool CHECK = false>
__global__ __launch_bounds__(K7_NT) __attribute__((amdgpu_waves_per_eu(8))) void k_bicg_init_km(
    K7Tiling kt, Geom g, OP a, const double *__restrict__ x,
                                                        const double *__restrict__ b, double *r64, T *r, T *rr,
                                                        const DevState *st, double *partials) {
    if (st->final || !st->done) return;
    __shared__ K7Lds<double> lx;
    __shared__ double lds[2 * (K7_NT / 64)];
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    Plane7<double> px;
    px.setup(x, &lx, Q, g);
    px.prologue(Q, sz);
    double v[2] = {0.0, 0.0};
    int bf = 0;
    const auto ca = a.column(Q.act ? Q.i : kt.box.lo[0], Q.act ? Q.j : kt.box.lo[1]);
    const unsigned szd = (unsigned)sz * 8u, szt = (unsigned)sz * sizeof(T);
    unsigned od = (unsigned)Q.col0 * 8u + (unsigned)Q.k0 * szd, ot = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szt;
    // b and the operator's cell coefficients are loaded a plane ahead too
    auto cq = ca.load(od, Q.k0);
    double bq = ld_at(b, od);
    __syncthreads();
    for (int k = Q.k0; k < Q.k1; k++, od += szd, ot += szt) {
        const bool more = k + 1 < Q.k1;
        const auto cc = cq;
        const double bc = bq;
        if (more) {
            px.prefetch(Q, k, sz);
            cq = ca.load(od + szd, k + 1);
            bq = ld_at(b, od + szd);
        }
        if (Q.act) {
            const double rv = bc - ca.apply(cc, px.c(), px.xp(bf), px.xm(bf), px.yp(bf), px.ym(bf), px.zp(), px.zm());
            v[0] += rv * rv;
            if (!CHECK) {
                if (r64) st_at(r64, od, rv);
                const T rt = (T)rv;
                if (r && (void *)r != (void *)r64) st_at(r, ot, rt);
                st_at(rr, ot, rt);
                v[1] += (double)rt * (double)rt;
            } else if (r64) {
                st_at(r64, od, rv);
            }
        }
        px.advance(bf, more);
        bf ^= 1;
        __syncthreads();
    }
    block_partials<K7_NT, 2>(v, partials, lds);
}

template <class T, class OP>
__global__ __launch_bounds__(K7_NT) void k_sor_op_km(K7Tiling kt, Geom g, OP a, T *x, const T *__restrict__ b,
                                                     T omega, int color, const int *done) {
    if (*done) return;
    __shared__ K7Lds<T> lx;
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    Plane7<T> px;
    px.setup(x, &lx, Q, g);
    px.prologue(Q, sz);
    int bf = 0;
    const auto ca = a.column(Q.act ? Q.i : kt.box.lo[0], Q.act ? Q.j : kt.box.lo[1]);
    const unsigned szt = (unsigned)sz * sizeof(T);
    unsigned ot = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szt;
    auto cq = ca.load(ot, Q.k0);
    T bq = ld_at(b, ot);
    __syncthreads();
    for (int k = Q.k0; k < Q.k1; k++, ot += szt) {
        const bool more = k + 1 < Q.k1;
        const auto cc = cq;
        const T bc = bq;
        if (more) {
            px.prefetch(Q, k, sz);
            cq = ca.load(ot + szt, k + 1);
            bq = ld_at(b, ot + szt);
        }
        if (Q.act && ((Q.i + Q.j + k + g.osum) & 1) == color) {
            const T xc = px.c();
            // kernel_Struct3d7p_SOR's update (src/dev/FalmEqDevCall.cu:195-215)
            const T s = ca.apply(cc, xc, px.xp(bf), px.xm(bf), px.yp(bf), px.ym(bf), px.zp(), px.zm());
            const T cv = (bc - s) / decltype(ca)::diag(cc);
            st_at(x, ot, xc + omega * cv);
        }
        px.advance(bf, more);
        bf ^= 1;
        __syncthreads();
    }
}

// two vectors' elements of one cell side by side
template <class T> struct PairOf;
template <> struct PairOf<float> { using type = float2; };
template <> struct PairOf<double> { using type = double2; };
template <class T> using PairT = typename PairOf<T>::type;
template <class T> __device__ __forceinline__ PairT<T> pair_of(T a, T b) {
    PairT<T> v;
    v.x = a;
    v.y = b;
    return v;
}

// The end of a BiCGStab pass with the one-pass SOR preconditioner M (black
// half-sweep from zero, then the red one; k_bicg_pcmv_km below): the pass's
// solution update is x += M(z) with z = (0 or y) + alpha p + omega s, y the
// pass's earlier alpha p + omega s (present from the second iteration on) and
// alpha p + omega s the last iteration's, summed as the p kernel sums y.  M is
// linear, so this is what adding alpha M(p) + omega M(s) every iteration
// gives, up to rounding.  p is in p0 or p1 (the last p kernel recorded which);
// rout, when given, receives the last residual s - omega t.  One k-marching
// pass: z is formed on the tile and its one-cell halo from y, p, s at each
// cell, the black half-sweep's values go to LDS (x and y neighbours) and a
// register queue (k neighbours), and every red cell is swept from its six
// black neighbours, cells outside the interior contributing zero.  It runs
// once a pass is over and its update not yet in x (the init kernels'
// finalize clears that).
template <class T, class OP>
__global__ __launch_bounds__(K7_NT) void k_bicg_close_km(K7Tiling kt, Geom g, OP a, const T *__restrict__ y,
                                                         const T *__restrict__ p0, const T *__restrict__ p1,
                                                         const PairT<T> *__restrict__ stv, T *__restrict__ rout,
                                                         const DevState *st, T omega_pc, double *__restrict__ xd,
                                                         DevState *hm) {
    using P = PairT<T>;
    // hm: the host's pinned mirror of the state, written when the closing
    // kernel ends a batch of the double solve (nothing after it changes
    // the state)
    if (hm && blockIdx.x == 0 && threadIdx.x == 0) *hm = *st;
    if (!st->done || st->final || !st->pending) return;
    __shared__ K7Lds<T> lw;
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    const double ya = st->alpha, yw = st->omega;
    const bool yread = st->it_pass >= 2;
    const T *const pv = st->plast ? p1 : p0;
    const T om = (T)yw;
    const auto vec = [&](T yy, T pp, T ss) -> T {
        return (T)((yread ? (double)yy : 0.0) + (ya * (double)pp + yw * (double)ss));
    };
    const auto ca = a.column(Q.act ? Q.i : kt.box.lo[0], Q.act ? Q.j : kt.box.lo[1]);
    using Col = typename std::decay<decltype(ca)>::type;
    using Cell = typename Col::Cell;
    static_assert(sizeof(decltype(ca.diag_at(0u))) == sizeof(T), "operator and vector precision differ");
    const int lo0 = kt.box.lo[0], hi0 = lo0 + kt.box.n[0], lo1 = kt.box.lo[1], hi1 = lo1 + kt.box.n[1];
    const int lo2 = kt.box.lo[2], hi2 = lo2 + kt.box.n[2];
    const K7Halo h = k7_halo(Q);
    const bool hok = h.slot && h.gi >= 0 && h.gi < g.nx && h.gj >= 0 && h.gj < g.ny;
    const bool hin = hok && h.gi >= lo0 && h.gi < hi0 && h.gj >= lo1 && h.gj < hi1;
    const unsigned szt = (unsigned)sz * sizeof(T);
    const unsigned cb = (unsigned)Q.col0 * sizeof(T);
    const unsigned hb = hok ? (unsigned)(h.gi + g.nx * h.gj) * sizeof(T) : cb;
    const int cpar = (Q.i + Q.j + g.osum) & 1, hpar = (h.gi + h.gj + g.osum) & 1;
    // the operator's view of the halo cell's column (a valid column when the
    // halo cell lies outside the array; its values are then unused)
    const auto ch = a.column(hok ? h.gi : (Q.act ? Q.i : kt.box.lo[0]), hok ? h.gj : (Q.act ? Q.j : kt.box.lo[1]));
    // the black half-sweep's value (zero on red cells and outside the interior)
    const auto wval = [&](T vv, T d, bool in, int par, int k) -> T {
        return (in && ((par + k) & 1) == FALM_COLOR_BLACK) ? T(0) + omega_pc * (vv / d) : T(0);
    };
    const auto kin = [&](int k) { return k >= lo2 && k < hi2; };
    const auto load = [&](unsigned ob, T &yy, T &pp, T &ss) {
        yy = ld_at(y, ob);
        pp = ld_at(pv, ob);
        ss = ld_at(stv, 2 * ob).x;
    };
    // registers: w at k-1 and k, z and the cell coefficients at k, the
    // column's raw values at k+1 and the halo cell's at the next LDS plane
    T wm, wc, vc;
    Cell ec;
    T ny, np, ns;
    Cell ne;
    T hy = T(0), hp = T(0), hs = T(0), hd = T(1);
    T sc;  // the column's s at k (rout)
    {
        unsigned ob = cb + (unsigned)(Q.k0 - 1) * szt;
        T ym, pm, sm, yc, pc, s0;
        load(ob, ym, pm, sm);
        const T dm = ca.diag_k(ob, Q.k0 - 1);
        load(ob + szt, yc, pc, s0);
        ec = ca.load(ob + szt, Q.k0);
        load(ob + 2 * szt, ny, np, ns);
        ne = ca.load(ob + 2 * szt, Q.k0 + 1);
        if (h.slot) {
            const unsigned oh = hb + (unsigned)Q.k0 * szt;
            load(oh, hy, hp, hs);
            hd = ch.diag_k(oh, Q.k0);
        }
        wm = wval(vec(ym, pm, sm), dm, Q.act && kin(Q.k0 - 1), cpar, Q.k0 - 1);
        vc = vec(yc, pc, s0);
        sc = s0;
        wc = wval(vc, Col::diag(ec), Q.act, cpar, Q.k0);
        lw[0][Q.wy + 1][Q.lane + 1] = wc;
        if (h.slot) lw[0][h.row][h.col] = wval(vec(hy, hp, hs), hd, hin, hpar, Q.k0);
    }
    // x a plane ahead (act columns only); with rout, t a plane ahead
    constexpr unsigned XS = (unsigned)(sizeof(double) / sizeof(T));
    double xq = 0.0;
    if (Q.act) xq = ld_at(xd, (cb + (unsigned)Q.k0 * szt) * XS);
    const bool rw = rout != nullptr;
    T tq = T(0);
    if (rw) tq = ld_at(stv, 2 * (cb + (unsigned)Q.k0 * szt)).y;
    __syncthreads();
    int bf = 0;
    unsigned ob = cb + (unsigned)Q.k0 * szt;
    for (int k = Q.k0; k < Q.k1; k++, ob += szt) {
        const bool more = k + 1 < Q.k1;
        const double xc = xq;
        if (more && Q.act) xq = ld_at(xd, (ob + szt) * XS);
        const T tc = tq;
        if (rw && more) tq = ld_at(stv, 2 * (ob + szt)).y;
        T y2 = T(0), p2 = T(0), s2 = T(0);
        Cell e2 = ne;
        if (more) {
            load(ob + 2 * szt, y2, p2, s2);
            e2 = ca.load(ob + 2 * szt, k + 2);
            if (h.slot) {
                const unsigned oh = hb + (unsigned)(k + 1) * szt;
                load(oh, hy, hp, hs);
                hd = ch.diag_k(oh, k + 1);
            }
        }
        const T vn = vec(ny, np, ns);
        const T wn = wval(vn, Col::diag(ne), Q.act && kin(k + 1), cpar, k + 1);
        if (Q.act) {
            // the last residual, kernel_PBiCGStab_4's expression
            if (rw) st_at(rout, ob, sc - om * tc);
            T out = wc;
            if (((cpar + k) & 1) != FALM_COLOR_BLACK) {
                // kernel_Struct3d7p_SOR's update from x = 0 (src/dev/FalmEqDevCall.cu:195-215)
                const T s = ca.apply(ec, T(0), lw[bf][Q.wy + 1][Q.lane + 2], lw[bf][Q.wy + 1][Q.lane],
                                     lw[bf][Q.wy + 2][Q.lane + 1], lw[bf][Q.wy][Q.lane + 1], wn, wm);
                out = T(0) + omega_pc * ((vc - s) / Col::diag(ec));
            }
            st_at(xd, ob * XS, xc + (double)out);
        }
        if (more) {
            lw[bf ^ 1][Q.wy + 1][Q.lane + 1] = wn;
            if (h.slot) lw[bf ^ 1][h.row][h.col] = wval(vec(hy, hp, hs), hd, hin && kin(k + 1), hpar, k + 1);
        }
        wm = wc;
        wc = wn;
        vc = vn;
        ec = ne;
        sc = ns;
        ny = y2;
        np = p2;
        ns = s2;
        ne = e2;
        bf ^= 1;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// A Krylov vector, its one-pass SOR preconditioning and the operator product
// in one k-marching pass (pc_maxit = 1; the two kernels of a BiCGStab
// iteration, src/FalmEq.cpp:345-407):
//   MODE 0: r = s - omega t (the previous iteration's residual, formed here
//           and stored; a pass's first iteration reads r instead),
//           v = r (first iteration) or r + beta (p - omega q),
//           w = A M(v), partial sums of (rr, w)          -> alpha;
//           y = (0 or y) + alpha p + omega s of the previous iteration
//   MODE 1: v = r - alpha q,  w = A M(v), partial sums of (w, v), (w, w),
//           (v, v), (v, rr), (w, rr)       -> omega, |r|, (r, rr) of the
//           next residual r = v - omega w (k_fin_fold)
// M(v) is the black half-sweep from zero followed by the red one (black
// cells omega_pc v / a_c, red cells omega_pc (v - six black neighbours'
// products) / a_c, zero outside the interior; kernel_Struct3d7p_SOR's
// expression, src/dev/FalmEqDevCall.cu:195-215), so w is, bit for bit, what
// the separate half-sweeps and k_bicg_mv_km give.  w at a tile cell needs
// M(v) on the tile and its four one-cell strips in x and y; a red value
// there needs black values one cell further out; so v is formed on the tile
// plus a two-cell ring (PQ_W x PQ_H cells a plane) from its operands read
// there.  Planes pass through four LDS slots: at step s the slot of plane
// s+2 is filled (black cells M(v), red cells v), the red cells of plane s
// are turned into M(v) in place (they read black values of planes s-1..s+1
// only), and w of plane s-1 is formed from M(v) of plane s-1 in LDS and of
// the column's planes s-2 and s in registers: one barrier per step.
// v is written on the tile (p or s) and w (q or t); M(v) is not stored:
// y accumulates the pass's alpha p + omega s and k_bicg_close_km adds M of
// the sum to x at the end of the pass (M is linear).  MODE 0 reads p and q
// at the ring cells, which neighbouring workgroups write: both alternate
// between two buffers.  s and t, which the next MODE 0 and k_bicg_close_km
// read together, are stored as pairs of adjacent elements (one 8-byte
// access in float, 16 in double): MODE 1 writes the pair at plane s-1 with
// the column's s kept in registers.  Every pass's first iteration is the
// FIRST instance (the host starts every pass).
constexpr int PQ_R = 2;                                    // ring width
constexpr int PQ_W = K7_BX + 2 * PQ_R, PQ_H = K7_BY + 2 * PQ_R;
constexpr int PQ_RING2 = PQ_W * PQ_H - K7_NT;              // ring cells of v
constexpr int PQ_RING1 = 2 * K7_BX + 2 * K7_BY;            // red-sweep cells outside the tile
constexpr int PQ_R1T0 = K7_NT - PQ_RING1;                  // first thread of those
static_assert(PQ_RING2 <= K7_NT && PQ_R1T0 >= PQ_RING2 - 2 * PQ_W, "ring cells exceed threads");
static_assert((PQ_R1T0 + 2 * K7_BY) % 64 == 0, "ring-1 rows start on a wave");
template <class T>
using PQLds = T[4][PQ_H][PQ_W];
// k_bicg_pcmv_km runs one round of workgroups, pcmv_wg_per_cu of them
// resident on each CU: its chunk is the shortest (at least K7_KC planes)
// whose grid fits that many.  Float kernels run two 8-wave workgroups per CU
// (at most 128 VGPRs, which the separable-operator instances are held to);
// double ones one, with 64-plane chunks at 512 x 128 x 128 (measured as fast
// as two per CU, and the folded double s kernel needs more than 128 VGPRs).
#ifndef FALM_PCMV_WG_F32
#define FALM_PCMV_WG_F32 2
#endif
#ifndef FALM_PCMV_WG_F64
#define FALM_PCMV_WG_F64 1
#endif
constexpr int pcmv_wg_per_cu(bool dbl) { return dbl ? FALM_PCMV_WG_F64 : FALM_PCMV_WG_F32; }
template <class OP> struct PcmvWaves { static constexpr int value = 1; };
template <class TA> struct PcmvWaves<OpSep<TA>> {
    static constexpr int value = 2 * pcmv_wg_per_cu(sizeof(TA) == 8);
};

// region position (rx, ry) of thread s's ring cell (s < PQ_RING2): two rows
// below the tile, two above, then the two columns on either side
__device__ inline void pq_ring2(int s, int &rx, int &ry) {
    if (s < 2 * PQ_W) {
        ry = s / PQ_W;
        rx = s - ry * PQ_W;
    } else if (s < 4 * PQ_W) {
        const int u = s - 2 * PQ_W, q = u / PQ_W;
        ry = PQ_H - PQ_R + q;
        rx = u - q * PQ_W;
    } else {
        const int u = s - 4 * PQ_W, c = u & 3;
        ry = PQ_R + (u >> 2);
        rx = c < 2 ? c : PQ_W - 4 + c;
    }
}
// region position of the red-sweep cell outside the tile of thread s
// (s >= PQ_R1T0): the columns beside the tile, the row below, the row above
__device__ inline void pq_ring1(int s, int &rx, int &ry) {
    const int u = s - PQ_R1T0;
    if (u < 2 * K7_BY) {
        ry = PQ_R + (u >> 1);
        rx = (u & 1) ? PQ_W - PQ_R : PQ_R - 1;
    } else if (u < 2 * K7_BY + K7_BX) {
        ry = PQ_R - 1;
        rx = PQ_R + (u - 2 * K7_BY);
    } else {
        ry = PQ_H - PQ_R;
        rx = PQ_R + (u - 2 * K7_BY - K7_BX);
    }
}

// The diagonal and the cell coefficients at a byte offset, with the
// diagonal-class choice fixed at compile time (DCLS: from the column's
// per-run values, else the stored column), so every load is unconditional.
template <bool DCLS, class TA>
__device__ inline TA diag_of(const OpSepCol<TA> &c, unsigned ob, int k) {
    if constexpr (DCLS) return c.pick(k);
    else return ld_at(c.ac, ob);
}
template <bool DCLS, class TA>
__device__ inline TA diag_of(const OpCols<TA> &c, unsigned ob, int) {
    return ld_at(c.c[0], ob);
}
template <bool DCLS, class TA>
__device__ inline typename OpSepCol<TA>::Cell cell_of(const OpSepCol<TA> &c, unsigned ob, int k) {
    return typename OpSepCol<TA>::Cell{diag_of<DCLS>(c, ob, k), ld_uniform(c.ab, k), ld_uniform(c.at, k)};
}
template <bool DCLS, class TA>
__device__ inline typename OpCols<TA>::Cell cell_of(const OpCols<TA> &c, unsigned ob, int k) {
    return c.load(ob, k);
}

// Every global load and store of a step is issued by every thread (loads at
// clamped valid cells, stores of threads with nothing to store go to the
// thread's slot in `trash`): the compiler can then count the memory
// operations in flight exactly and waits for the previous step's loads
// only, not for those just issued.  trash holds K7_NT values per workgroup.
// stin: the (s, t) pairs MODE 0 reads; stout: those MODE 1 writes.
template <class T, class OP, int MODE, bool DCLS, bool FIRST>
__global__ __launch_bounds__(K7_NT) __attribute__((amdgpu_waves_per_eu(PcmvWaves<OP>::value))) void k_bicg_pcmv_km(
    K7Tiling kt, Geom g, OP a, const T *__restrict__ r, const T *__restrict__ pin, const T *__restrict__ qin,
    const T *__restrict__ rr, T *__restrict__ vout, T *__restrict__ wout, const DevState *st, T omega_pc,
    const double *apart, int anb, DevState *stw, double *partials, T *__restrict__ trash,
    const PairT<T> *__restrict__ stin, PairT<T> *__restrict__ stout, T *__restrict__ rout, T *__restrict__ yv,
    int pid) {
    using P = PairT<T>;
    if (st->done) return;
    __shared__ PQLds<T> lv;
    __shared__ double lds[5 * (K7_NT / 64)];
    const K7Pos Q = k7_pos(kt, g);
    const int tid = (int)threadIdx.x;
    const unsigned szt = (unsigned)g.nxy * sizeof(T);
    // q (and p) are used; FIRST: the launch is a pass's first iteration
    constexpr bool upd = MODE == 1 || !FIRST;
    T c1 = T(0), c2 = T(0);
    if (MODE == 0) {
        c1 = (T)st->beta;
        c2 = (T)st->omega;
    }  // MODE 1: alpha is summed below, while the prologue's loads are in flight
    // MODE 0: the previous iteration's alpha, omega for y; y is set
    // (not accumulated) on the pass's second iteration
    const double ya = MODE == 0 ? st->alpha : 0.0, yw = MODE == 0 ? st->omega : 0.0;
    const bool yset = MODE == 0 && st->it_pass == 1;
    if (MODE == 0 && blockIdx.x == 0 && threadIdx.x == 0) stw->plast = pid;
    const int lo0 = kt.box.lo[0], hi0 = lo0 + kt.box.n[0], lo1 = kt.box.lo[1], hi1 = lo1 + kt.box.n[1];
    const int lo2 = kt.box.lo[2], hi2 = lo2 + kt.box.n[2];
    const int k0 = Q.k0, k1 = Q.k1;
    const auto inxy = [&](int i, int j) { return i >= lo0 && i < hi0 && j >= lo1 && j < hi1; };
    const auto kin = [&](int k) { return k >= lo2 && k < hi2; };
    // a plane to load from: inside the interior and at most k1 + 1
    const auto kcl = [&](int k) {
        k = k > k1 + 1 ? k1 + 1 : k;
        return k < lo2 ? lo2 : (k >= hi2 ? hi2 - 1 : k);
    };
    const auto black = [&](int i, int j, int k) { return ((i + j + k + g.osum) & 1) == FALM_COLOR_BLACK; };
    // this thread's cells: its column (region position ro), a ring cell of
    // v (r2) and a red-sweep cell outside the tile (r1); positions outside
    // the interior read the thread's own (valid) column instead
    const int oi = Q.act ? Q.i : lo0, oj = Q.act ? Q.j : lo1;
    const int rxo = Q.lane + PQ_R, ryo = Q.wy + PQ_R;
    const auto ca = a.column(oi, oj);
    using Col = typename std::decay<decltype(ca)>::type;
    using Cell = typename Col::Cell;
    static_assert(sizeof(decltype(ca.diag_at(0u))) == sizeof(T), "operator and vector precision differ");
    const unsigned cb = (unsigned)(oi + g.nx * oj) * sizeof(T);
    int rx2 = 0, ry2 = 0, rx1 = 0, ry1 = 0;
    const bool r2 = tid < PQ_RING2, r1 = tid >= PQ_R1T0;
    if (r2) pq_ring2(tid, rx2, ry2);
    if (r1) pq_ring1(tid, rx1, ry1);
    const int gi2 = Q.i0 - PQ_R + rx2, gj2 = Q.j0 - PQ_R + ry2;
    const int gi1 = Q.i0 - PQ_R + rx1, gj1 = Q.j0 - PQ_R + ry1;
    const bool in2 = r2 && inxy(gi2, gj2), in1 = r1 && inxy(gi1, gj1);
    // colours at even planes (a cell's colour flips with the plane's parity)
    const bool beo = black(Q.i, Q.j, 0), be2 = black(gi2, gj2, 0), be1 = black(gi1, gj1, 0);
    const auto c2v = a.column(in2 ? gi2 : oi, in2 ? gj2 : oj);
    const auto c1v = a.column(in1 ? gi1 : oi, in1 ? gj1 : oj);
    const unsigned b2 = in2 ? (unsigned)(gi2 + g.nx * gj2) * sizeof(T) : cb;
    const unsigned b1 = in1 ? (unsigned)(gi1 + g.nx * gj1) * sizeof(T) : cb;
    // trash slots: T and pair stores share them (2 K7_NT T per workgroup)
    T *const tr = trash + (size_t)blockIdx.x * (2 * K7_NT) + tid;
    P *const tr2 = reinterpret_cast<P *>(trash) + (size_t)blockIdx.x * K7_NT + tid;

    // operands of v at one cell of a plane, and its diagonal: a is r, or in
    // MODE 0 after a pass's first iteration s (and b is t: r = s - omega t)
    struct Ld {
        T a, b, p, q, d;
    };
    const auto load = [&](const Col &c, unsigned ob, int k) {
        Ld x;
        x.b = T(0);
        if constexpr (MODE == 0 && !FIRST) {
            const P sp = ld_at(stin, 2 * ob);
            x.a = sp.x;
            x.b = sp.y;
        } else {
            x.a = ld_at(r, ob);
        }
        x.q = (MODE == 1 || !FIRST) ? ld_at(qin, ob) : T(0);
        x.p = (MODE == 0 && !FIRST) ? ld_at(pin, ob) : T(0);
        x.d = diag_of<DCLS>(c, ob, k);
        return x;
    };
    // the residual a cell's v is formed from (MODE 0: r = s - omega t, the
    // previous iteration's kernel_PBiCGStab_4 with one rounding)
    const auto res = [&](const Ld &x) -> T {
        if (MODE == 0) return upd ? fma(-c2, x.b, x.a) : x.a;
        return x.a;
    };
    const auto vec_of = [&](const Ld &x, T rv) -> T {
        if (MODE == 1) return rv - c1 * x.q;
        return upd ? rv + c1 * (x.p - c2 * x.q) : rv;
    };
    const auto vec = [&](const Ld &x) -> T { return vec_of(x, res(x)); };
    struct Fl {
        Ld o, h;  // the column's cell and the ring cell
        T y;      // MODE 0: the column's y
    };
    const auto fetch = [&](int f) {
        const int kk = kcl(f);
        const unsigned pf = (unsigned)kk * szt;
        Fl x;
        x.o = load(ca, cb + pf, kk);
        x.h = load(c2v, b2 + pf, kk);
        x.y = (MODE == 0 && !FIRST) ? ld_at(yv, cb + pf) : T(0);
        return x;
    };
    // slot of plane f: black cells M(v), red cells v, zero outside the
    // interior; returns the column's v at plane f.  MODE 0 stores v (p), r
    // and y = (0 or y) + alpha p + omega s of the previous iteration (what
    // the update kernel accumulates; a pass's first iteration leaves y as it
    // is: the second sets it); MODE 1's v (s) is stored with w (t) three
    // planes later.
    const auto own = [&](int f, const Fl &x, bool kf) -> T {
        const T rv = res(x.o);
        const T vo = (Q.act && kf) ? vec_of(x.o, rv) : T(0);
        if (MODE == 0) {
            const bool wr = Q.act && f >= k0 && f < k1;
            const unsigned ob = cb + (unsigned)f * szt;
            *(wr ? (T *)((char *)vout + ob) : tr) = vo;
            if (!FIRST) {
                *(wr ? (T *)((char *)rout + ob) : tr) = rv;
                const double yo = (yset ? 0.0 : (double)x.y) + (ya * (double)x.o.p + yw * (double)x.o.a);
                *(wr ? (T *)((char *)yv + ob) : tr) = (T)yo;
            }
        }
        return vo;
    };
    const auto fill = [&](int f, const Fl &x) -> T {
        const int sl = f & 3;
        const bool kf = kin(f);
        const T vo = own(f, x, kf);
        lv[sl][ryo][rxo] = black(Q.i, Q.j, f) ? T(0) + omega_pc * (vo / x.o.d) : vo;
        if (r2) {
            const T vh = (in2 && kf) ? vec(x.h) : T(0);
            lv[sl][ry2][rx2] = black(gi2, gj2, f) ? T(0) + omega_pc * (vh / x.h.d) : vh;
        }
        return vo;
    };
    // Step s's SOR work: plane s+2 into its slot (as fill) and the red cells
    // of plane s (tile and its strips) to M(v), in place; returns the
    // column's v at s+2 and sets h to its M(v) at s.  A cell has the same
    // colour at s and s+2, so each lane divides once: black lanes for plane
    // s+2, red lanes for plane s; a ring thread owns one ring cell (a v ring
    // cell, black at s+2, or a red-sweep cell, red at s), never both.  Black
    // lanes also form the red sum from their neighbours in slot s, which red
    // lanes rewrite in this step: that sum is not used.  The red update is
    // kernel_Struct3d7p_SOR's from x = 0 (src/dev/FalmEqDevCall.cu:195-215).
    const auto fill_red = [&](int s, const Fl &x, const Cell &eo, const Cell &e1, T &h) -> T {
        const int f = s + 2;
        const int sf = f & 3, sl = s & 3, sm = (s - 1) & 3, sp = (s + 1) & 3;
        const bool kf = kin(f), km = kin(s), odd = (s & 1) != 0;
        const T vo = own(f, x, kf);
        const bool bo = odd ? !beo : beo;
        const T ho = lv[sl][ryo][rxo];
        const T so = ca.apply(eo, T(0), lv[sl][ryo][rxo + 1], lv[sl][ryo][rxo - 1], lv[sl][ryo + 1][rxo],
                              lv[sl][ryo - 1][rxo], lv[sp][ryo][rxo], lv[sm][ryo][rxo]);
        const T qo = T(0) + omega_pc * ((bo ? vo : ho - so) / (bo ? x.o.d : Col::diag(eo)));
        lv[sf][ryo][rxo] = bo ? qo : vo;
        const bool ro = !bo && Q.act && km;
        if (ro) lv[sl][ryo][rxo] = qo;
        h = ro ? qo : ho;
        if (r2 || r1) {
            T num, den;
            bool w2 = false, w1 = false;
            if (r2) {
                const T vh = (in2 && kf) ? vec(x.h) : T(0);
                w2 = odd ? !be2 : be2;
                if (!w2) lv[sf][ry2][rx2] = vh;
                num = vh;
                den = x.h.d;
            } else {
                const T v1 = lv[sl][ry1][rx1];
                const T s1 = c1v.apply(e1, T(0), lv[sl][ry1][rx1 + 1], lv[sl][ry1][rx1 - 1], lv[sl][ry1 + 1][rx1],
                                       lv[sl][ry1 - 1][rx1], lv[sp][ry1][rx1], lv[sm][ry1][rx1]);
                w1 = in1 && km && !(odd ? !be1 : be1);
                num = v1 - s1;
                den = Col::diag(e1);
            }
            const T q = T(0) + omega_pc * (num / den);
            if (w2) lv[sf][ry2][rx2] = q;
            if (w1) lv[sl][ry1][rx1] = q;
        }
        return vo;
    };
    const auto cell_at = [&](const Col &c, unsigned base, int k) {
        const int kk = kcl(k);
        return cell_of<DCLS>(c, base + (unsigned)kk * szt, kk);
    };

    // prologue: planes k0-2 .. k0 in their slots, plane k0+1's loads in flight
    const Fl f0 = fetch(k0 - 2), f1 = fetch(k0 - 1), f2 = fetch(k0);
    Fl fA = fetch(k0 + 1), fB{};  // fB is loaded in the first step
    Cell eA = cell_at(ca, cb, k0 - 1), eB = eA, e1A = cell_at(c1v, b1, k0 - 1), e1B = e1A;
    T rA = T(0), rB = T(0);
    if (MODE == 1) {
        // alpha = rho / (rr, q) from the first kernel's partial sums (what
        // k_fin_alpha computes), in every workgroup; the first stores it
        double tot[1];
        sum_partials<1, K7_NT>(apart, anb, tot);
        const double alpha = st->rho / tot[0];
        if (blockIdx.x == 0 && threadIdx.x == 0) stw->alpha = alpha;
        c1 = (T)alpha;
    }
    // at the start of step s: the column's v at planes s-1, s, s+1 (v0, v1,
    // v2; MODE 1's dot (w, v)) and M(v) at planes s-2, s-1 (hm, hc)
    T v0 = fill(k0 - 2, f0), v1 = fill(k0 - 1, f1), v2 = fill(k0, f2);
    T hm = T(0), hc = T(0);
    __syncthreads();
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // Step s with two buffer sets: the `use` set holds what step s-1 loaded
    // (plane s+2's operands, rr at plane s-1, the cells of plane s; eo holds
    // plane s-1's cells), the `load` set receives plane s+3's operands, rr at
    // plane s and the cells of plane s+1.  The loop runs two steps per trip
    // with the sets swapped, so a value is first touched one step after its
    // load was issued: rotating one set of registers at the end of every
    // step would wait for the loads just issued.  Past the chunk's last
    // plane the fills go to a slot no later step reads, and their stores to
    // trash.
    const auto step = [&](int s, const Fl &fu, Fl &fl, T ru, T &rl, const Cell &eu, Cell &eo, const Cell &e1u,
                          Cell &e1l) {
        const Cell eq = eo;
        fl = fetch(s + 3);
        eo = cell_at(ca, cb, s + 1);
        e1l = cell_at(c1v, b1, s + 1);
        {
            const int k = s < k0 ? k0 : (s >= k1 ? k1 - 1 : s);
            rl = ld_at(rr, cb + (unsigned)k * szt);
        }
        T h;
        const T vn = fill_red(s, fu, eu, e1u, h);
        // w at plane s-1 (the column's M(v): s-2, s-1, s)
        const int sl = (s - 1) & 3;
        const T wv = ca.apply(eq, hc, lv[sl][ryo][rxo + 1], lv[sl][ryo][rxo - 1], lv[sl][ryo + 1][rxo],
                              lv[sl][ryo - 1][rxo], h, hm);
        const bool wr = Q.act && s - 1 >= k0;
        if (MODE == 0) *(wr ? (T *)((char *)wout + cb + (unsigned)(s - 1) * szt) : tr) = wv;
        else *(wr ? (P *)((char *)stout + 2 * (cb + (unsigned)(s - 1) * szt)) : tr2) = pair_of(v0, wv);
        if (MODE == 0) {
            acc[0] += wr ? (double)ru * (double)wv : 0.0;
        } else {
            acc[0] += wr ? (double)wv * (double)v0 : 0.0;
            acc[1] += wr ? (double)wv * (double)wv : 0.0;
            // (s, s), (s, rr), (t, rr): |r|^2 and (r, rr) of the next
            // residual r = s - omega t follow from these and the two above
            acc[2] += wr ? (double)v0 * (double)v0 : 0.0;
            acc[3] += wr ? (double)v0 * (double)ru : 0.0;
            acc[4] += wr ? (double)wv * (double)ru : 0.0;
        }
        hm = hc;
        hc = h;
        v0 = v1;
        v1 = v2;
        v2 = vn;
        __syncthreads();
    };
    int s = k0 - 1;
    for (; s + 1 <= k1; s += 2) {
        step(s, fA, fB, rA, rB, eA, eB, e1A, e1B);
        step(s + 1, fB, fA, rB, rA, eB, eA, e1B, e1A);
    }
    if (s <= k1) step(s, fA, fB, rA, rB, eA, eB, e1A, e1B);
    if (MODE == 0) {
        double v1s[1] = {acc[0]};
        block_partials<K7_NT, 1>(v1s, partials, lds);
    } else {
        block_partials<K7_NT, 5>(acc, partials, lds);
    }
}

// ---------------------------------------------------------------------
// Three red-black SOR half-sweeps (colours c0, 1-c0, c0) in one pass over
// memory, out of place (x_in -> x_out).  A workgroup owns an S3_BX x S3_BY
// column of cells and marches its planes in k with the three sweeps skewed:
// while plane s gets sweep 1, plane s-1 gets sweep 2 and plane s-2 sweep 3,
// each in LDS.  A cell's update reads its six neighbours, which have the
// other colour, so sweep 2 at plane s-1 needs sweep-1 values on planes
// s-2..s and sweep 3 at s-2 needs sweep-2 values on s-3..s-1: the skew gives
// both.  In x and y the sweeps are computed redundantly on rings of the
// tile (sweep 1 on the tile plus 2 cells, sweep 2 plus 1, sweep 3 on the
// tile), from the same input values a neighbouring workgroup uses, so the
// written cells equal those of three separate sweeps bit for bit.  The same
// holds in k across the chunk ends.  Per cell this reads x, b and the
// diagonal once and writes x once, where three separate half-sweeps read
// and write x three times.
constexpr int S3_BX = 64, S3_BY = 8, S3_NT = S3_BX * S3_BY;
constexpr int S3_R = 3;                     // x ring (sweep 1 on ring 2 reads ring 3)
constexpr int S3_W = S3_BX + 2 * S3_R, S3_H = S3_BY + 2 * S3_R;
constexpr int S3_PL = S3_W * S3_H;          // x cells per plane
constexpr int S3_NXP = 5;                   // x planes in flight: s+2 .. s-2
constexpr int S3_RB = 2;                    // b / diagonal ring
constexpr int S3_BW = S3_BX + 2 * S3_RB, S3_BH = S3_BY + 2 * S3_RB;
constexpr int S3_BPL = S3_BW * S3_BH;       // b cells per plane
constexpr int S3_NBP = 3;                   // b planes in flight: s+1 .. s-1
constexpr int S3_KC = 16;                   // output planes per chunk
constexpr int S3_XPT = (S3_PL + S3_NT - 1) / S3_NT;   // x cells per thread per plane
constexpr int S3_BPT = (S3_BPL + S3_NT - 1) / S3_NT;  // b cells per thread per plane

struct S3Tiling {
    Box box;  // the interior
    int tx, ty, tk;
    int blocks;
};
inline S3Tiling make_s3_tiling(const Box &b) {
    S3Tiling t;
    t.box = b;
    t.tx = (b.n[0] + S3_BX - 1) / S3_BX;
    t.ty = (b.n[1] + S3_BY - 1) / S3_BY;
    t.tk = (b.n[2] + S3_KC - 1) / S3_KC;
    t.blocks = (b.n[0] > 0 && b.n[1] > 0 && b.n[2] > 0) ? t.tx * t.ty * t.tk : 0;
    return t;
}

template <class T>
__global__ __launch_bounds__(S3_NT) void k_sor3_km(S3Tiling kt, Geom g, OpSep<T> a, const T *__restrict__ xin,
                                                   T *__restrict__ xout, const T *__restrict__ b, T omega, int c0,
                                                   const int *done) {
    if (*done) return;
    __shared__ T sx[S3_NXP][S3_PL];
    __shared__ T sb[S3_NBP][S3_BPL], sd[S3_NBP][S3_BPL];
    __shared__ T saw[S3_W], sae[S3_W], sas[S3_H], san[S3_H];
    const int L = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x));
    const int bx = L % kt.tx, by = (L / kt.tx) % kt.ty, kc = L / (kt.tx * kt.ty);
    const int i0 = kt.box.lo[0] + bx * S3_BX, j0 = kt.box.lo[1] + by * S3_BY;
    const int ilo = kt.box.lo[0], ihi = kt.box.lo[0] + kt.box.n[0];  // interior [lo, hi)
    const int jlo = kt.box.lo[1], jhi = kt.box.lo[1] + kt.box.n[1];
    const int klo = kt.box.lo[2], khi = kt.box.lo[2] + kt.box.n[2];
    const int k0 = klo + kc * S3_KC, k1 = min(k0 + S3_KC, khi);  // output planes
    const int t = (int)threadIdx.x;
    const int sz = g.nxy;
    // 1-D operator coefficients of the tile's x and y range
    for (int q = t; q < S3_W + S3_H; q += S3_NT) {
        if (q < S3_W) {
            const int gi = min(max(i0 - S3_R + q, 0), g.nx - 1);
            saw[q] = a.aw[gi];
            sae[q] = a.ae[gi];
        } else {
            const int gj = min(max(j0 - S3_R + (q - S3_W), 0), g.ny - 1);
            sas[q - S3_W] = a.as[gj];
            san[q - S3_W] = a.an[gj];
        }
    }
    // this thread's load slots: x on ring 3, b and the diagonal on ring 2
    // (cells outside the array load nothing; no update reads them)
    unsigned xo[S3_XPT], bo[S3_BPT];
    bool xok[S3_XPT], bok[S3_BPT];
#pragma unroll
    for (int q = 0; q < S3_XPT; q++) {
        const int c = t + q * S3_NT;
        const int gi = i0 - S3_R + c % S3_W, gj = j0 - S3_R + c / S3_W;
        xok[q] = c < S3_PL && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        xo[q] = xok[q] ? (unsigned)(gi + g.nx * gj) * sizeof(T) : 0u;
    }
#pragma unroll
    for (int q = 0; q < S3_BPT; q++) {
        const int c = t + q * S3_NT;
        const int gi = i0 - S3_RB + c % S3_BW, gj = j0 - S3_RB + c / S3_BW;
        bok[q] = c < S3_BPL && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        bo[q] = bok[q] ? (unsigned)(gi + g.nx * gj) * sizeof(T) : 0u;
    }
    const unsigned szb = (unsigned)sz * sizeof(T);
    auto load_x = [&](int k, T (&v)[S3_XPT]) {
        const bool kin = k >= 0 && k < g.nz;
#pragma unroll
        for (int q = 0; q < S3_XPT; q++) v[q] = (kin && xok[q]) ? ld_at(xin, xo[q] + (unsigned)k * szb) : T(0);
    };
    auto load_b = [&](int k, T (&vb)[S3_BPT], T (&vd)[S3_BPT]) {
        const bool kin = k >= 0 && k < g.nz;
#pragma unroll
        for (int q = 0; q < S3_BPT; q++) {
            const bool ok = kin && bok[q];
            vb[q] = ok ? ld_at(b, bo[q] + (unsigned)k * szb) : T(0);
            vd[q] = ok ? ld_at(a.ac, bo[q] + (unsigned)k * szb) : T(1);
        }
    };
    auto store_x = [&](int k, const T (&v)[S3_XPT]) {
        T *p = sx[((k % S3_NXP) + S3_NXP) % S3_NXP];
#pragma unroll
        for (int q = 0; q < S3_XPT; q++)
            if (t + q * S3_NT < S3_PL) p[t + q * S3_NT] = v[q];
    };
    auto store_b = [&](int k, const T (&vb)[S3_BPT], const T (&vd)[S3_BPT]) {
        const int sl = ((k % S3_NBP) + S3_NBP) % S3_NBP;
#pragma unroll
        for (int q = 0; q < S3_BPT; q++)
            if (t + q * S3_NT < S3_BPL) {
                sb[sl][t + q * S3_NT] = vb[q];
                sd[sl][t + q * S3_NT] = vd[q];
            }
    };
    // one SOR update of x-buffer cell (xi, yi) of plane k (colour already
    // checked by the caller): k_sor_op's expression
    auto update = [&](int k, int xi, int yi, T atk, T abk) {
        T *xc_p = sx[k % S3_NXP];
        const T *xm = sx[(k + S3_NXP - 1) % S3_NXP], *xp = sx[(k + 1) % S3_NXP];
        const int c = xi + S3_W * yi;
        const int sl = k % S3_NBP, cbi = (xi - (S3_R - S3_RB)) + S3_BW * (yi - (S3_R - S3_RB));
        const T xc = xc_p[c];
        const T ac = sd[sl][cbi];
        const T s = ac * xc + sae[xi] * xc_p[c + 1] + saw[xi] * xc_p[c - 1] + san[yi] * xc_p[c + S3_W] +
                    sas[yi] * xc_p[c - S3_W] + atk * xp[c] + abk * xm[c];
        const T cv = (sb[sl][cbi] - s) / ac;
        xc_p[c] = xc + omega * cv;
    };
    // sweep with colour col on ring r of plane k, interior cells only.  A
    // ring row of even width holds w/2 cells of each colour: thread c takes
    // the c-th cell of the colour, so no lane idles on the other colour.
    auto sweep = [&](int k, int r, int col) {
        if (k < klo || k >= khi) return;  // uniform: ghost planes are not updated
        const int w = S3_BX + 2 * r, h = S3_BY + 2 * r, half = w / 2;
        const int gi0 = i0 - r;
        const T atk = a.at[k], abk = a.ab[k];  // before the loop: one wait for both
        for (int c = t; c < half * h; c += S3_NT) {
            const int yr = c / half, xr = 2 * (c - yr * half);
            const int gj = j0 - r + yr;
            const int gi = gi0 + xr + ((col - gi0 - gj - k - g.osum) & 1);
            if (gi < ilo || gi >= ihi || gj < jlo || gj >= jhi) continue;
            update(k, gi - (i0 - S3_R), gj - (j0 - S3_R), atk, abk);
        }
    };
    // Sweep 3 works per column: thread t owns column (i0 + t % BX, j0 + t / BX)
    // and keeps its final value of the previous plane, b and the diagonal
    // of its cell in registers, so neither plane s-3 nor b of plane s-2
    // occupies LDS.
    const int oi = i0 + t % S3_BX, oj = j0 + t / S3_BX;
    const bool own = oi < ihi && oj < jhi;
    const int oxi = oi - (i0 - S3_R), oyi = oj - (j0 - S3_R);
    const unsigned oo = own ? (unsigned)(oi + g.nx * oj) * sizeof(T) : 0u;
    T xprev = T(0), ob = T(0), od = T(1), nb_ = T(0), nd_ = T(1);
    auto load_own = [&](int k, T &vb_, T &vd_) {
        if (own && k >= klo && k < khi) {
            vb_ = ld_at(b, oo + (unsigned)k * szb);
            vd_ = ld_at(a.ac, oo + (unsigned)k * szb);
        }
    };
    // prologue: x planes k0-3 .. k0-1 and b plane k0-2 in LDS, x plane k0
    // and b plane k0-1 in registers
    T vx[S3_XPT], vb[S3_BPT], vd[S3_BPT];
    for (int k = k0 - 3; k < k0; k++) {
        load_x(k, vx);
        store_x(k, vx);
    }
    load_b(k0 - 2, vb, vd);
    store_b(k0 - 2, vb, vd);
    load_x(k0, vx);
    load_b(k0 - 1, vb, vd);
    __syncthreads();
    const int c1 = 1 - c0;
    // Iteration s: sweep 1 on plane s, sweep 2 on s-1, sweep 3 on s-2.  The
    // LDS slots written at the top of the iteration (x plane s+2, b plane
    // s+1) were last read during iteration s-1, before its final barrier.
    for (int s = k0 - 2; s < k1 + 2; s++) {
        if (s + 2 <= k1 + 2) store_x(s + 2, vx);
        if (s + 1 <= k1 + 1) store_b(s + 1, vb, vd);
        if (s + 3 <= k1 + 2) load_x(s + 3, vx);
        if (s + 2 <= k1 + 1) load_b(s + 2, vb, vd);
        ob = nb_;
        od = nd_;
        load_own(s - 1, nb_, nd_);  // the own cell of the next sweep-3 plane
        sweep(s, 2, c0);
        __syncthreads();
        if (s - 1 >= k0 - 1) sweep(s - 1, 1, c1);
        __syncthreads();
        const int k = s - 2;
        if (k == k0 - 1 && own) xprev = sx[k % S3_NXP][oxi + S3_W * oyi];  // final below the chunk
        if (k >= k0 && own) {
            // sweep 3 on the own column's cell: its k-1 neighbour is the
            // final value kept from the previous plane
            T *xc_p = sx[k % S3_NXP];
            const int c = oxi + S3_W * oyi;
            T v = xc_p[c];
            if (((oi + oj + k + g.osum) & 1) == c0) {
                const T sr = od * v + sae[oxi] * xc_p[c + 1] + saw[oxi] * xc_p[c - 1] + san[oyi] * xc_p[c + S3_W] +
                             sas[oyi] * xc_p[c - S3_W] + a.at[k] * sx[(k + 1) % S3_NXP][c] + a.ab[k] * xprev;
                const T cv = (ob - sr) / od;
                v = v + omega * cv;
            }
            st_at(xout, oo + (unsigned)k * szb, v);
            xprev = v;
        }
        __syncthreads();
    }
}

// q = A pp with (rr, q) (mode 1) or t = A ss with (t, s), (t, t) (mode 2)
template <class T, class OP, int MODE>
__global__ __launch_bounds__(K7_NT) void k_bicg_mv_km(K7Tiling kt, Geom g, OP a, const T *__restrict__ xin,
                                                      T *__restrict__ out, const T *__restrict__ w, const DevState *st,
                                                      double *partials) {
    if (st->done) return;
    __shared__ K7Lds<T> lx;
    __shared__ double lds[2 * (K7_NT / 64)];
    const K7Pos Q = k7_pos(kt, g);
    const int sz = g.nxy;
    Plane7<T> px;
    px.setup(xin, &lx, Q, g);
    px.prologue(Q, sz);
    double v[2] = {0.0, 0.0};
    int bf = 0;
    const auto ca = a.column(Q.act ? Q.i : kt.box.lo[0], Q.act ? Q.j : kt.box.lo[1]);
    const unsigned szt = (unsigned)sz * sizeof(T);
    unsigned ot = (unsigned)Q.col0 * sizeof(T) + (unsigned)Q.k0 * szt;
    auto cq = ca.load(ot, Q.k0);
    T wq = ld_at(w, ot);
    __syncthreads();
    for (int k = Q.k0; k < Q.k1; k++, ot += szt) {
        const bool more = k + 1 < Q.k1;
        const auto cc = cq;
        const T wv = wq;
        if (more) {
            px.prefetch(Q, k, sz);
            cq = ca.load(ot + szt, k + 1);
            wq = ld_at(w, ot + szt);
        }
        if (Q.act) {
            const T qv = ca.apply(cc, px.c(), px.xp(bf), px.xm(bf), px.yp(bf), px.ym(bf), px.zp(), px.zm());
            st_at(out, ot, qv);
            if (MODE == 1) {
                v[0] += (double)wv * (double)qv;
            } else {
                v[0] += (double)qv * (double)wv;
                v[1] += (double)qv * (double)qv;
            }
        }
        px.advance(bf, more);
        bf ^= 1;
        __syncthreads();
    }
    if (MODE == 1) {
        double v1[1] = {v[0]};
        block_partials<K7_NT, 1>(v1, partials, lds);
    } else {
        block_partials<K7_NT, 2>(v, partials, lds);
    }
}

__global__ __launch_bounds__(FIN_NT) void k_fin_init(DevState *st, const double *partials, int nb) {
    if (st->final || !st->done) return;
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if (threadIdx.x == 0) {
        const double err = sqrt(tot[0]) / st->gN;
        st->err = err;
        st->rho = tot[1];
        st->pending = 0;  // the init kernel has read x with the pass's sum added
        if (st->mixed && ((st->passes > 0 && err <= st->tol) || st->total >= st->maxit)) {
            st->final = 1;
            return;
        }
        st->rrho = 1.0;
        st->alpha = 0.0;
        st->omega = 1.0;
        st->beta = 0.0;
        st->it_pass = 0;
        st->first = 1;
        if (fabs(st->rho) < (double)FLT_MIN) {  // src/FalmEq.cpp:345-351
            if (st->mixed) {
                // the float residual vanished: no float pass can improve x
                st->status = err <= st->tol ? FALM_OK : FALM_ERR_BREAKDOWN;
            } else {
                st->err = st->rho;
                st->status = FALM_ERR_BREAKDOWN;
            }
            st->final = 1;
            return;
        }
        if (st->mixed) {
            const double target = st->inner_rtol * err;
            const double floor = FALM_MIXED_TOL_MARGIN * st->tol;
            st->tol_inner = target > floor ? target : floor;
            const int left = st->maxit - st->total;
            st->maxit_pass = left < st->inner_maxit ? left : st->inner_maxit;
        }
        st->passes += 1;
        st->done = 0;
    }
}

// The mixed solve's test after a pass: the true residual's norm ends the
// solve (tolerance met, or maxit spent); otherwise the pass is over and the
// host starts the next one with the full init kernels.  It ends every round
// of the mixed solve, so it also copies the state into the host's pinned
// mirror (hm): the host's poll then needs no copy of its own.
__global__ __launch_bounds__(FIN_NT) void k_fin_check(DevState *st, const double *partials, int nb,
                                                      DevState *hm) {
    if (!(st->final || !st->done)) {
        double tot[2];
        sum_partials<2>(partials, nb, tot);
        if (threadIdx.x == 0) {
            const double err = sqrt(tot[0]) / st->gN;
            st->err = err;
            st->pending = 0;
            if ((st->passes > 0 && err <= st->tol) || st->total >= st->maxit) st->final = 1;
        }
    }
    if (threadIdx.x == 0) *hm = *st;
}

// p = r (first) or r + beta(p - omega q) (kernel_PBiCGStab_1), then the
// first preconditioner sweep pp = M0(p)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_bicg_p(Tiling tl, Geom g, OP a, const T *__restrict__ r, T *__restrict__ p,
                                               const T *__restrict__ q, T *__restrict__ pp, const DevState *st,
                                               int pc_type, T omega_pc) {
    if (st->done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        T pv;
        if (st->first) {
            pv = r[cc];
        } else {
            const T beta = (T)st->beta, omega = (T)st->omega;
            pv = r[cc] + beta * (p[cc] - omega * q[cc]);
        }
        p[cc] = pv;
        pp[cc] = pc_first<T>(pc_type, omega_pc, pv, a.diag(cc), ((i + j + k + g.osum) & 1) == FALM_COLOR_BLACK);
    });
}

// s = r - alpha q (kernel_PBiCGStab_2), then ss = M0(s)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_bicg_s(Tiling tl, Geom g, OP a, const T *__restrict__ r,
                                               const T *__restrict__ q, T *__restrict__ s, T *__restrict__ ss,
                                               const DevState *st, int pc_type, T omega_pc) {
    if (st->done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const T alpha = (T)st->alpha;
        const T sv = r[cc] - alpha * q[cc];
        s[cc] = sv;
        ss[cc] = pc_first<T>(pc_type, omega_pc, sv, a.diag(cc), ((i + j + k + g.osum) & 1) == FALM_COLOR_BLACK);
    });
}

// red-black SOR half sweep on the cells of `color` (kernel_Struct3d7p_SOR,
// src/dev/FalmEqDevCall.cu:175-211)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_sor_op(Tiling tl, Geom g, OP a, T *x, const T *__restrict__ b, T omega,
                                               int color, const int *done) {
    if (*done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        if (((i + j + k + g.osum) & 1) != color) return;
        const int cc = g.id(i, j, k);
        const T xc = x[cc];
        const T s = apply_op<OP, T, T, T>(a, x, cc, i, j, k, g.nx, g.nxy);
        const T cv = (b[cc] - s) / a.diag(cc);
        x[cc] = xc + omega * cv;
    });
}

// kernel_Struct3d7p_Jacobi (src/dev/FalmEqDevCall.cu:101-132)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_jacobi_op(Tiling tl, Geom g, OP a, T *__restrict__ x,
                                                  const T *__restrict__ xp, const T *__restrict__ b,
                                                  const int *done) {
    if (*done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const T xc = xp[cc];
        x[cc] = xc + (b[cc] - apply_op<OP, T, T, T>(a, xp, cc, i, j, k, g.nx, g.nxy)) / a.diag(cc);
    });
}

template <class T>
__global__ __launch_bounds__(NT) void k_copy_interior(Tiling tl, Geom g, const T *__restrict__ src,
                                                      T *__restrict__ dst, const int *done) {
    if (*done) return;
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        dst[cc] = src[cc];
    });
}

// q = A pp ; partial (rr, q)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_bicg_mvdot(Tiling tl, Geom g, OP a, const T *__restrict__ pp,
                                                   T *__restrict__ q, const T *__restrict__ rr, const DevState *st,
                                                   double *partials) {
    if (st->done) return;
    __shared__ double lds[NT / 64];
    int i, j, k, t0, t1;
    double v[1] = {0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        const T qv = apply_op<OP, T, T, T>(a, pp, cc, i, j, k, g.nx, g.nxy);
        q[cc] = qv;
        v[0] += (double)rr[cc] * (double)qv;
    }
    block_partials<NT, 1>(v, partials, lds);
}

__global__ __launch_bounds__(FIN_NT) void k_fin_alpha(DevState *st, const double *partials, int nb) {
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) st->alpha = st->rho / tot[0];
}

// t = A ss ; partials (t, s), (t, t)
template <class T, class OP>
__global__ __launch_bounds__(NT) void k_bicg_mvdot2(Tiling tl, Geom g, OP a, const T *__restrict__ ss,
                                                    T *__restrict__ t, const T *__restrict__ s, const DevState *st,
                                                    double *partials) {
    if (st->done) return;
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k, t0, t1;
    double v[2] = {0.0, 0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        const T tv = apply_op<OP, T, T, T>(a, ss, cc, i, j, k, g.nx, g.nxy);
        t[cc] = tv;
        v[0] += (double)tv * (double)s[cc];
        v[1] += (double)tv * (double)tv;
    }
    block_partials<NT, 2>(v, partials, lds);
}

__global__ __launch_bounds__(FIN_NT) void k_fin_omega(DevState *st, const double *partials, int nb) {
    if (st->done) return;
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if (threadIdx.x == 0) st->omega = tot[0] / tot[1];
}

// x += alpha pp + omega ss (kernel_PBiCGStab_3, x in double);
// r = s - omega t (kernel_PBiCGStab_4); partials |r|^2, (r, rr).
template <class T>
__global__ __launch_bounds__(NT) void k_bicg_update(Tiling tl, Geom g, double *__restrict__ x,
                                                    const T *__restrict__ pp, const T *__restrict__ ss,
                                                    const T *__restrict__ s, const T *__restrict__ t,
                                                    T *__restrict__ r, const T *__restrict__ rr, const DevState *st,
                                                    double *partials) {
    if (st->done) return;
    __shared__ double lds[2 * (NT / 64)];
    int i, j, k, t0, t1;
    double v[2] = {0.0, 0.0};
    const double alpha = st->alpha, omega = st->omega;
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        x[cc] += alpha * (double)pp[cc] + omega * (double)ss[cc];
        const T rv = s[cc] - (T)omega * t[cc];
        r[cc] = rv;
        v[0] += (double)rv * (double)rv;
        v[1] += (double)rv * (double)rr[cc];
    }
    block_partials<NT, 2>(v, partials, lds);
}

// k_bicg_update on pairs of x-adjacent cells: 8-byte (float) or 16-byte
// (double) accesses, which stream at ~5.2 TB/s where one float per lane
// reaches ~4.0 (measured on MI355X with the same bytes).  The tiling runs over
// cell pairs (tp.box.n[0] = pairs per row, tp.box.lo[0] = 0); the pair at
// index ip covers cells lo0 + 2 ip and lo0 + 2 ip + 1.  Needs an even number
// of interior cells per row, even cell offsets at every pair (lo0 and nx
// even) and 16-byte aligned x; the sums run per thread over its pairs.
// workgroups of the pair update (each sums the second operator product's
// partials, so fewer and longer-running ones)
constexpr int UPD_BLOCKS = 1024;

template <class T>
__global__ __launch_bounds__(NT) void k_bicg_update2(Tiling tp, int lo0, Geom g, double *__restrict__ x,
                                                     const T *__restrict__ pp, const T *__restrict__ ss,
                                                     const T *__restrict__ s, const T *__restrict__ t,
                                                     T *__restrict__ r, const T *__restrict__ rr,
                                                     const DevState *st, double *partials, const double *opart,
                                                     int onb, DevState *stw) {
    using T2 = PairT<T>;
    if (st->done) return;
    __shared__ double lds[2 * (NT / 64)];
    int t0, t1;
    double v[2] = {0.0, 0.0};
    const double alpha = st->alpha;
    // a pair's operands, loaded one tile ahead of their use
    struct Pr {
        T2 p2, s2, sv, tv, r0;
        double2 xv;
        int cc;
        bool ok;
    };
    const auto load = [&](int tile) -> Pr {
        Pr q;
        int ip, j, k;
        q.ok = tile_cell_at(tp, tile, ip, j, k);
        if (q.ok) {
            q.cc = g.id(lo0 + 2 * ip, j, k);
            q.p2 = *(const T2 *)(pp + q.cc);
            q.s2 = *(const T2 *)(ss + q.cc);
            q.sv = *(const T2 *)(s + q.cc);
            q.tv = *(const T2 *)(t + q.cc);
            q.r0 = *(const T2 *)(rr + q.cc);
            q.xv = *(const double2 *)(x + q.cc);
        }
        return q;
    };
    tile_range(tp.tiles, t0, t1);
    // the first tile's loads are in flight while omega is summed
    Pr cur = load(t0 < t1 ? t0 : 0);
    cur.ok = cur.ok && t0 < t1;
    double omega;
    if (opart) {
        // omega = (t, s) / (t, t) from the second operator kernel's partial
        // sums (what k_fin_omega computes), in every workgroup; the first
        // stores it
        double tot[2];
        sum_partials<2, NT>(opart, onb, tot);
        omega = tot[0] / tot[1];
        if (blockIdx.x == 0 && threadIdx.x == 0) stw->omega = omega;
    } else {
        omega = st->omega;
    }
    const T om = (T)omega;
    for (int tile = t0; tile < t1; tile++) {
        Pr nxt;
        nxt.ok = false;
        if (tile + 1 < t1) nxt = load(tile + 1);
        if (cur.ok) {
            double2 xv = cur.xv;
            xv.x += alpha * (double)cur.p2.x + omega * (double)cur.s2.x;
            xv.y += alpha * (double)cur.p2.y + omega * (double)cur.s2.y;
            *(double2 *)(x + cur.cc) = xv;
            T2 rv;
            rv.x = cur.sv.x - om * cur.tv.x;
            rv.y = cur.sv.y - om * cur.tv.y;
            *(T2 *)(r + cur.cc) = rv;
            v[0] += (double)rv.x * (double)rv.x;
            v[1] += (double)rv.x * (double)cur.r0.x;
            v[0] += (double)rv.y * (double)rv.y;
            v[1] += (double)rv.y * (double)cur.r0.y;
        }
        cur = nxt;
    }
    block_partials<NT, 2>(v, partials, lds);
}

// err = |r|/N, it++, stop test, next rho and beta (src/FalmEq.cpp:345-407)
__global__ __launch_bounds__(FIN_NT) void k_fin_update(DevState *st, const double *partials, int nb) {
    if (st->done) return;
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if (threadIdx.x == 0) {
        st->rrho = st->rho;
        st->err = sqrt(tot[0]) / st->gN;
        st->it_pass += 1;
        st->total += 1;
        st->first = 0;
        st->pending = 1;
        if (!(st->it_pass < st->maxit_pass && st->err > st->tol_inner)) {
            st->done = 1;
        } else {
            st->rho = tot[1];
            if (fabs(st->rho) < (double)FLT_MIN) {
                // mixed: the pass ends and the outer loop decides
                if (!st->mixed) {
                    st->err = st->rho;
                    st->status = FALM_ERR_BREAKDOWN;
                }
                st->done = 1;
            } else {
                st->beta = (st->rho * st->alpha) / (st->rrho * st->omega);
            }
        }
    }
}

// k_fin_update for the folded update: the s kernel's partial sums (t, s),
// (t, t), (s, s), (s, rr), (t, rr) give omega and, for r = s - omega t,
// |r|^2 = (s, s) - omega (t, s) (omega minimises |s - omega t|) and
// (r, rr) = (s, rr) - omega (t, rr), so the stop test and beta come before r
// is formed (the next p kernel forms it).
__global__ __launch_bounds__(FIN_NT) void k_fin_fold(DevState *st, const double *partials, int nb) {
    if (st->done) return;
    double tot[5];
    sum_partials<5>(partials, nb, tot);
    if (threadIdx.x == 0) {
        const double omega = tot[0] / tot[1];
        const double r2 = tot[2] - omega * tot[0];
        st->omega = omega;
        st->rrho = st->rho;
        st->err = sqrt(r2 > 0.0 ? r2 : 0.0) / st->gN;
        st->it_pass += 1;
        st->total += 1;
        st->first = 0;
        st->pending = 1;
        if (!(st->it_pass < st->maxit_pass && st->err > st->tol_inner)) {
            st->done = 1;
        } else {
            st->rho = tot[3] - omega * tot[4];
            if (fabs(st->rho) < (double)FLT_MIN) {
                if (!st->mixed) {
                    st->err = st->rho;
                    st->status = FALM_ERR_BREAKDOWN;
                }
                st->done = 1;
            } else {
                st->beta = (st->rho * st->alpha) / (st->rrho * omega);
            }
        }
    }
}

// plain SOR / Jacobi solvers: r = b - A x and partial |r|^2
template <class OP>
__global__ __launch_bounds__(NT) void k_res_norm(Tiling tl, Geom g, OP a, const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ r,
                                                 const DevState *st, double *partials) {
    if (st->done) return;
    __shared__ double lds[NT / 64];
    int i, j, k, t0, t1;
    double v[1] = {0.0};
    tile_range(tl.tiles, t0, t1);
    for (int tile = t0; tile < t1; tile++) {
        if (!tile_cell_at(tl, tile, i, j, k)) continue;
        const int cc = g.id(i, j, k);
        const double rv = b[cc] - apply_op<OP, double, double, double>(a, x, cc, i, j, k, g.nx, g.nxy);
        r[cc] = rv;
        v[0] += rv * rv;
    }
    block_partials<NT, 1>(v, partials, lds);
}

__global__ __launch_bounds__(FIN_NT) void k_fin_resnorm(DevState *st, const double *partials, int nb) {
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) {
        st->err = sqrt(tot[0]) / st->gN;
        st->it_pass += 1;
        if (!(st->it_pass < st->maxit_pass && st->err > st->tol_inner)) st->done = 1;
    }
}

template <class TS, class TD>
__global__ void k_convert(const TS *__restrict__ s, TD *__restrict__ d, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        d[t] = (TD)s[t];
}

// A new solve.  Krylov solves start with no pass running (done = 1) and the
// init kernels armed (final = 0); the stationary solvers start iterating.
__global__ void k_state_reset(DevState *st, double gN, int mixed, double tol, int maxit, double inner_rtol,
                              int inner_maxit, int done0) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->gN = gN;
        st->mixed = mixed;
        st->done = done0;
        st->final = 0;
        st->status = FALM_OK;
        st->it_pass = 0;
        st->total = 0;
        st->passes = 0;
        st->err = 0.0;
        st->tol = tol;
        st->tol_inner = tol;
        st->maxit = maxit;
        st->maxit_pass = maxit;
        st->inner_rtol = inner_rtol;
        st->inner_maxit = inner_maxit;
        st->pending = 0;
    }
}

// --------------------------------------------- separable operator check
// 1-D candidates from one interior row per axis
template <class TA>
__global__ void k_sep_extract(Geom g, Mat7<TA> a, TA *aw, TA *ae, TA *as, TA *an, TA *ab, TA *at) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int gc = g.gc;
    if (t < g.nx) {
        const int cc = g.id(t, gc, gc);
        aw[t] = a.c[1][cc];
        ae[t] = a.c[2][cc];
    }
    if (t < g.ny) {
        const int cc = g.id(gc, t, gc);
        as[t] = a.c[3][cc];
        an[t] = a.c[4][cc];
    }
    if (t < g.nz) {
        const int cc = g.id(gc, gc, t);
        ab[t] = a.c[5][cc];
        at[t] = a.c[6][cc];
    }
}

__device__ inline bool same_bits(double x, double y) { return __double_as_longlong(x) == __double_as_longlong(y); }
__device__ inline bool same_bits(float x, float y) { return __float_as_int(x) == __float_as_int(y); }

// every interior cell's off-diagonals equal the 1-D candidates bit for bit
template <class TA>
__global__ __launch_bounds__(NT) void k_sep_check(Tiling tl, Geom g, Mat7<TA> a, const TA *aw, const TA *ae,
                                                  const TA *as, const TA *an, const TA *ab, const TA *at,
                                                  int *mismatch) {
    for_each_cell(tl, [&](int i, int j, int k) {
        const int cc = g.id(i, j, k);
        const bool ok = same_bits(a.c[1][cc], aw[i]) && same_bits(a.c[2][cc], ae[i]) && same_bits(a.c[3][cc], as[j]) &&
                        same_bits(a.c[4][cc], an[j]) && same_bits(a.c[5][cc], ab[k]) && same_bits(a.c[6][cc], at[k]);
        if (!ok) *mismatch = 1;  // benign race: every writer stores 1
    });
}

// every interior cell's diagonal equals the one its runs' classes give, bit
// for bit (through the same column view the k-marching kernels use)
__global__ __launch_bounds__(NT) void k_diag_check(Tiling tl, Geom g, OpSep<double> o, int *mismatch) {
    for_each_cell(tl, [&](int i, int j, int k) {
        if (!same_bits(o.ac[g.id(i, j, k)], o.column(i, j).pick(k))) *mismatch = 1;  // benign race
    });
}

}  // namespace falm

using namespace falm;

struct falm_solver {
    falm_domain dom;
    falm_solver_config cfg;
    size_t n;
    Geom g;
    Tiling tl;
    Tiling tp;     // tiles of x-adjacent cell pairs (k_bicg_update2); no tiles when pairs are not aligned
    K7Tiling kt;   // k-marching column tiles of the interior (stencil kernels)
    K7Tiling ktp;  // the same for k_bicg_pcmv_km (its own chunk length)
    S3Tiling s3;   // tiles of the fused three-sweep SOR kernel
    DevState *st = nullptr;
    DevState *hst = nullptr;  // pinned mirror
    double *partials = nullptr;
    double *partials2 = nullptr;  // the second operator product's partial sums (read by the update kernel)
    int *flag = nullptr, *hflag = nullptr;
    // Krylov vectors: double (F64) or float (MIXED)
    void *rr = nullptr, *p = nullptr, *q = nullptr, *s = nullptr, *pp = nullptr, *ss = nullptr, *t = nullptr,
         *w = nullptr, *r32 = nullptr;
    // operator copies set by falm_solver_set_matrix
    const void *a_src = nullptr;  // the caller's matrix the copies were made from
    bool sep = false;
    double *sep64 = nullptr;      // 1-D arrays: aw,ae [nx], as,an [ny], ab,at [nz]
    float *sep32 = nullptr;
    // diagonal classes (OpSep): per-axis run starts and run classes, and the
    // class-triple values in double and float
    bool dclass = false;
    int drs[3][DIAG_CLS - 1] = {};
    int drc[3] = {};
    double *dtab64 = nullptr;
    float *dtab32 = nullptr;
    float *a32 = nullptr;         // mixed: float columns (7 if not separable, else diagonal only)
    double *r64own = nullptr;     // residual when the caller passes r = NULL (double solvers)
    int last_it = 0;              // iterations of the previous solve
    int pflip = 0;                // fused SOR path: p lives in w (1) or p (0)
    int qflip = 0;                // deferred path: q lives in pp (1) or q (0)
    void *trash = nullptr;        // k_bicg_pcmv_km's stores of threads with nothing to store
    void *stp = nullptr;  // the fused path's (s, t) pairs (2 n elements)
    // a solve split in two calls (begin / resume): the polling loop's state
    // between them
    int cur_batch = 0, cur_enqueued = 0, cur_polls = 0;
    bool cur_pass_start = false;
    bool first_ok = false;  // the last solve was over after its first batch
};

namespace falm {

static double interior_count(const falm_solver *S) {
    return (double)(S->dom.gshape[0] - 2 * S->dom.gc) * (double)(S->dom.gshape[1] - 2 * S->dom.gc) *
           (double)(S->dom.gshape[2] - 2 * S->dom.gc);
}

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
    const size_t ns = 2 * ((size_t)S->g.nx + S->g.ny + S->g.nz);
    FALM_HIP(hipMalloc(&S->sep64, ns * sizeof(double)));
    FALM_HIP(hipMalloc(&S->sep32, ns * sizeof(float)));
    FALM_HIP(hipMalloc(&S->dtab64, DIAG_CLS * DIAG_CLS * DIAG_CLS * sizeof(double)));
    FALM_HIP(hipMalloc(&S->dtab32, DIAG_CLS * DIAG_CLS * DIAG_CLS * sizeof(float)));
    FALM_HIP(hipMalloc(&S->st, sizeof(DevState)));
    FALM_HIP(hipMemset(S->st, 0, sizeof(DevState)));
    // coherent: the finalize kernels write it directly (the host reads it
    // after a stream synchronisation)
    FALM_HIP(hipHostMalloc(&S->hst, sizeof(DevState), hipHostMallocCoherent));
    FALM_HIP(hipMalloc(&S->flag, sizeof(int)));
    FALM_HIP(hipHostMalloc(&S->hflag, sizeof(int), hipHostMallocDefault));
    size_t nb = (size_t)(S->tl.tiles > 0 ? red_blocks(S->tl) : 1);
    if ((size_t)S->kt.blocks > nb) nb = (size_t)S->kt.blocks;
    if ((size_t)S->ktp.blocks > nb) nb = (size_t)S->ktp.blocks;
    FALM_HIP(hipMalloc(&S->partials, 2 * nb * sizeof(double)));
    FALM_HIP(hipMalloc(&S->trash, (size_t)(S->ktp.blocks > 0 ? S->ktp.blocks : 1) * K7_NT * 2 * sizeof(double)));
    if (S->cfg.pc_type == FALM_LS_SOR && S->cfg.pc_maxit <= 1) {  // the fused path (deferred_pc)
        FALM_HIP(hipMalloc(&S->stp, 2 * vb));
        FALM_HIP(hipMemset(S->stp, 0, 2 * vb));
    }
    FALM_HIP(hipMalloc(&S->partials2, 5 * nb * sizeof(double)));  // 5 sums: the folded s kernel
    // the zero fills above run on the null stream, which the caller's
    // (possibly non-blocking) stream does not wait for
    FALM_HIP(hipDeviceSynchronize());
    return FALM_OK;
}

// mirrored: the last state-changing kernel enqueued wrote the host mirror
static falm_status read_state(falm_solver *S, hipStream_t s, bool mirrored = false) {
    if (!mirrored) FALM_HIP(hipMemcpyAsync(S->hst, S->st, sizeof(DevState), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

static int check_period(const falm_solver *S) {
    if (S->cfg.check_every > 0) return S->cfg.check_every;
    double n = (double)S->tl.box.size();
    int c = (int)(4.0e6 / (n > 1 ? n : 1));
    return c < 1 ? 1 : (c > 32 ? 32 : c);
}

// views of the operator the solver uses
template <class TA>
static OpSep<TA> sep_view(const falm_solver *S, const TA *diag, const TA *arr) {
    OpSep<TA> o;
    const int nx = S->g.nx, ny = S->g.ny, nz = S->g.nz;
    o.ac = diag;
    o.aw = arr;
    o.ae = arr + nx;
    o.as = arr + 2 * nx;
    o.an = arr + 2 * nx + ny;
    o.ab = arr + 2 * nx + 2 * ny;
    o.at = arr + 2 * nx + 2 * ny + nz;
    if (S->dclass) {
        if constexpr (std::is_same<TA, double>::value) o.dtab = S->dtab64;
        else o.dtab = S->dtab32;
        o.x1 = S->drs[0][0], o.x2 = S->drs[0][1], o.x3 = S->drs[0][2];
        o.y1 = S->drs[1][0], o.y2 = S->drs[1][1], o.y3 = S->drs[1][2];
        o.z1 = S->drs[2][0], o.z2 = S->drs[2][1], o.z3 = S->drs[2][2];
        o.cxr = S->drc[0], o.cyr = S->drc[1], o.czr = S->drc[2];
    }
    return o;
}
template <class TA>
static OpCols<TA> cols_view(const falm_solver *S, const void *a) {
    OpCols<TA> o;
    for (int q = 0; q < 7; q++) o.c[q] = col<TA>(a, S->n, q);
    return o;
}

#define FALM_GRID(S) tiling_grid((S)->tl), tiling_block(), 0
#define FALM_RED(S) red_grid((S)->tl), tiling_block(), 0
#define FALM_KM(S) dim3((S)->kt.blocks), dim3(K7_NT), 0
#define FALM_FIN dim3(1), dim3(FIN_NT), 0

// bytes per cell of the operator's diagonal that a k-marching kernel
// streams: none when a separable operator takes it from its classes
template <class OP>
static double diag_bytes(const falm_solver *S) {
    using TA = decltype(std::declval<OP>().diag_at(0u));
    if constexpr (std::is_same<OP, OpSep<TA>>::value) return S->dclass ? 0.0 : (double)sizeof(TA);
    return (double)sizeof(TA);
}

// preconditioner sweeps after the fused first one (SORPC: pc_maxit x
// {black, red}; JacobiPC: pc_maxit sweeps; both run at least once,
// src/FalmEq.cpp:147-170,258-302).  Returns the vector holding the result:
// x itself, or `spare` when the three sweeps of pc_maxit = 2 on a separable
// double operator run as one out-of-place pass (k_sor3_km).
template <class T, class OP>
static T *pc_rest(falm_solver *S, const OP &a, T *x, const T *b, T *spare, hipStream_t s, falm_status &st) {
    const falm_solver_config &c = S->cfg;
    const int m = c.pc_maxit < 1 ? 1 : c.pc_maxit;
    const int *done = &S->st->done;
    const T w = (T)c.pc_relax;
    st = FALM_OK;
    if (c.pc_type == FALM_LS_SOR) {
        // the fused pass pays off in double (measured on MI355X at
        // 512x128x128: 127 us against 3 x 58 us); in float the three
        // separate half-sweeps (3 x 27 us) stay faster than its 80-95 us
        if constexpr (std::is_same<T, double>::value && std::is_same<OP, OpSep<T>>::value) {
            if (m == 2 && S->s3.blocks > 0) {
                TimedLaunch tm(FALM_TIME_SOR, s);
                hipLaunchKernelGGL((k_sor3_km<T>), dim3(S->s3.blocks), dim3(S3_NT), 0, s, S->s3, S->g, a,
                                   (const T *)x, spare, b, w, FALM_COLOR_RED, done);
                tm.done((double)S->tl.box.size() * 4.0 * sizeof(T));
                if (hipGetLastError() != hipSuccess) st = FALM_ERR_HIP;
                return spare;
            }
        }
        for (int it = 0; it < m; it++) {
            if (it > 0) {
                TimedLaunch tm(FALM_TIME_SOR, s);
                hipLaunchKernelGGL((k_sor_op_km<T, OP>), FALM_KM(S), s, S->kt, S->g, a, x, b, w, FALM_COLOR_BLACK,
                                   done);
                tm.done((double)S->tl.box.size() * 4.0 * sizeof(T));
            }
            TimedLaunch tm(FALM_TIME_SOR, s);
            hipLaunchKernelGGL((k_sor_op_km<T, OP>), FALM_KM(S), s, S->kt, S->g, a, x, b, w, FALM_COLOR_RED, done);
            tm.done((double)S->tl.box.size() * 4.0 * sizeof(T));
        }
    } else if (c.pc_type == FALM_LS_JACOBI) {
        for (int it = 1; it < m; it++) {
            hipLaunchKernelGGL((k_copy_interior<T>), FALM_GRID(S), s, S->tl, S->g, (const T *)x, (T *)S->w, done);
            hipLaunchKernelGGL((k_jacobi_op<T, OP>), FALM_GRID(S), s, S->tl, S->g, a, x, (const T *)S->w, b, done);
        }
    }
    if (hipGetLastError() != hipSuccess) st = FALM_ERR_HIP;
    return x;
}

// With the one-pass SOR preconditioner (pc_maxit <= 1) an iteration is the
// two k_bicg_pcmv_km kernels and k_fin_fold: the preconditioning is deferred
// to the end of a pass (y accumulates alpha p + omega s, M is linear) and
// the update r = s - omega t, y += alpha p + omega s is folded into the next
// iteration's p kernel (|r| and (r, rr) from the s kernel's sums); the
// pass's last update goes into k_bicg_close_km.  Other preconditioners run
// the separate p, sweep, operator and update kernels.
static bool deferred_pc(const falm_solver *S) {
    return S->cfg.pc_type == FALM_LS_SOR && S->cfg.pc_maxit <= 1;
}

// x += M((0 or y) + alpha p + omega s) once a pass is over, and the last
// residual s - omega t into rout when given (no-op while the pass runs or
// once its update is in x)
template <class T, class OP>
static falm_status launch_close(falm_solver *S, const OP &a, double *x, T *rout, hipStream_t s, bool mirror) {
    TimedLaunch tm(FALM_TIME_PC_APPLY, s);
    using P = PairT<T>;
    hipLaunchKernelGGL((k_bicg_close_km<T, OP>), FALM_KM(S), s, S->kt, S->g, a, (const T *)S->ss, (const T *)S->p,
                       (const T *)S->w, (const P *)S->stp, rout, (const DevState *)S->st, (T)S->cfg.pc_relax, x,
                       mirror ? S->hst : (DevState *)nullptr);
    // y, p, (s, t) (and the diagonal) read, x read and written; with rout
    // r written
    tm.done((double)S->tl.box.size() *
            ((rout ? 5.0 : 4.0) * sizeof(T) + 2.0 * sizeof(double) + diag_bytes<OP>(S)));
    FALM_CHECK_LAUNCH("k_bicg_close_km");
    return FALM_OK;
}

// one BiCGStab iteration (src/FalmEq.cpp:345-407); first: the host knows it
// starts a pass
template <class T, class OP>
static falm_status bicg_iteration(falm_solver *S, const OP &a, double *x, T *r, hipStream_t s, bool first) {
    const falm_solver_config &c = S->cfg;
    const Tiling &tl = S->tl;
    const Geom &g = S->g;
    const int nb = S->kt.blocks;
    T *rr = (T *)S->rr, *p = (T *)S->p, *q = (T *)S->q, *sv = (T *)S->s, *pp = (T *)S->pp, *ss = (T *)S->ss,
      *t = (T *)S->t;
    const T wpc = (T)c.pc_relax;
    if (deferred_pc(S)) {
        // p, q alternate between two buffers (the p kernel reads both at the
        // ring cells of its tiles); a pass's first iteration reads neither,
        // so the alternation may run on across iterations a finished pass
        // skips (the p kernel records which buffer it wrote)
        using P = PairT<T>;
        const int pid = S->pflip ? 0 : 1;  // 1: p goes to w
        T *pin = (T *)(S->pflip ? S->w : S->p), *pout = (T *)(S->pflip ? S->p : S->w);
        T *qin = (T *)(S->qflip ? S->pp : S->q), *qout = (T *)(S->qflip ? S->q : S->pp);
        S->pflip ^= 1;
        S->qflip ^= 1;
        const double box = (double)tl.box.size(), db = diag_bytes<OP>(S);
        const int nbp = S->ktp.blocks;
        const dim3 gp((unsigned)nbp), bp(K7_NT);
        T *trash = (T *)S->trash;
        // a pass's first iteration reads its residual from rr (r~ = r at
        // the start of a pass; the init kernels of the mixed solve do not
        // write a float copy of r)
        const T *rin = first ? (const T *)rr : (const T *)r;
        // the diagonal from its classes (a compile-time choice in the kernel)
        bool dcls = false;
        if constexpr (std::is_same<OP, OpSep<T>>::value) dcls = a.dtab != nullptr;
        // the solve's first iteration starts a pass (a later pass's first
        // iteration is found on the device and runs the general kernel)
        const auto launch = [&](auto dc, auto fi) {
            constexpr bool D = decltype(dc)::value, F = decltype(fi)::value;
            {
                TimedLaunch tm(FALM_TIME_PCMV, s);
                hipLaunchKernelGGL((k_bicg_pcmv_km<T, OP, 0, D, F>), gp, bp, 0, s, S->ktp, g, a, rin,
                                   (const T *)pin, (const T *)qin, (const T *)rr, pout, qout, (const DevState *)S->st,
                                   wpc, (const double *)nullptr, 0, S->st, S->partials, trash, (const P *)S->stp,
                                   (P *)nullptr, r, ss, pid);
                // a pass's first iteration: rr (its residual too; and the
                // diagonal) read, p and q written; then (s, t), p, q, rr, y
                // read and p, q, r, y written
                tm.done(box * ((first ? 3.0 : 10.0) * sizeof(T) + db));
            }
            {
                TimedLaunch tm(FALM_TIME_PCMV_S, s);
                hipLaunchKernelGGL((k_bicg_pcmv_km<T, OP, 1, D, false>), gp, bp, 0, s, S->ktp, g, a, rin,
                                   (const T *)nullptr, (const T *)qout, (const T *)rr, (T *)nullptr, (T *)nullptr,
                                   (const DevState *)S->st, wpc, (const double *)S->partials, nbp, S->st,
                                   S->partials2, trash, (const P *)nullptr, (P *)S->stp, (T *)nullptr,
                                   (T *)nullptr, 0);
                // r, q, rr (and the diagonal) read, (s, t) written
                tm.done(box * (5.0 * sizeof(T) + db));
            }
        };
        if (dcls) {
            if (first) launch(std::true_type{}, std::true_type{});
            else launch(std::true_type{}, std::false_type{});
        } else {
            if (first) launch(std::false_type{}, std::true_type{});
            else launch(std::false_type{}, std::false_type{});
        }
        FALM_CHECK_LAUNCH("k_bicg_pcmv_km");
        hipLaunchKernelGGL(k_fin_fold, FALM_FIN, s, S->st, (const double *)S->partials2, nbp);
        FALM_CHECK_LAUNCH("k_fin_fold");
        return FALM_OK;
    }
    falm_status pst;
    hipLaunchKernelGGL((k_bicg_p<T, OP>), FALM_GRID(S), s, tl, g, a, (const T *)r, p, (const T *)q, pp,
                       (const DevState *)S->st, c.pc_type, wpc);
    FALM_CHECK_LAUNCH("k_bicg_p");
    // the preconditioned p lands in pp or, out of place, in w
    T *const pcp = pc_rest<T, OP>(S, a, pp, p, (T *)S->w, s, pst);
    FALM_TRY(pst);
    TimedLaunch tm1(FALM_TIME_MV, s);
    hipLaunchKernelGGL((k_bicg_mv_km<T, OP, 1>), FALM_KM(S), s, S->kt, g, a, (const T *)pcp, q, (const T *)rr,
                       (const DevState *)S->st, S->partials);
    tm1.done((double)tl.box.size() * (3.0 * sizeof(T) + diag_bytes<OP>(S)));
    hipLaunchKernelGGL(k_fin_alpha, FALM_FIN, s, S->st, (const double *)S->partials, nb);
    FALM_CHECK_LAUNCH("k_bicg_mvdot");
    hipLaunchKernelGGL((k_bicg_s<T, OP>), FALM_GRID(S), s, tl, g, a, (const T *)r, (const T *)q, sv, ss,
                       (const DevState *)S->st, c.pc_type, wpc);
    FALM_CHECK_LAUNCH("k_bicg_s");
    // the preconditioned s lands in ss or, out of place, in pp (free once
    // p's sweeps have read it)
    T *const pcs = pc_rest<T, OP>(S, a, ss, sv, pcp == pp ? ss : pp, s, pst);
    FALM_TRY(pst);
    TimedLaunch tm3(FALM_TIME_MV2, s);
    hipLaunchKernelGGL((k_bicg_mv_km<T, OP, 2>), FALM_KM(S), s, S->kt, g, a, (const T *)pcs, t, (const T *)sv,
                       (const DevState *)S->st, S->partials2);
    tm3.done((double)tl.box.size() * (3.0 * sizeof(T) + diag_bytes<OP>(S)));
    FALM_CHECK_LAUNCH("k_bicg_mvdot2");
    // the pair update computes omega from mv2's partials itself (its own go
    // to the other buffer); the single-cell update reads it from the state
    const bool pairs = S->tp.tiles > 0 && ((uintptr_t)x & 15) == 0;
    if (!pairs) hipLaunchKernelGGL(k_fin_omega, FALM_FIN, s, S->st, (const double *)S->partials2, nb);
    TimedLaunch tm2(FALM_TIME_SOLVER_UPDATE, s);
    int nbu = red_blocks(tl);
    if (pairs) {
        nbu = S->tp.tiles < UPD_BLOCKS ? S->tp.tiles : UPD_BLOCKS;
        hipLaunchKernelGGL((k_bicg_update2<T>), dim3((unsigned)nbu), tiling_block(), 0, s, S->tp, tl.box.lo[0],
                           g, x, (const T *)pcp, (const T *)pcs, (const T *)sv,
                           (const T *)t, r, (const T *)rr, (const DevState *)S->st, S->partials,
                           (const double *)S->partials2, nb, S->st);
    } else {
        hipLaunchKernelGGL((k_bicg_update<T>), FALM_RED(S), s, tl, g, x, (const T *)pcp, (const T *)pcs,
                           (const T *)sv, (const T *)t, r, (const T *)rr, (const DevState *)S->st, S->partials);
    }
    tm2.done((double)tl.box.size() * (2.0 * sizeof(double) + 6.0 * sizeof(T)));
    hipLaunchKernelGGL(k_fin_update, FALM_FIN, s, S->st, (const double *)S->partials, nbu);
    FALM_CHECK_LAUNCH("k_bicg_update");
    return FALM_OK;
}

// Iterations enqueued before the first poll: the previous solve's count
// (time steps in sequence need similar counts), else the polling period.
static int first_batch(const falm_solver *S) {
    const int k = S->last_it > 0 ? S->last_it : check_period(S);
    return k < 1 ? 1 : (k > 64 ? 64 : k);
}

template <class T, class OP64>
static falm_status launch_init(falm_solver *S, const OP64 &a, const double *x, const double *b, double *r64, T *rT,
                               hipStream_t s) {
    TimedLaunch tm(FALM_TIME_SOLVER_INIT, s);
    hipLaunchKernelGGL((k_bicg_init_km<T, OP64>), FALM_KM(S), s, S->kt, S->g, a, x, b, r64, rT, (T *)S->rr,
                       (const DevState *)S->st, S->partials);
    // x, b (and the diagonal) read; r (double, if kept), r and rr (T) written
    tm.done((double)S->tl.box.size() *
            (2.0 * sizeof(double) + diag_bytes<OP64>(S) + (r64 ? sizeof(double) : 0.0) +
             ((rT && (void *)rT != (void *)r64) ? sizeof(T) : 0.0) + sizeof(T)));
    hipLaunchKernelGGL(k_fin_init, FALM_FIN, s, S->st, (const double *)S->partials, S->kt.blocks);
    FALM_CHECK_LAUNCH("k_bicg_init");
    return FALM_OK;
}

// the mixed solve's test after a pass (k_bicg_init_km<CHECK>): x, b read,
// the residual written into r64 when given
template <class OP64>
static falm_status launch_check(falm_solver *S, const OP64 &a, const double *x, const double *b, double *r64,
                                hipStream_t s) {
    TimedLaunch tm(FALM_TIME_SOLVER_INIT, s);
    hipLaunchKernelGGL((k_bicg_init_km<float, OP64, true>), FALM_KM(S), s, S->kt, S->g, a, x, b, r64,
                       (float *)nullptr, (float *)nullptr, (const DevState *)S->st, S->partials);
    tm.done((double)S->tl.box.size() * ((r64 ? 3.0 : 2.0) * sizeof(double) + diag_bytes<OP64>(S)));
    hipLaunchKernelGGL(k_fin_check, FALM_FIN, s, S->st, (const double *)S->partials, S->kt.blocks, S->hst);
    FALM_CHECK_LAUNCH("k_bicg_init_km<check>");
    return FALM_OK;
}

// mode (SolveMode): the whole solve, or split in two calls around work the
// caller enqueues between them.  SOLVE_BEGIN enqueues the first batch of
// iterations and returns without waiting; SOLVE_RESUME polls, finishes the
// solve and sets *rerun when the first batch had not finished it (the flag
// solver_gate() names was still zero when the caller's work ran).
template <class OP64>
static falm_status solve_bicgstab_f64(falm_solver *S, const OP64 &a, double *x, const double *b, double *r,
                                      hipStream_t s, falm_solve_info *info, int mode) {
    const falm_solver_config &c = S->cfg;
    int &batch = S->cur_batch, &enqueued = S->cur_enqueued;
    auto enqueue_batch = [&]() -> falm_status {
        for (int k = 0; k < batch && enqueued < c.maxit; k++, enqueued++)
            FALM_TRY((bicg_iteration<double, OP64>(S, a, x, r, s, enqueued == 0)));
        // the pass's sum into x, if the pass is over (before the poll, so x
        // is complete when the solve returns)
        if (deferred_pc(S)) FALM_TRY((launch_close<double, OP64>(S, a, x, r == S->r64own ? nullptr : r, s, true)));
        return FALM_OK;
    };
    if (mode != SOLVE_RESUME) {
        hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, interior_count(S), 0, c.tol, c.maxit,
                           0.0, 0, 1);
        FALM_CHECK_LAUNCH("k_state_reset");
        // on the fused path the first iteration reads its residual from rr,
        // so r is written only when it is the caller's
        double *const r0 = (deferred_pc(S) && r == S->r64own) ? nullptr : r;
        FALM_TRY((launch_init<double, OP64>(S, a, x, b, r0, r0, s)));
        batch = first_batch(S);
        enqueued = 0;
        FALM_TRY(enqueue_batch());
        if (mode == SOLVE_BEGIN) return FALM_OK;
    }
    bool first_read = true;
    while (true) {
        // the closing kernel mirrors the state (fused path)
        FALM_TRY(read_state(S, s, deferred_pc(S)));
        if (first_read) S->first_ok = S->hst->done;
        first_read = false;
        if (S->hst->done || enqueued >= c.maxit) break;
        batch = check_period(S);
        FALM_TRY(enqueue_batch());
    }
    info->it = S->hst->it_pass;
    info->err = S->hst->err;
    info->outer_it = 1;
    info->status = S->hst->status;
    if (info->status == FALM_OK && !(S->hst->err <= c.tol)) info->status = FALM_ERR_NOT_CONVERGED;
    S->last_it = info->it;
    return FALM_OK;
}

// Mixed precision: float BiCGStab passes on the defect b - A x, x
// accumulated in double.  Every pass ends with the init kernels, which
// compute the true residual in double and either end the solve or start the
// next pass, all on the device; the host polls once per batch.
template <class OP64, class OP32>
static falm_status solve_bicgstab_mixed(falm_solver *S, const OP64 &a64, const OP32 &a32, double *x,
                                        const double *b, double *r, hipStream_t s, falm_solve_info *info, int mode) {
    const falm_solver_config &c = S->cfg;
    const int inner_maxit = c.inner_maxit > 0 ? c.inner_maxit : 50;
    const double inner_rtol = c.inner_rtol > 0 ? c.inner_rtol : 1e-4;
    float *r32 = (float *)S->r32;
    // every round either advances the running pass by at least one
    // iteration or ends/starts a pass; the device ends the solve at maxit
    // pass_start: the next iteration enqueued starts a pass (after an init
    // that the host launched for a finished pass)
    int &batch = S->cur_batch, &polls = S->cur_polls;
    bool &pass_start = S->cur_pass_start;
    auto enqueue_round = [&]() -> falm_status {
        for (int k = 0; k < batch; k++)
            FALM_TRY((bicg_iteration<float, OP32>(S, a32, x, r32, s, pass_start && k == 0)));
        pass_start = false;
        // (the float residual is not kept: the init kernels form the next
        // pass's from x)
        if (deferred_pc(S)) FALM_TRY((launch_close<float, OP32>(S, a32, x, (float *)nullptr, s, false)));
        // a finished pass is tested with the norm (and the residual, if the
        // caller wants it); a pass that must follow starts from the full
        // init, which the host launches (it recomputes the same residual):
        // every pass's first iteration is known to the host
        FALM_TRY((launch_check<OP64>(S, a64, x, b, r, s)));
        return FALM_OK;
    };
    if (mode != SOLVE_RESUME) {
        hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, interior_count(S), 1, c.tol, c.maxit,
                           inner_rtol, inner_maxit, 1);
        FALM_CHECK_LAUNCH("k_state_reset");
        FALM_TRY((launch_init<float, OP64>(S, a64, x, b, r, deferred_pc(S) ? (float *)nullptr : r32, s)));
        batch = first_batch(S);
        polls = 0;
        pass_start = true;
        FALM_TRY(enqueue_round());
        if (mode == SOLVE_BEGIN) return FALM_OK;
    }
    bool first_read = true;
    while (true) {
        FALM_TRY(read_state(S, s, true));  // k_fin_check mirrors the state
        if (first_read) S->first_ok = S->hst->final;
        first_read = false;
        polls++;
        if (S->hst->final) break;
        if (S->hst->done) {
            FALM_TRY((launch_init<float, OP64>(S, a64, x, b, r, deferred_pc(S) ? (float *)nullptr : r32, s)));
            pass_start = true;
        }
        batch = check_period(S);
        if (polls > 2 * c.maxit + 2) {
            set_error("falm_solver_solve: mixed solve made no progress");
            return FALM_ERR_BREAKDOWN;
        }
        FALM_TRY(enqueue_round());
    }
    const DevState &h = *S->hst;
    info->it = h.total;
    info->err = h.err;
    info->outer_it = h.passes;
    info->status = h.status != FALM_OK ? h.status : (h.err <= c.tol ? FALM_OK : FALM_ERR_NOT_CONVERGED);
    S->last_it = h.total;
    return FALM_OK;
}

// FalmEq::SOR (src/FalmEq.cpp:173-240) and FalmEq::Jacobi (:81-127)
template <class OP>
static falm_status solve_stationary(falm_solver *S, const OP &a, double *x, const double *b, double *r,
                                    hipStream_t s, falm_solve_info *info, int mode) {
    const falm_solver_config &c = S->cfg;
    const int period = check_period(S);
    const int *done = &S->st->done;
    const int nb = red_blocks(S->tl);
    int &enqueued = S->cur_enqueued;
    auto enqueue_batch = [&]() -> falm_status {
        for (int k = 0; k < period && enqueued < c.maxit; k++, enqueued++) {
            if (c.type == FALM_LS_SOR) {
                hipLaunchKernelGGL((k_sor_op<double, OP>), FALM_GRID(S), s, S->tl, S->g, a, x, b, c.relax,
                                   FALM_COLOR_BLACK, done);
                hipLaunchKernelGGL((k_sor_op<double, OP>), FALM_GRID(S), s, S->tl, S->g, a, x, b, c.relax,
                                   FALM_COLOR_RED, done);
            } else {
                hipLaunchKernelGGL((k_copy_interior<double>), FALM_GRID(S), s, S->tl, S->g, (const double *)x,
                                   (double *)S->w, done);
                hipLaunchKernelGGL((k_jacobi_op<double, OP>), FALM_GRID(S), s, S->tl, S->g, a, x,
                                   (const double *)S->w, b, done);
            }
            hipLaunchKernelGGL((k_res_norm<OP>), FALM_RED(S), s, S->tl, S->g, a, (const double *)x, b, r,
                               (const DevState *)S->st, S->partials);
            hipLaunchKernelGGL(k_fin_resnorm, FALM_FIN, s, S->st, (const double *)S->partials, nb);
            FALM_CHECK_LAUNCH("k_res_norm");
        }
        return FALM_OK;
    };
    if (mode != SOLVE_RESUME) {
        hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, S->st, interior_count(S), 0, c.tol, c.maxit,
                           0.0, 0, 0);
        FALM_CHECK_LAUNCH("k_state_reset");
        enqueued = 0;
        FALM_TRY(enqueue_batch());
        if (mode == SOLVE_BEGIN) return FALM_OK;
    }
    bool first_read = true;
    while (true) {
        FALM_TRY(read_state(S, s));
        if (first_read) S->first_ok = S->hst->done;
        first_read = false;
        if (S->hst->done || enqueued >= c.maxit) break;
        FALM_TRY(enqueue_batch());
    }
    info->it = S->hst->it_pass;
    info->err = S->hst->err;
    info->outer_it = 1;
    info->status = S->hst->err <= c.tol ? FALM_OK : FALM_ERR_NOT_CONVERGED;
    return FALM_OK;
}

// Diagonal classes of a separable operator (OpSep): each interior index of
// each axis is classed by its (minus, plus) coupling pair, the classes must
// form at most DIAG_CLS runs per axis, one cell per class triple gives that
// triple's diagonal, and the classes are kept only if every interior cell's
// stored diagonal equals its triple's value bit for bit.
static falm_status diag_classes(falm_solver *S, const double *a, hipStream_t s) {
    const Geom &g = S->g;
    const Box &b = S->tl.box;
    const int n[3] = {g.nx, g.ny, g.nz};
    const size_t ns = 2 * ((size_t)g.nx + g.ny + g.nz);
    std::vector<double> h(ns);
    FALM_HIP(hipMemcpyAsync(h.data(), S->sep64, ns * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    int rep[3][DIAG_CLS], ncls[3] = {0, 0, 0};
    int rs[3][DIAG_CLS - 1], rc[3] = {0, 0, 0};
    size_t base = 0;  // the axis's minus array in h (its plus array follows)
    for (int m = 0; m < 3; m++) {
        const double *mi = h.data() + base, *pl = mi + n[m];
        int nrun = 0, prev = -1;
        for (int q = 0; q < DIAG_CLS - 1; q++) rs[m][q] = INT_MAX;
        for (int i = b.lo[m]; i < b.lo[m] + b.n[m]; i++) {
            int c = 0;
            while (c < ncls[m] && !(std::memcmp(&mi[i], &mi[rep[m][c]], sizeof(double)) == 0 &&
                                    std::memcmp(&pl[i], &pl[rep[m][c]], sizeof(double)) == 0))
                c++;
            if (c == ncls[m]) {
                if (c == DIAG_CLS) return FALM_OK;  // too many classes: the diagonal stays streamed
                rep[m][c] = i;
                ncls[m]++;
            }
            if (c != prev) {
                if (nrun == DIAG_CLS) return FALM_OK;  // too many runs
                if (nrun > 0) rs[m][nrun - 1] = i;
                rc[m] |= c << (8 * nrun);
                nrun++;
                prev = c;
            }
        }
        base += 2 * (size_t)n[m];
    }
    double t64[DIAG_CLS * DIAG_CLS * DIAG_CLS] = {};
    float t32[DIAG_CLS * DIAG_CLS * DIAG_CLS] = {};
    for (int cx = 0; cx < ncls[0]; cx++)
        for (int cy = 0; cy < ncls[1]; cy++)
            for (int cz = 0; cz < ncls[2]; cz++) {
                const int q = (cx * DIAG_CLS + cy) * DIAG_CLS + cz;
                const size_t cell = (size_t)g.id(rep[0][cx], rep[1][cy], rep[2][cz]);
                FALM_HIP(hipMemcpyAsync(&t64[q], a + cell, sizeof(double), hipMemcpyDeviceToHost, s));
            }
    FALM_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < DIAG_CLS * DIAG_CLS * DIAG_CLS; q++) t32[q] = (float)t64[q];
    FALM_HIP(hipMemcpyAsync(S->dtab64, t64, sizeof(t64), hipMemcpyHostToDevice, s));
    FALM_HIP(hipMemcpyAsync(S->dtab32, t32, sizeof(t32), hipMemcpyHostToDevice, s));
    std::memcpy(S->drs, rs, sizeof(rs));
    std::memcpy(S->drc, rc, sizeof(rc));
    S->dclass = true;  // for the view the check uses
    FALM_HIP(hipMemsetAsync(S->flag, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_diag_check, FALM_GRID(S), s, S->tl, g, sep_view<double>(S, a, S->sep64), S->flag);
    FALM_CHECK_LAUNCH("k_diag_check");
    FALM_HIP(hipMemcpyAsync(S->hflag, S->flag, sizeof(int), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    S->dclass = (*S->hflag == 0);
    return FALM_OK;
}

static falm_status prepare_matrix(falm_solver *S, const void *a, hipStream_t s) {
    const Geom &g = S->g;
    const int nmax = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
    double *d = S->sep64;
    hipLaunchKernelGGL((k_sep_extract<double>), dim3((nmax + 255) / 256), dim3(256), 0, s, g,
                       mat7<double>(a, S->n), d, d + g.nx, d + 2 * g.nx, d + 2 * g.nx + g.ny,
                       d + 2 * g.nx + 2 * g.ny, d + 2 * g.nx + 2 * g.ny + g.nz);
    FALM_CHECK_LAUNCH("k_sep_extract");
    FALM_HIP(hipMemsetAsync(S->flag, 0, sizeof(int), s));
    hipLaunchKernelGGL((k_sep_check<double>), FALM_GRID(S), s, S->tl, g, mat7<double>(a, S->n), d, d + g.nx,
                       d + 2 * g.nx, d + 2 * g.nx + g.ny, d + 2 * g.nx + 2 * g.ny, d + 2 * g.nx + 2 * g.ny + g.nz,
                       S->flag);
    FALM_CHECK_LAUNCH("k_sep_check");
    FALM_HIP(hipMemcpyAsync(S->hflag, S->flag, sizeof(int), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    S->sep = (*S->hflag == 0);
    S->dclass = false;
    if (S->sep) FALM_TRY(diag_classes(S, (const double *)a, s));
    if (S->cfg.precision == FALM_SOLVE_MIXED) {
        const size_t ncol = S->sep ? 1 : 7;
        hipLaunchKernelGGL((k_convert<double, float>), ew_grid(ncol * S->n), dim3(EW_BLOCK), 0, s, (const double *)a,
                           S->a32, ncol * S->n);
        const size_t ns = 2 * ((size_t)g.nx + g.ny + g.nz);
        hipLaunchKernelGGL((k_convert<double, float>), ew_grid(ns), dim3(EW_BLOCK), 0, s, (const double *)S->sep64,
                           S->sep32, ns);
        FALM_CHECK_LAUNCH("k_convert");
    }
    S->a_src = a;
    return FALM_OK;
}

falm_status solver_set_tol(falm_solver *S, double tol) {
    if (!S || !(tol > 0.0)) return FALM_ERR_INVALID;
    S->cfg.tol = tol;
    return FALM_OK;
}

falm_status solver_set_maxit(falm_solver *S, int maxit) {
    if (!S || maxit < 1) return FALM_ERR_INVALID;
    S->cfg.maxit = maxit;
    return FALM_OK;
}

}  // namespace falm

namespace falm {

// falm_solver_solve in one of the SolveMode forms.  The argument checks and
// the matrix preparation run in the call that starts the solve.
static falm_status solve_dispatch(falm_solver *S, const void *a, void *x, const void *b, void *r, hipStream_t s,
                                  falm_solve_info *info, int mode) {
    if (!S || !a || !x || !b) {
        set_error("falm_solver_solve: null argument");
        return FALM_ERR_INVALID;
    }
    if (x == r || x == b || r == b) {
        set_error("falm_solver_solve: x, b and r must be distinct");
        return FALM_ERR_ALIAS;
    }
    falm_solve_info tmp;
    if (!info) info = &tmp;
    if (!r && S->cfg.precision != FALM_SOLVE_MIXED) {
        // the double solvers iterate on r itself: use the solver's own vector
        if (!S->r64own) {
            FALM_HIP(hipMalloc(&S->r64own, S->n * sizeof(double)));
            FALM_HIP(hipMemsetAsync(S->r64own, 0, S->n * sizeof(double), s));
        }
        r = S->r64own;
    }
    double *xd = (double *)x, *rd = (double *)r;
    const double *bd = (const double *)b;
    const bool prepared = (S->a_src == a);
    if (!prepared && S->cfg.precision == FALM_SOLVE_MIXED) {
        // no declared matrix: build the float copy from this one now
        FALM_TRY(prepare_matrix(S, a, s));
    }
    const bool sep = (S->a_src == a) && S->sep;
    if (S->cfg.type == FALM_LS_PBICGSTAB) {
        if (S->cfg.precision == FALM_SOLVE_MIXED) {
            if (sep)
                return solve_bicgstab_mixed(S, sep_view<double>(S, (const double *)a, S->sep64),
                                            sep_view<float>(S, S->a32, S->sep32), xd, bd, rd, s, info, mode);
            return solve_bicgstab_mixed(S, cols_view<double>(S, a), cols_view<float>(S, S->a32), xd, bd, rd, s, info,
                                        mode);
        }
        if (sep)
            return solve_bicgstab_f64(S, sep_view<double>(S, (const double *)a, S->sep64), xd, bd, rd, s, info, mode);
        return solve_bicgstab_f64(S, cols_view<double>(S, a), xd, bd, rd, s, info, mode);
    }
    if (sep)
        return solve_stationary(S, sep_view<double>(S, (const double *)a, S->sep64), xd, bd, rd, s, info, mode);
    return solve_stationary(S, cols_view<double>(S, a), xd, bd, rd, s, info, mode);
}

falm_status solver_solve_mode(falm_solver *S, const void *a, void *x, const void *b, void *r, hipStream_t s,
                              falm_solve_info *info, int mode, bool *rerun) {
    const falm_status st = solve_dispatch(S, a, x, b, r, s, info, mode);
    if (rerun) *rerun = !S->first_ok;
    return st;
}

// The device flag that is set once the solve begun by SOLVE_BEGIN is over:
// the mixed solve's "final" (its passes ended), else "done".  Null when the
// last solve was not over after its first batch (the next one likely will
// not be either: the caller then runs the solve whole).
const int *solver_gate(const falm_solver *S) {
    if (!S->first_ok) return nullptr;
    return S->cfg.type == FALM_LS_PBICGSTAB && S->cfg.precision == FALM_SOLVE_MIXED ? &S->st->final : &S->st->done;
}

}  // namespace falm

extern "C" {

falm_status falm_solver_create(const falm_domain *dom, const falm_solver_config *cfg, falm_solver **out) {
    FALM_TRY(check_domain(dom));
    if (!cfg || !out) return FALM_ERR_INVALID;
    *out = nullptr;
    // the k-marching kernels address the double vectors (x, b, the true
    // residual) with 32-bit byte offsets
    FALM_TRY(check_byte_offsets(ncell_of(dom), sizeof(double)));
    if (cfg->type != FALM_LS_PBICGSTAB && cfg->type != FALM_LS_SOR && cfg->type != FALM_LS_JACOBI) {
        set_error("falm_solver_create: solver type %d not supported", cfg->type);
        return FALM_ERR_UNSUPPORTED;
    }
    if (cfg->precision != FALM_SOLVE_F64 && cfg->precision != FALM_SOLVE_MIXED) {
        set_error("falm_solver_create: precision %d not supported", cfg->precision);
        return FALM_ERR_UNSUPPORTED;
    }
    // the fused BiCGStab path (one-pass SOR preconditioner) stores vector
    // pairs: two elements per cell
    if (cfg->type == FALM_LS_PBICGSTAB && cfg->pc_type == FALM_LS_SOR && cfg->pc_maxit <= 1)
        FALM_TRY(check_byte_offsets(ncell_of(dom), 2 * (cfg->precision == FALM_SOLVE_MIXED ? sizeof(float)
                                                                                           : sizeof(double))));
    if (cfg->precision == FALM_SOLVE_MIXED && cfg->type != FALM_LS_PBICGSTAB) {
        set_error("falm_solver_create: mixed precision needs the BiCGStab solver");
        return FALM_ERR_UNSUPPORTED;
    }
    if (cfg->maxit < 1) {
        set_error("falm_solver_create: maxit must be >= 1");
        return FALM_ERR_INVALID;
    }
    int dev = 0, cus = 0;
    FALM_HIP(hipGetDevice(&dev));
    FALM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus < 1) cus = 1;
    falm_solver *S = new falm_solver();
    S->dom = *dom;
    S->cfg = *cfg;
    S->n = ncell_of(dom);
    S->g = make_geom(dom);
    S->tl = make_tiling(interior_box(S->g));
    S->kt = make_k7_tiling(interior_box(S->g));
    S->ktp = make_k7_tiling_fit(interior_box(S->g), cus * pcmv_wg_per_cu(cfg->precision != FALM_SOLVE_MIXED));
    {
        // cell pairs start at even offsets when the interior starts at an
        // even x index on rows of even length
        Box pb = interior_box(S->g);
        const bool even = pb.lo[0] % 2 == 0 && pb.n[0] % 2 == 0 && S->g.nx % 2 == 0;
        pb.n[0] = (pb.n[0] + 1) / 2;
        pb.lo[0] = 0;
        S->tp = make_tiling(pb);
        if (!even || S->tl.tiles == 0) S->tp.tiles = 0;
    }
    S->s3 = make_s3_tiling(interior_box(S->g));
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
    return prepare_matrix(S, a, (hipStream_t)stream);
}

falm_status falm_solver_solve(falm_solver *S, const void *a, void *x, const void *b, void *r, void *stream,
                              falm_solve_info *info) {
    return falm::solve_dispatch(S, a, x, b, r, (hipStream_t)stream, info, falm::SOLVE_FULL);
}

falm_status falm_solver_destroy(falm_solver *S) {
    if (!S) return FALM_OK;
    void *bufs[] = {S->rr, S->p, S->q, S->s, S->pp, S->ss, S->t, S->w, S->r32, S->a32, S->stp,
                    S->sep64, S->sep32, S->dtab64, S->dtab32, S->st, S->partials, S->partials2, S->flag,
                    S->r64own, S->trash};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (S->hst) (void)hipHostFree(S->hst);
    if (S->hflag) (void)hipHostFree(S->hflag);
    delete S;
    return FALM_OK;
}

}  // extern "C"
