// linalg.hpp — 7-point stencil operators and launch declarations shared by
// linalg.hip, solver.hip and the step driver.
#pragma once

#include <vector>

#include "common.hpp"

namespace falm {

// The 7 coefficient columns [c,w,e,s,n,b,t] (src/dev/MVDevCall.cu:25-31).
template <class TA>
struct Mat7 {
    const TA *__restrict__ c[7];
};

template <class TA>
inline Mat7<TA> mat7(const void *a, size_t ncell) {
    Mat7<TA> m;
    for (int q = 0; q < 7; q++) m.c[q] = col<TA>(a, ncell, q);
    return m;
}

// ac*xc + ae*xe + aw*xw + an*xn + as*xs + at*xt + ab*xb, summed in the
// reference's order (src/dev/MVDevCall.cu:39).
template <class TA, class TX, class R>
__device__ inline R apply7(const Mat7<TA> &a, const TX *__restrict__ x, int cc, int sy, int sz) {
    const R ac = (R)a.c[0][cc], aw = (R)a.c[1][cc], ae = (R)a.c[2][cc], as = (R)a.c[3][cc], an = (R)a.c[4][cc],
            ab = (R)a.c[5][cc], at = (R)a.c[6][cc];
    const R xc = (R)x[cc], xe = (R)x[cc + 1], xw = (R)x[cc - 1], xn = (R)x[cc + sy], xs = (R)x[cc - sy],
            xt = (R)x[cc + sz], xb = (R)x[cc - sz];
    return ac * xc + ae * xe + aw * xw + an * xn + as * xs + at * xt + ab * xb;
}

// Operator views used by the solver kernels.  OpCols reads the seven stored
// columns; OpSep reads a stored diagonal and six 1-D coefficient arrays
// (west/east indexed by i, south/north by j, bottom/top by k) — the exact
// stored values of a matrix whose off-diagonal columns each depend on one
// index only, which is what the Poisson assembly gives on a uniform
// tensor-product mesh.
template <class TA>
struct OpCols {
    const TA *__restrict__ c[7];
    __device__ inline void get(int cc, int, int, int, TA &ac, TA &aw, TA &ae, TA &as, TA &an, TA &ab,
                               TA &at) const {
        ac = c[0][cc]; aw = c[1][cc]; ae = c[2][cc]; as = c[3][cc]; an = c[4][cc]; ab = c[5][cc]; at = c[6][cc];
    }
    __device__ inline TA diag(int cc) const { return c[0][cc]; }
    // the same at a byte offset ob of the cell (ld_at addressing)
    __device__ inline void get_at(unsigned ob, int, int, int, TA &ac, TA &aw, TA &ae, TA &as, TA &an, TA &ab,
                                  TA &at) const {
        ac = ld_at(c[0], ob); aw = ld_at(c[1], ob); ae = ld_at(c[2], ob); as = ld_at(c[3], ob);
        an = ld_at(c[4], ob); ab = ld_at(c[5], ob); at = ld_at(c[6], ob);
    }
    __device__ inline TA diag_at(unsigned ob) const { return ld_at(c[0], ob); }
    __device__ inline TA diag_k(unsigned ob, int) const { return ld_at(c[0], ob); }
    // the view a k-marching kernel keeps for one column (nothing to hoist)
    __device__ inline OpCols column(int, int) const { return *this; }
    // a cell's coefficients, loaded ahead of use (k-marching prefetch)
    struct Cell {
        TA v[7];
    };
    __device__ inline Cell load(unsigned ob, int) const {
        Cell q;
#pragma unroll
        for (int m = 0; m < 7; m++) q.v[m] = ld_at(c[m], ob);
        return q;
    }
    template <class R>
    __device__ inline R apply(const Cell &q, R xc, R xe, R xw, R xn, R xs, R xt, R xb) const {
        return (R)q.v[0] * xc + (R)q.v[2] * xe + (R)q.v[1] * xw + (R)q.v[4] * xn + (R)q.v[3] * xs + (R)q.v[6] * xt +
               (R)q.v[5] * xb;
    }
    __device__ inline static TA diag(const Cell &q) { return q.v[0]; }
};

constexpr int DIAG_CLS = 4;  // diagonal classes per axis at most (OpSep)

// One column (i, j) of a separable operator: the four in-plane coefficients
// in registers, loaded once before a kernel walks the column's planes.
//
// With diagonal classes (see OpSep) the column also holds its diagonal on
// each z run, and a plane's diagonal is picked by comparing the plane index
// with the run starts (kernel arguments) instead of being loaded.
template <class TA>
struct OpSepCol {
    TA w, e, s, n;
    const TA *__restrict__ ac;
    const TA *__restrict__ ab, *__restrict__ at;
    TA d0, d1, d2, d3;  // the diagonal on z runs 0..3
    int z1, z2, z3;     // first plane of z runs 1..3
    bool cls;
    // selects of values read unconditionally (a select of member addresses
    // would keep the column view in scratch memory)
    __device__ inline TA pick(int k) const {
        const TA a0 = d0, a1 = d1, a2 = d2, a3 = d3;
        TA v = k >= z1 ? a1 : a0;
        v = k >= z2 ? a2 : v;
        return k >= z3 ? a3 : v;
    }
    __device__ inline void get_at(unsigned ob, int, int, int k, TA &c0, TA &w_, TA &e_, TA &s_, TA &n_, TA &b_,
                                  TA &t_) const {
        c0 = cls ? pick(k) : ld_at(ac, ob); w_ = w; e_ = e; s_ = s; n_ = n; b_ = ld_uniform(ab, k);
        t_ = ld_uniform(at, k);
    }
    __device__ inline TA diag_at(unsigned ob) const { return ld_at(ac, ob); }
    // the diagonal of the column's cell in plane k (byte offset ob)
    __device__ inline TA diag_k(unsigned ob, int k) const { return cls ? pick(k) : ld_at(ac, ob); }
    struct Cell {
        TA c0, b, t;
    };
    // k is the plane of a k-marching kernel: the same in every lane
    __device__ inline Cell load(unsigned ob, int k) const {
        return Cell{diag_k(ob, k), ld_uniform(ab, k), ld_uniform(at, k)};
    }
    // the reference's summation order (src/dev/MVDevCall.cu:39)
    template <class R>
    __device__ inline R apply(const Cell &q, R xc, R xe, R xw, R xn, R xs, R xt, R xb) const {
        return (R)q.c0 * xc + (R)e * xe + (R)w * xw + (R)n * xn + (R)s * xs + (R)q.t * xt + (R)q.b * xb;
    }
    __device__ inline static TA diag(const Cell &q) { return q.c0; }
};

//
// Diagonal classes: on a uniform mesh a cell's diagonal is set by which of
// its six couplings the assembly kept, so it takes a few values.  Each axis's
// indices are classed by their (minus, plus) coupling pair; the classes along
// an axis form at most DIAG_CLS runs of consecutive indices.  When the solver
// has checked that every interior cell's stored diagonal equals
// dtab[(cx*DIAG_CLS + cy)*DIAG_CLS + cz] bit for bit (cx, cy, cz: the
// classes of the cell's runs), dtab is set and the k-marching kernels take
// the diagonal from it; ac stays the stored column.
template <class TA>
struct OpSep {
    const TA *__restrict__ ac;
    const TA *__restrict__ aw, *__restrict__ ae, *__restrict__ as, *__restrict__ an, *__restrict__ ab,
        *__restrict__ at;
    const TA *__restrict__ dtab = nullptr;
    // per axis: first index of runs 1..3 (INT_MAX when absent) and the class
    // of run r in bits 8r..8r+7 (scalar members: the view is a kernel argument)
    int x1 = 0, x2 = 0, x3 = 0, y1 = 0, y2 = 0, y3 = 0, z1 = 0, z2 = 0, z3 = 0;
    int cxr = 0, cyr = 0, czr = 0;
    __device__ static inline int run_class(int v, int s1, int s2, int s3, int cr) {
        const int sh = v >= s2 ? (v >= s3 ? 24 : 16) : (v >= s1 ? 8 : 0);
        return (cr >> sh) & 0xff;
    }
    __device__ inline void get(int cc, int i, int j, int k, TA &c0, TA &w, TA &e, TA &s, TA &n, TA &b,
                               TA &t) const {
        c0 = ac[cc]; w = aw[i]; e = ae[i]; s = as[j]; n = an[j]; b = ab[k]; t = at[k];
    }
    __device__ inline TA diag(int cc) const { return ac[cc]; }
    __device__ inline void get_at(unsigned ob, int i, int j, int k, TA &c0, TA &w, TA &e, TA &s, TA &n, TA &b,
                                  TA &t) const {
        c0 = ld_at(ac, ob); w = aw[i]; e = ae[i]; s = as[j]; n = an[j]; b = ab[k]; t = at[k];
    }
    __device__ inline TA diag_at(unsigned ob) const { return ld_at(ac, ob); }
    // i, j must lie in the array (callers clamp inactive threads)
    __device__ inline OpSepCol<TA> column(int i, int j) const {
        OpSepCol<TA> c;
        c.w = aw[i]; c.e = ae[i]; c.s = as[j]; c.n = an[j];
        c.ac = ac; c.ab = ab; c.at = at;
        c.cls = dtab != nullptr;
        c.z1 = z1;
        c.z2 = z2;
        c.z3 = z3;
        c.d0 = c.d1 = c.d2 = c.d3 = TA(0);
        if (c.cls) {
            const TA *row =
                dtab + (run_class(i, x1, x2, x3, cxr) * DIAG_CLS + run_class(j, y1, y2, y3, cyr)) * DIAG_CLS;
            c.d0 = row[czr & 0xff];
            c.d1 = row[(czr >> 8) & 0xff];
            c.d2 = row[(czr >> 16) & 0xff];
            c.d3 = row[(czr >> 24) & 0xff];
        }
        return c;
    }
};

template <class OP, class TA, class TX, class R>
__device__ inline R apply_op(const OP &op, const TX *__restrict__ x, int cc, int i, int j, int k, int sy, int sz) {
    TA ac, aw, ae, as, an, ab, at;
    op.get(cc, i, j, k, ac, aw, ae, as, an, ab, at);
    const R xc = (R)x[cc], xe = (R)x[cc + 1], xw = (R)x[cc - 1], xn = (R)x[cc + sy], xs = (R)x[cc - sy],
            xt = (R)x[cc + sz], xb = (R)x[cc - sz];
    return (R)ac * xc + (R)ae * xe + (R)aw * xw + (R)an * xn + (R)as * xs + (R)at * xt + (R)ab * xb;
}

// the same sum with the seven x values given (k-marching kernels)
template <class OP, class TA, class R>
__device__ inline R apply_vals(const OP &op, int cc, int i, int j, int k, R xc, R xe, R xw, R xn, R xs, R xt, R xb) {
    TA ac, aw, ae, as, an, ab, at;
    op.get(cc, i, j, k, ac, aw, ae, as, an, ab, at);
    return (R)ac * xc + (R)ae * xe + (R)aw * xw + (R)an * xn + (R)as * xs + (R)at * xt + (R)ab * xb;
}

// apply_vals with the operator's per-cell columns at byte offset ob
template <class OP, class TA, class R>
__device__ inline R apply_vals_at(const OP &op, unsigned ob, int i, int j, int k, R xc, R xe, R xw, R xn, R xs, R xt,
                                  R xb) {
    TA ac, aw, ae, as, an, ab, at;
    op.get_at(ob, i, j, k, ac, aw, ae, as, an, ab, at);
    return (R)ac * xc + (R)ae * xe + (R)aw * xw + (R)an * xn + (R)as * xs + (R)at * xt + (R)ab * xb;
}

template <class T>
falm_status launch_sor(const falm_domain *dom, const void *a, void *x, const void *b, double omega, int color,
                       const int *done, hipStream_t s);
template <class T>
falm_status launch_jacobi(const falm_domain *dom, const void *a, void *x, const void *xp, const void *b,
                          const int *done, hipStream_t s);
template <class T>
falm_status launch_residual(const falm_domain *dom, const void *a, const void *x, const void *b, void *r,
                            hipStream_t s);
falm_status launch_dot(const falm_domain *dom, int prec, const void *a, const void *b, int ncol, bool norm,
                       double *dev_out, double *partials, hipStream_t s);
falm_status host_minmax(const falm_domain *dom, int prec, int op, const void *a, int ncol, int col, double *out,
                        hipStream_t s);

}  // namespace falm
