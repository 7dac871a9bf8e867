// km.hpp — k-marching column tiles for 7-point stencils.
//
// One-cell-per-thread stencil kernels on gfx950 are bound by vector-memory
// instruction issue: every neighbour load costs about as much as a streamed
// one, and double-precision neighbour loads twice that.  Here a workgroup of
// K7_BX x K7_BY threads owns that column of cells and walks a chunk of planes
// in k.  For each stencil array a Plane7 keeps
//   * the column's values at k-1, k, k+1 in registers (the k neighbours), and
//   * plane k with a one-cell halo in LDS (the x and y neighbours),
// so a cell costs one streamed load per array plus a share of the halo.  The
// next plane's loads are issued before the current plane is computed and the
// LDS planes are double-buffered: one barrier per plane.
#pragma once

#include "common.hpp"

namespace falm {

constexpr int K7_BX = 64, K7_BY = 8, K7_NT = K7_BX * K7_BY;
constexpr int K7_HALO = 2 * K7_BY + 2 * K7_BX;  // one halo cell per thread at most
static_assert(K7_HALO <= K7_NT, "halo slots exceed threads");

struct K7Tiling {
    Box box;
    int tx, ty, tk;  // column tiles along x, y and k chunks
    int kc;          // planes per chunk
    int blocks;
};

// planes per chunk: K7_KC, doubled until the grid fits max_blocks workgroups
#ifndef FALM_K7_KC
#define FALM_K7_KC 16
#endif
constexpr int K7_KC = FALM_K7_KC;
inline K7Tiling make_k7_tiling(const Box &b, int max_blocks = RED_BLOCKS) {
    K7Tiling t;
    t.box = b;
    t.tx = (b.n[0] + K7_BX - 1) / K7_BX;
    t.ty = (b.n[1] + K7_BY - 1) / K7_BY;
    t.kc = K7_KC;
    for (;;) {
        t.tk = (b.n[2] + t.kc - 1) / t.kc;
        t.blocks = t.tx * t.ty * t.tk;
        if (t.blocks <= max_blocks || t.kc >= b.n[2]) break;
        t.kc *= 2;
    }
    if (b.n[0] <= 0 || b.n[1] <= 0 || b.n[2] <= 0) t.blocks = 0;
    return t;
}

// the shortest chunk (at least K7_KC planes) whose grid fits max_blocks
inline K7Tiling make_k7_tiling_fit(const Box &b, int max_blocks) {
    K7Tiling t;
    t.box = b;
    t.tx = (b.n[0] + K7_BX - 1) / K7_BX;
    t.ty = (b.n[1] + K7_BY - 1) / K7_BY;
    const int cols = t.tx * t.ty > 0 ? t.tx * t.ty : 1;
    int tk = max_blocks / cols;
    if (tk < 1) tk = 1;
    t.kc = (b.n[2] + tk - 1) / tk;
    if (t.kc < K7_KC) t.kc = K7_KC;
    t.tk = (b.n[2] + t.kc - 1) / t.kc;
    t.blocks = t.tx * t.ty * t.tk;
    if (b.n[0] <= 0 || b.n[1] <= 0 || b.n[2] <= 0) t.blocks = 0;
    return t;
}

struct K7Pos {
    int lane, wy;   // thread's column within the tile
    int i0, j0;     // tile origin
    int i, j;       // thread's cell column
    int k0, k1;     // plane range [k0, k1)
    int col0;       // offset of the column's plane-0 cell to load from
    bool act;       // the column is inside the box
};

__device__ inline K7Pos k7_pos(const K7Tiling &kt, const Geom &g) {
    K7Pos P;
    P.lane = (int)(threadIdx.x & (K7_BX - 1));
    P.wy = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / K7_BX));
    const int L = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x));
    const int bx = L % kt.tx, by = (L / kt.tx) % kt.ty, kc = L / (kt.tx * kt.ty);
    P.i0 = kt.box.lo[0] + bx * K7_BX;
    P.j0 = kt.box.lo[1] + by * K7_BY;
    P.i = P.i0 + P.lane;
    P.j = P.j0 + P.wy;
    P.act = P.i < kt.box.lo[0] + kt.box.n[0] && P.j < kt.box.lo[1] + kt.box.n[1];
    // a thread past the box still provides its cell to the LDS plane (its
    // neighbours inside the box read it); one past the array loads a valid
    // cell instead (no box cell reads it)
    const bool ld = P.i < g.nx && P.j < g.ny;
    P.col0 = (ld ? P.i : kt.box.lo[0]) + g.nx * (ld ? P.j : kt.box.lo[1]);
    P.k0 = kt.box.lo[2] + kc * kt.kc;
    P.k1 = min(P.k0 + kt.kc, kt.box.lo[2] + kt.box.n[2]);
    return P;
}

template <class T>
using K7Lds = T[2][K7_BY + 2][K7_BX + 2];

// The halo cell a thread fills in the LDS plane: the x halo (rows 1..BY,
// columns 0 and BX+1) on the first 2*BY threads, the y halo (rows 0 and BY+1,
// columns 1..BX) on the next 2*BX; (gi, gj) is its cell column.
struct K7Halo {
    int row, col, gi, gj;
    bool slot;
};

__device__ inline K7Halo k7_halo(const K7Pos &P) {
    K7Halo h;
    const int s = (int)threadIdx.x;
    h.slot = s < K7_HALO;
    h.row = h.col = 0;
    if (s < 2 * K7_BY) {
        h.row = 1 + (s >> 1);
        h.col = (s & 1) ? K7_BX + 1 : 0;
    } else if (h.slot) {
        const int r = s - 2 * K7_BY, q = r / K7_BX;
        h.row = q ? K7_BY + 1 : 0;
        h.col = 1 + (r - q * K7_BX);
    }
    h.gi = P.i0 - 1 + h.col;
    h.gj = P.j0 - 1 + h.row;
    return h;
}

// One stencil array walked through the planes of a K7 column tile.  Loads
// use byte offsets from the array's base (ld_at), so arrays walked with the
// same column share one address register.
template <class T>
struct Plane7 {
    const T *__restrict__ a;
    K7Lds<T> *lds;
    T q0, q1, q2;  // column values at k-1, k, k+1
    T nq, hv;      // prefetched: column at k+2, this thread's halo cell at k+1
    unsigned cb, hb;  // byte offsets of the column's and the halo cell's plane-0 cells
    int hrow, hcol, lane, wy;
    bool hslot, hok;

    __device__ inline void setup(const T *arr, K7Lds<T> *l, const K7Pos &P, const Geom &g) {
        a = arr;
        lds = l;
        lane = P.lane;
        wy = P.wy;
        const K7Halo h = k7_halo(P);
        hslot = h.slot;
        hrow = h.row;
        hcol = h.col;
        const int gi = h.gi, gj = h.gj;
        hok = hslot && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        cb = (unsigned)P.col0 * sizeof(T);
        // a halo slot outside the array reads the column's own cell (value unused)
        hb = hok ? (unsigned)(gi + g.nx * gj) * sizeof(T) : cb;
    }
    // registers for plane k0 and LDS buffer 0 (the caller synchronises)
    __device__ inline void prologue(const K7Pos &P, int sz) {
        const unsigned szb = (unsigned)sz * sizeof(T);
        q0 = ld_at(a, cb + (unsigned)(P.k0 - 1) * szb);
        q1 = ld_at(a, cb + (unsigned)P.k0 * szb);
        q2 = ld_at(a, cb + (unsigned)(P.k0 + 1) * szb);
        hv = T(0);
        if (hslot) hv = ld_at(a, hb + (unsigned)P.k0 * szb);  // the halo threads only (the first waves)
        fill(0, q1);
    }
    // loads for plane k+1 (call only when k+1 < k1).  The raw values are kept
    // and masked only when they go to LDS, so nothing here waits for a load.
    __device__ inline void prefetch(const K7Pos &P, int k, int sz) {
        const unsigned szb = (unsigned)sz * sizeof(T);
        nq = ld_at(a, cb + (unsigned)(k + 2) * szb);
        if (hslot) hv = ld_at(a, hb + (unsigned)(k + 1) * szb);
    }
    __device__ inline void fill(int b, T centre) {
        lds[0][b][wy + 1][lane + 1] = centre;
        if (hslot) lds[0][b][hrow][hcol] = hok ? hv : T(0);
    }
    // after plane k: LDS for plane k+1 and the register shift
    __device__ inline void advance(int b, bool more) {
        if (more) fill(b ^ 1, q2);
        q0 = q1;
        q1 = q2;
        q2 = nq;
    }
    __device__ inline T c() const { return q1; }
    __device__ inline T zm() const { return q0; }
    __device__ inline T zp() const { return q2; }
    __device__ inline T xm(int b) const { return lds[0][b][wy + 1][lane]; }
    __device__ inline T xp(int b) const { return lds[0][b][wy + 1][lane + 2]; }
    __device__ inline T ym(int b) const { return lds[0][b][wy][lane + 1]; }
    __device__ inline T yp(int b) const { return lds[0][b][wy + 2][lane + 1]; }
};

}  // namespace falm
