// common.hpp — shared host/device helpers of the gfx950 library.
//
// Layout: the reference's Matrix layout (src/matrix.h:20), one column per
// component, x fastest over the padded local shape.  Kernels receive one
// pointer per column so that no 64-bit column arithmetic happens on device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <initializer_list>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/falm_amd.h"

namespace falm {

// ------------------------------------------------------------------ errors
void set_error(const char *fmt, ...);
falm_status hip_status(hipError_t e, const char *what);

#define FALM_HIP(expr)                                                      \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) return ::falm::hip_status(_e, #expr);         \
    } while (0)

#define FALM_TRY(expr)                                                      \
    do {                                                                    \
        falm_status _s = (expr);                                            \
        if (_s != FALM_OK) return _s;                                       \
    } while (0)

#define FALM_CHECK_LAUNCH(name) FALM_HIP(hipGetLastError())

// ----------------------------------------------------------------- domain
// Device-side view of a falm_domain (src/region.h:8-31 + CPM::gc).
struct Geom {
    int nx, ny, nz;     // padded local shape
    int gc;
    int osum;           // sum of pdm.offset: parity term of the RB colour
    int nxy;            // nx*ny
    __host__ __device__ inline int id(int i, int j, int k) const { return i + nx * j + nxy * k; }
};

inline Geom make_geom(const falm_domain *d) {
    Geom g;
    g.nx = d->shape[0];
    g.ny = d->shape[1];
    g.nz = d->shape[2];
    g.gc = d->gc;
    g.osum = d->offset[0] + d->offset[1] + d->offset[2];
    g.nxy = g.nx * g.ny;
    return g;
}

inline size_t ncell_of(const falm_domain *d) {
    return (size_t)d->shape[0] * (size_t)d->shape[1] * (size_t)d->shape[2];
}

falm_status check_domain(const falm_domain *d);
// FALM_ERR_INVALID unless every pointer is hipMalloc'd device memory (the
// adjoint kernels accumulate into such outputs with FP64 atomics)
falm_status vjp_outputs_on_device(std::initializer_list<const void *> outs);

// A box of cells [lo, lo+n) in padded local indices.
struct Box {
    int lo[3];
    int n[3];
    __host__ __device__ inline long long size() const { return (long long)n[0] * n[1] * n[2]; }
};

inline Box interior_box(const Geom &g) {
    return Box{{g.gc, g.gc, g.gc}, {g.nx - 2 * g.gc, g.ny - 2 * g.gc, g.nz - 2 * g.gc}};
}
// Region(pdm.shape, gc) with shape += 1 and offset -= 1 (src/FalmCFD.cpp:119-126).
inline Box face_box(const Geom &g) {
    return Box{{g.gc - 1, g.gc - 1, g.gc - 1},
               {g.nx - 2 * g.gc + 1, g.ny - 2 * g.gc + 1, g.nz - 2 * g.gc + 1}};
}

// ----------------------------------------------------------- launch shape
// Stencil launches use 64x4 threads (one wavefront per x row segment) over a
// 1-D grid.  The 1-D block id is remapped so that each of the 8 XCDs of an
// MI355X (workgroups are dealt to XCDs round-robin) gets one contiguous
// slab of tiles: neighbouring tiles then share that XCD's L2.
constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NXCD = 8;

struct Tiling {
    Box box;
    int tx, ty;          // tiles along x, y
    int tiles;           // tx*ty*box.n[2]
    __host__ __device__ inline int ntiles() const { return tiles; }
};

inline Tiling make_tiling(const Box &b) {
    Tiling t;
    t.box = b;
    t.tx = (b.n[0] + TX - 1) / TX;
    t.ty = (b.n[1] + TY - 1) / TY;
    t.tiles = t.tx * t.ty * b.n[2];
    return t;
}

// physical block id -> logical tile id, XCD-contiguous
__device__ inline int xcd_tile(int bid, int nblocks) {
    int per = (nblocks + NXCD - 1) / NXCD;
    int xcd = bid % NXCD;
    int loc = bid / NXCD;
    int full = nblocks % NXCD;          // XCDs that receive `per` blocks
    // blocks are dealt round robin: XCDs [0, full) get `per` blocks and the
    // rest get per-1 when nblocks is not a multiple of NXCD
    if (full == 0) return xcd * per + loc;
    if (xcd < full) return xcd * per + loc;
    return full * per + (xcd - full) * (per - 1) + loc;
}

// Tile order: x tile fastest, then y, then k, so consecutive workgroups
// stream through contiguous memory (a k-fastest order that keeps stencil
// neighbours in flight together measured slower: every tile then starts a
// new DRAM page).
__host__ __device__ inline void tile_coords(const Tiling &t, int tile, int &bx, int &by, int &kk) {
    const int txy = t.tx * t.ty;
    kk = tile / txy;
    const int r = tile - kk * txy;
    by = r / t.tx;
    bx = r - by * t.tx;
}

// cell coordinates of this thread in tile `tile`; false outside the box.
// A wavefront covers one x row of the tile, so j and k are wave-uniform:
// readfirstlane makes that visible to the compiler, which then keeps them
// (and loads indexed by them alone) in scalar registers.
__device__ inline bool tile_cell_at(const Tiling &t, int tile, int &i, int &j, int &k) {
    int bx, by;
    tile = __builtin_amdgcn_readfirstlane(tile);
    tile_coords(t, tile, bx, by, k);
    i = bx * TX + (int)(threadIdx.x & (TX - 1));
    j = by * TY + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (i >= t.box.n[0] || j >= t.box.n[1]) return false;
    i += t.box.lo[0];
    j += t.box.lo[1];
    k += t.box.lo[2];
    return true;
}

// Reduction kernels run at most RED_BLOCKS workgroups; each walks a
// contiguous run of tiles (of the XCD-contiguous order), so the finalize
// kernel sums at most RED_BLOCKS partials.
constexpr int RED_BLOCKS = 4096;
inline dim3 red_grid(const Tiling &t) { return dim3((unsigned)(t.tiles < RED_BLOCKS ? t.tiles : RED_BLOCKS)); }
inline int red_blocks(const Tiling &t) { return t.tiles < RED_BLOCKS ? t.tiles : RED_BLOCKS; }
__device__ inline void tile_range(int ntiles, int &t0, int &t1) {
    const int L = xcd_tile(blockIdx.x, gridDim.x);
    const int C = (ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
    t0 = L * C;
    t1 = t0 + C < ntiles ? t0 + C : ntiles;
}

// The non-reducing kernels take one tile per workgroup (one_tile_grid):
// measured on MI355X, walking several tiles per workgroup made the stencil
// kernels slower.  f(i, j, k) is called for this thread's cell of the tile
// (return in f ends the thread's work).
template <class F>
__device__ inline void for_each_cell(const Tiling &t, F f) {
    int i, j, k;
    if (tile_cell_at(t, xcd_tile(blockIdx.x, gridDim.x), i, j, k)) f(i, j, k);
}

// cell coordinates of this thread for one tile per workgroup; false outside
// the box
__device__ inline bool tile_cell(const Tiling &t, int &i, int &j, int &k) {
    return tile_cell_at(t, xcd_tile(blockIdx.x, gridDim.x), i, j, k);
}

// one tile per workgroup (for_each_cell, tile_cell)
inline dim3 tiling_grid(const Tiling &t) { return dim3((unsigned)t.tiles); }
inline dim3 one_tile_grid(const Tiling &t) { return dim3((unsigned)t.tiles); }
inline dim3 tiling_block() { return dim3(TX * TY); }

// Access at an unsigned 32-bit byte offset from a wave-uniform base pointer.
// The compiler then emits the SGPR-base form of the global load/store (one
// VGPR offset shared by every array indexed at the same cell, constant
// neighbour offsets folded into the instruction), where a signed element
// index costs a 64-bit address computation per access.  Arrays addressed
// this way must be below 4 GiB (checked by the launchers, by
// falm_solver_create and by falm_case_create).
template <class T>
__device__ __forceinline__ T ld_at(const T *p, unsigned off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + off);
}
template <class T>
__device__ __forceinline__ void st_at(T *p, unsigned off, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + off) = v;
}

// Read of a wave-uniform element of a table the kernel never writes (the
// per-axis coefficient tables), through the constant address space: the
// compiler then issues a scalar load, counted by lgkmcnt.  A plain load of
// such a value is a vector load when the kernel also stores to global
// memory, and waiting for it (vmcnt) also waits for every vector load issued
// before it: in a k-marching loop, for the prefetches of the next planes.
// p and i must be the same in every lane of the wave.
template <class T>
__device__ __forceinline__ T ld_uniform(const T *p, int i) {
    return ((const __attribute__((address_space(4))) T *)p)[i];
}

// the ld_at/st_at kernels address one field column of ncell elements
inline falm_status check_byte_offsets(size_t ncell, size_t elem) {
    if (ncell * elem >= (size_t(1) << 32)) {
        set_error("a field column exceeds 4 GiB (the stencil kernels use 32-bit byte offsets)");
        return FALM_ERR_UNSUPPORTED;
    }
    return FALM_OK;
}

// 1-D element-wise launches
constexpr int EW_BLOCK = 256;
inline dim3 ew_grid(size_t n) {
    size_t b = (n + EW_BLOCK - 1) / EW_BLOCK;
    if (b > 2147483647ull) b = 2147483647ull;
    return dim3((unsigned)b);
}

// ------------------------------------------------------------- reductions
// Wavefront (64 lanes) then block reduction of doubles.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <int NT>
__device__ inline double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < NT / 64; w++) r += lds[w];
    }
    __syncthreads();
    return r;  // valid in thread 0
}

// Multi-value block sum (NV values per thread), result valid in thread 0.
template <int NT, int NV>
__device__ inline void block_sum_n(double (&v)[NV], double *lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = wave_sum(v[q]);
        if (lane == 0) lds[q * (NT / 64) + wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double r = 0.0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) r += lds[q * (NT / 64) + w];
            v[q] = r;
        }
    }
    __syncthreads();
}

// Deterministic two-phase grid reduction.  Phase 1 (in the producing
// kernel): each workgroup reduces its values and writes one partial per
// value with a plain store.  Phase 2 (a separate one-workgroup kernel that
// the stream orders after phase 1): the partials are summed in a fixed
// order.  No device-scope fences or atomics are needed: a per-workgroup
// release fence would write back the XCD's L2 in every one of tens of
// thousands of workgroups.
template <int NT, int NV>
__device__ inline void block_partials(double (&v)[NV], double *partials, double *lds) {
    block_sum_n<NT, NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) partials[(size_t)q * gridDim.x + blockIdx.x] = v[q];
    }
}

// A launch gated on a device flag: the kernel does nothing unless *gate is
// nonzero (gate null: always runs).  The step driver enqueues the kernels
// after the Poisson solve gated on the solver's "finished" flag before it
// knows whether the first batch of iterations finished the solve.
__device__ inline bool gated_off(const int *gate) { return gate && !*gate; }

constexpr int FIN_NT = 1024;  // threads of a finalize workgroup

// Phase 2 helper for an NT-thread workgroup: totals of nb partials per
// value, valid in every thread.  Used by the one-workgroup finalize kernels,
// and by consumer kernels whose every workgroup needs a reduction's result:
// each sums the same partials in the same order, so all get the same bits
// (no finalize launch between producer and consumer).
template <int NV, int NT = FIN_NT>
__device__ inline void sum_partials(const double *partials, int nb, double (&tot)[NV]) {
    __shared__ double lds[NV * (NT / 64)];
    __shared__ double bc[NV];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += NT) s += partials[(size_t)q * nb + b];
        acc[q] = s;
    }
    block_sum_n<NT, NV>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) bc[q] = acc[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) tot[q] = bc[q];
}

// ------------------------------------------------------------ precision
template <class T> struct PrecOf;
template <> struct PrecOf<double> { static constexpr int value = FALM_F64; };
template <> struct PrecOf<float> { static constexpr int value = FALM_F32; };

inline size_t prec_bytes(int prec) { return prec == FALM_F32 ? 4 : 8; }

// column pointer helper
template <class T> inline T *col(void *p, size_t ncell, int c) { return (T *)p + (size_t)c * ncell; }
template <class T> inline const T *col(const void *p, size_t ncell, int c) {
    return (const T *)p + (size_t)c * ncell;
}

// -------------------------------------------------------------- scratch
// Library-owned device scratch for one-shot reductions (per device).
struct ReduceScratch {
    double *partials = nullptr;  // 8 * max_blocks doubles
    double *result = nullptr;    // 8 doubles
    double *host = nullptr;      // pinned host mirror of result
    size_t max_blocks = 0;
};
falm_status reduce_scratch(size_t nblocks, ReduceScratch **out);
// device address of falm_dot_noalloc's workspace on the current device,
// taken on first use (falm_init calls it for its device)
falm_status noalloc_setup(double **out);

// ------------------------------------------------------------ timing
struct TimedLaunch {
    int family;
    hipStream_t stream;
    bool on;
    TimedLaunch(int fam, hipStream_t s);
    void done(double bytes);
    // bytes known only after the launch (e.g. a box the kernels chose);
    // counts toward the family's total when the launch was timed
    void add_bytes(double bytes);
};

}  // namespace falm
