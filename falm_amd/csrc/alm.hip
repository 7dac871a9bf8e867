// alm.hip — actuator-line model (src/alm/*) on gfx950.
//
// k_alm_points: one thread per actuator point does what the reference does
//   in three launches (kernel_UpdateTurbineAngles, kernel_UpdateAPX,
//   kernel_CalcAPForce, src/alm/almDevCall.cu:43-253).
// k_alm_factors + k_alm_project: the Gaussian projection
//   (kernel_DistributeAPForce, :255-294) as a cell-major gather.  The kernel
//   is separable on the tensor-product grid, exp(-r^2/eps^2) =
//   ex(x_i) ey(y_j) ez(z_k), so each point's three 1-D factor rows are
//   evaluated once per step and a cell-point pair costs two products
//   instead of an exp.  A fixed grid walks 32x8-cell tiles of the force box
//   (of the whole interior on a full write); each tile first culls, in LDS,
//   the points whose kernel value is exactly zero in double precision over
//   the whole tile (r^2/eps^2 > 746), keeping the survivors in ascending id
//   order; a finite cutoff (radius in units of eps) culls harder.
// k_alm_totals: per-turbine torque/thrust in a fixed order (the reference
//   uses thrust::reduce, src/alm/almDevCall.cu:362-370), kept on device.
#include <cstddef>
#include <cmath>
#include <vector>

#include "alm.hpp"

namespace falm {

constexpr double kPi = 3.14159265358979323846;  // Falm::Pi (src/falmath.h:10)

struct TurbDev {
    // bvel: base velocity set with falm_alm_set_base (0 unless set);
    // yaw, aoff: the rotor's static yaw and the platform angle's static part
    // (radians)
    double base[3], bvel[3], hub[3], tip_rate, motion[3], surge[3], pitch_deg[8];
    double yaw, aoff;
    int angle_type;
};

struct AlmDev {
    int nt, nb, napb, nap, nat;
    double radius, hub_radius, eps;
    int flags;
    int nx, ny, nz, gc;
    const TurbDev *tb;
    const double *r, *chord, *twist, *attack, *cl, *cd, *twist_off;
    const double *x, *y, *z;
    double *ap_xyz, *ap_force, *ap_torque, *ap_thrust, *torque, *thrust;
    int *ap_ijk, *ap_rank;
    double *fx, *fy, *fz;  // per-point axis factors exp(-(x_i - X)^2/eps^2): [nap][nx], [nap][ny], [nap][nz]
    double cutr;           // culling radius (length)
    // cell boxes (ilo, ihi, jlo, jhi, klo, khi; empty when ilo > ihi):
    // [0,6) cells that may receive force this step, [6,12) its union with
    // the previous step's, [12,18) the previous step's
    int *box;
};

// per-point, per-axis Gaussian factors of the projection
__global__ void k_alm_factors(AlmDev A) {
    const int L = A.nx + A.ny + A.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nap * L) return;
    const int ap = t / L;
    int r = t - ap * L;
    const double eps2 = A.eps * A.eps;
    double v;
    double *out;
    if (r < A.nx) {
        v = A.x[r] - A.ap_xyz[3 * ap];
        out = A.fx + (size_t)ap * A.nx + r;
    } else if ((r -= A.nx) < A.ny) {
        v = A.y[r] - A.ap_xyz[3 * ap + 1];
        out = A.fy + (size_t)ap * A.ny + r;
    } else {
        r -= A.ny;
        v = A.z[r] - A.ap_xyz[3 * ap + 2];
        out = A.fz + (size_t)ap * A.nz + r;
    }
    *out = exp(-(v * v) / eps2);
}

// one_angle_frame_rotation, src/falmath.h:55-81
__device__ inline void rot(const double x[3], const double ang[3], int type, double out[3]) {
    const double X = x[0], Y = x[1], Z = x[2];
    if (type == 1) {
        const double C = cos(ang[0]), S = sin(ang[0]);
        out[0] = X; out[1] = C * Y + S * Z; out[2] = -S * Y + C * Z;
    } else if (type == 2) {
        const double C = cos(ang[1]), S = sin(ang[1]);
        out[0] = C * X - S * Z; out[1] = Y; out[2] = S * X + C * Z;
    } else if (type == 3) {
        const double C = cos(ang[2]), S = sin(ang[2]);
        out[0] = C * X + S * Y; out[1] = -S * X + C * Y; out[2] = Z;
    } else {
        out[0] = X; out[1] = Y; out[2] = Z;
    }
}

// one_angle_frame_rotation_dt, src/falmath.h:83-131
__device__ inline void rot_dt(const double x[3], const double v[3], const double ang[3], const double om[3], int type,
                              double out[3]) {
    const double X = x[0], Y = x[1], Z = x[2], U = v[0], V = v[1], W = v[2];
    if (type == 1) {
        const double C = cos(ang[0]), S = sin(ang[0]), O = om[0];
        out[0] = 0. + U; out[1] = O * (-S * Y + C * Z) + (C * V + S * W); out[2] = O * (-C * Y - S * Z) + (-S * V + C * W);
    } else if (type == 2) {
        const double C = cos(ang[1]), S = sin(ang[1]), O = om[1];
        out[0] = O * (-S * X - C * Z) + (C * U - S * W); out[1] = 0. + V; out[2] = O * (C * X - S * Z) + (S * U + C * W);
    } else if (type == 3) {
        const double C = cos(ang[2]), S = sin(ang[2]), O = om[2];
        out[0] = O * (-S * X + C * Y) + (C * U + S * V); out[1] = O * (-C * X - S * Y) + (-S * U + C * V); out[2] = 0. + W;
    } else {
        out[0] = U; out[1] = V; out[2] = W;
    }
}

// Rotor frame -> inertial frame for a yawed rotor on a moving platform: x
// turned by -yaw about z (the rotor's own yaw, no reference counterpart),
// then by -ang about the platform's axis (one_angle_frame_rotation with the
// negated angles, src/alm/almDevCall.cu:88-104).  yaw = 0 is the reference.
__device__ inline void to_inertial(const double x[3], double yaw, const double nang[3], int type, double out[3]) {
    if (yaw != 0.0) {
        const double ny[3] = {0.0, 0.0, -yaw};
        double y[3];
        rot(x, ny, 3, y);
        rot(y, nang, type, out);
    } else {
        rot(x, nang, type, out);
    }
}

// The platform frame's relative velocity turned into the yawed rotor frame.
__device__ inline void to_rotor(double v[3], double yaw) {
    if (yaw != 0.0) {
        const double py[3] = {0.0, 0.0, yaw};
        const double w[3] = {v[0], v[1], v[2]};
        rot(w, py, 3, v);
    }
}

// find_index, src/alm/almDevCall.cu:12-26: the first i with
// seq[i] <= v < seq[i+1].  The cell-centre arrays are strictly increasing,
// so a bisection finds the same i as the reference's linear scan.
__device__ inline int find_index(const double *seq, double v, int size) {
    if (v < seq[0]) return -1;
    if (v >= seq[size - 1] || v != v) return size - 1;  // NaN: the scan finds nothing
    int lo = 0, hi = size - 1;  // seq[lo] <= v < seq[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seq[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// trilinear_interpolation, src/alm/almDevCall.cu:28-41
__device__ inline double trilinear(double xd, double yd, double zd, const double c[8]) {
    const double c01 = c[0] * (1. - xd) + c[1] * xd;
    const double c23 = c[2] * (1. - xd) + c[3] * xd;
    const double c45 = c[4] * (1. - xd) + c[5] * xd;
    const double c67 = c[6] * (1. - xd) + c[7] * xd;
    const double c0123 = c01 * (1. - yd) + c23 * yd;
    const double c4567 = c45 * (1. - yd) + c67 * yd;
    return c0123 * (1. - zd) + c4567 * zd;
}

// The 1-D cell-centre arrays and the attack-angle table are staged in LDS
// when they fit (kPointsLds bytes), so the searches below are LDS-latency
// chains instead of dependent global loads.
constexpr int kPointsLds = 48 * 1024;

__device__ void alm_box_wg(const AlmDev &A, const double *pxyz, const double *X, const double *Y, const double *Z);

// fused (one workgroup holds every point): the force box and the
// per-turbine totals follow in the same launch, from LDS copies of the
// points' positions, torque and thrust
constexpr int kFusedPoints = 256;

template <class T>
__global__ __launch_bounds__(256) void k_alm_points(AlmDev A, const T *u0, const T *u1, const T *u2, double t,
                                                    int staged, int fused) {
    extern __shared__ double stage[];
    __shared__ double sxyz[3 * kFusedPoints], stq[kFusedPoints], sth[kFusedPoints];
    // The point's parameters are loaded first, so that they are in flight
    // with the staging loads, and every result is stored at the end: the
    // point's work is then three global round trips (parameters and tables,
    // the eight velocity corners, the polar entries) instead of one per
    // store/load alternation.
    const int ap = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = ap < A.nap;
    const int apc = live ? ap : 0;
    const int napt = A.napb * A.nb;
    const int tid = apc / napt, bid = (apc % napt) / A.napb, sec = apc % A.napb;
    const TurbDev &T0 = A.tb[tid];
    const int atype = T0.angle_type;
    const double tip = T0.tip_rate;
    const double mo0 = T0.motion[0], mo1 = T0.motion[1], mo2 = T0.motion[2];
    const double su0 = T0.surge[0], su1 = T0.surge[1], su2 = T0.surge[2];
    double base[3] = {T0.base[0], T0.base[1], T0.base[2]}, bvel[3] = {T0.bvel[0], T0.bvel[1], T0.bvel[2]};
    const double hub[3] = {T0.hub[0], T0.hub[1], T0.hub[2]};
    const double apr = A.r[apc], chord = A.chord[apc];
    // APFrame::get_airfoil_params (src/alm/apHandler.h:121-141); blade pitch
    // and per-section twist offset have no reference counterpart (0 default)
    const double twist = A.twist[apc] + T0.pitch_deg[bid < 8 ? bid : 7] + A.twist_off[sec];
    const double *X = A.x, *Y = A.y, *Z = A.z, *AT = A.attack;
    if (staged) {
        const int L = A.nx + A.ny + A.nz + A.nat;
        for (int q = threadIdx.x; q < L; q += blockDim.x) {
            double v;
            if (q < A.nx) v = A.x[q];
            else if (q < A.nx + A.ny) v = A.y[q - A.nx];
            else if (q < A.nx + A.ny + A.nz) v = A.z[q - A.nx - A.ny];
            else v = A.attack[q - A.nx - A.ny - A.nz];
            stage[q] = v;
        }
        __syncthreads();
        X = stage;
        Y = stage + A.nx;
        Z = stage + A.nx + A.ny;
        AT = stage + A.nx + A.ny + A.nz;
    }
    [&]() {  // one point (a return ends this point's work only)
    if (!live) return;
    // kernel_UpdateTurbineAngles (+ base surge, no reference counterpart)
    double ang[3] = {0, 0, 0}, om[3] = {0, 0, 0};
    if (atype) {
        const double sa = mo0 * sin(mo1 * t + mo2), oa = mo0 * mo1 * cos(mo1 * t + mo2);
        const double sb = T0.aoff != 0.0 ? T0.aoff + sa : sa;
        ang[0] = atype == 1 ? sb : 0.0;
        ang[1] = atype == 2 ? sb : 0.0;
        ang[2] = atype == 3 ? sb : 0.0;
        om[0] = atype == 1 ? oa : 0.0;
        om[1] = atype == 2 ? oa : 0.0;
        om[2] = atype == 3 ? oa : 0.0;
    }
    if (su0 != 0.0) {
        base[0] = base[0] + su0 * sin(su1 * t + su2);
        bvel[0] = bvel[0] + su0 * su1 * cos(su1 * t + su2);
    }
    const double nang[3] = {-ang[0], -ang[1], -ang[2]};
    // kernel_UpdateAPX
    const double theta0 = (2 * kPi / A.nb) * bid;
    const double theta = tip * t + theta0;
    double c1[3] = {hub[0], hub[1], hub[2]};
    c1[1] += apr * cos(theta);
    c1[2] += apr * sin(theta);
    double P[3];
    to_inertial(c1, T0.yaw, nang, atype, P);
    for (int m = 0; m < 3; m++) P[m] += base[m];
    if (fused)
        for (int m = 0; m < 3; m++) sxyz[3 * ap + m] = P[m];
    const int I = find_index(X, P[0], A.nx), J = find_index(Y, P[1], A.ny), K = find_index(Z, P[2], A.nz);
    const int gc = A.gc;
    const bool in = I >= gc - 1 && I < A.nx - gc - 1 && J >= gc - 1 && J < A.ny - gc - 1 && K >= gc - 1 &&
                    K < A.nz - gc - 1;
    // kernel_CalcAPForce
    double F[3] = {0.0, 0.0, 0.0}, torque = 0.0, thrust = 0.0;
    if (in) {
        const double dr = (A.radius - A.hub_radius) / A.napb;
        const int i0 = I, i1 = I + 1, j0 = J, j1 = J + 1, k0 = K, k1 = K + 1;
        const double x0 = X[i0], x1 = X[i1], y0 = Y[j0], y1 = Y[j1], z0 = Z[k0], z1 = Z[k1];
        const double xd = fabs(P[0] - x0) / fabs(x1 - x0);
        const double yd = fabs(P[1] - y0) / fabs(y1 - y0);
        const double zd = fabs(P[2] - z0) / fabs(z1 - z0);
        const int nx = A.nx, nxy = A.nx * A.ny;
        const int id[8] = {i0 + nx * j0 + nxy * k0, i1 + nx * j0 + nxy * k0, i0 + nx * j1 + nxy * k0,
                           i1 + nx * j1 + nxy * k0, i0 + nx * j0 + nxy * k1, i1 + nx * j0 + nxy * k1,
                           i0 + nx * j1 + nxy * k1, i1 + nx * j1 + nxy * k1};
        const T *uc[3] = {u0, u1, u2};
        double uap[3];
        for (int m = 0; m < 3; m++) {
            double c[8];
            for (int q = 0; q < 8; q++) c[q] = (double)uc[m][id[q]];
            uap[m] = trilinear(xd, yd, zd, c);
        }
        double rel[3], vrel[3], utt[3];
        for (int m = 0; m < 3; m++) {
            rel[m] = P[m] - base[m];
            vrel[m] = uap[m] - bvel[m];
        }
        rot_dt(rel, vrel, ang, om, atype, utt);
        to_rotor(utt, T0.yaw);
        const double ux = utt[0];
        const double ut = tip * apr + utt[1] * sin(theta) - utt[2] * cos(theta);
        const double urel2 = ux * ux + ut * ut;
        const double phi = atan(ux / ut);
        const double att = phi * 180. / kPi - twist;
        double cl = 0.0, cd = 0.0;
        const int nat = A.nat;
        if (att < AT[0]) {
            cl = A.cl[ap];
            cd = A.cd[ap];
        } else if (att >= AT[nat - 1]) {
            cl = A.cl[ap + (nat - 1) * A.nap];
            cd = A.cd[ap + (nat - 1) * A.nap];
        } else if (att == att) {
            // the interval AT[a] <= att < AT[a+1]: the reference scans for it
            // (src/alm/apHandler.h:121-141); on an ascending table a bisection
            // finds the same one.  NaN matches no interval (cl = cd = 0).
            int a = 0, hi = nat - 1;
            while (hi - a > 1) {
                const int mid = (a + hi) >> 1;
                if (AT[mid] <= att) a = mid;
                else hi = mid;
            }
            const double p = (att - AT[a]) / (AT[a + 1] - AT[a]);
            cl = (1. - p) * A.cl[ap + a * A.nap] + p * A.cl[ap + (a + 1) * A.nap];
            cd = (1. - p) * A.cd[ap + a * A.nap] + p * A.cd[ap + (a + 1) * A.nap];
        }
        const double fl = .5 * cl * urel2 * chord * dr;
        const double fd = .5 * cd * urel2 * chord * dr;
        const double fx = fl * cos(phi) + fd * sin(phi);
        double ft = (A.flags & FALM_ALM_FIX_DRAG_TANGENT) ? fl * sin(phi) - fd * cos(phi) : fl * sin(phi) - fd * sin(phi);
        const int sg = (tip > 0) - (tip < 0);
        ft *= sg;
        const double ftt[3] = {-fx, ft * sin(theta), -ft * cos(theta)};
        to_inertial(ftt, T0.yaw, nang, atype, F);
        torque = fabs(ft) * apr;
        thrust = fx;
    }
    for (int m = 0; m < 3; m++) A.ap_xyz[3 * ap + m] = P[m];
    A.ap_ijk[3 * ap] = I;
    A.ap_ijk[3 * ap + 1] = J;
    A.ap_ijk[3 * ap + 2] = K;
    A.ap_rank[ap] = in ? 0 : -1;
    for (int m = 0; m < 3; m++) A.ap_force[3 * ap + m] = F[m];
    A.ap_torque[ap] = torque;
    A.ap_thrust[ap] = thrust;
    if (fused) {
        stq[ap] = torque;
        sth[ap] = thrust;
    }
    }();
    if (!fused) return;
    __syncthreads();
    alm_box_wg(A, sxyz, X, Y, Z);
    // k_alm_totals: per-turbine sums in ascending point order
    if ((int)threadIdx.x < A.nt) {
        const int napt = A.napb * A.nb, tid = threadIdx.x;
        double tq = 0.0, th = 0.0;
        for (int a = 0; a < napt; a++) {
            tq += stq[tid * napt + a];
            th += sth[tid * napt + a];
        }
        A.torque[tid] = tq;
        A.thrust[tid] = th;
    }
}

// The cells that can receive force: centres within cutr (plus one cell) of
// the bounding box of the points along every axis.  One workgroup.
__device__ void alm_box_wg(const AlmDev &A, const double *pxyz, const double *X, const double *Y, const double *Z) {
    __shared__ double smn[3][4], smx[3][4];
    __shared__ int slo[3][4], shi[3][4];
    __shared__ double lim[6];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int ap = threadIdx.x; ap < A.nap; ap += blockDim.x)
        for (int m = 0; m < 3; m++) {
            mn[m] = fmin(mn[m], pxyz[3 * ap + m]);
            mx[m] = fmax(mx[m], pxyz[3 * ap + m]);
        }
    for (int m = 0; m < 3; m++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[m] = fmin(mn[m], __shfl_down(mn[m], off, 64));
            mx[m] = fmax(mx[m], __shfl_down(mx[m], off, 64));
        }
        if (lane == 0) {
            smn[m][wid] = mn[m];
            smx[m][wid] = mx[m];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int m = threadIdx.x;
        double a = smn[m][0], b = smx[m][0];
        for (int w = 1; w < 4; w++) {
            a = fmin(a, smn[m][w]);
            b = fmax(b, smx[m][w]);
        }
        lim[2 * m] = a - A.cutr;
        lim[2 * m + 1] = b + A.cutr;
    }
    __syncthreads();
    const double *C[3] = {X, Y, Z};
    const int N[3] = {A.nx, A.ny, A.nz};
    int lo[3], hi[3];
    for (int m = 0; m < 3; m++) {
        lo[m] = 0x7fffffff;
        hi[m] = -1;
        for (int i = threadIdx.x; i < N[m]; i += blockDim.x) {
            const double c = C[m][i];
            if (c >= lim[2 * m] && c <= lim[2 * m + 1]) {
                lo[m] = min(lo[m], i);
                hi[m] = max(hi[m], i);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo[m] = min(lo[m], __shfl_down(lo[m], off, 64));
            hi[m] = max(hi[m], __shfl_down(hi[m], off, 64));
        }
        if (lane == 0) {
            slo[m][wid] = lo[m];
            shi[m][wid] = hi[m];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cur[6];
        bool empty = false;
        for (int m = 0; m < 3; m++) {
            int a = slo[m][0], b = shi[m][0];
            for (int w = 1; w < 4; w++) {
                a = min(a, slo[m][w]);
                b = max(b, shi[m][w]);
            }
            if (b < a) {
                empty = true;
                continue;
            }
            // one more cell on each side, within the interior
            a = max(a - 1, A.gc);
            b = min(b + 1, N[m] - A.gc - 1);
            cur[2 * m] = a;
            cur[2 * m + 1] = b;
            if (b < a) empty = true;
        }
        if (empty)
            for (int m = 0; m < 3; m++) {
                cur[2 * m] = 1;
                cur[2 * m + 1] = 0;
            }
        int *B = A.box;
        for (int q = 0; q < 6; q++) B[12 + q] = B[q];
        const bool pe = B[12] > B[13], ce = cur[0] > cur[1];
        for (int m = 0; m < 3; m++) {
            B[2 * m] = cur[2 * m];
            B[2 * m + 1] = cur[2 * m + 1];
            B[6 + 2 * m] = pe ? cur[2 * m] : ce ? B[12 + 2 * m] : min(cur[2 * m], B[12 + 2 * m]);
            B[6 + 2 * m + 1] = pe ? cur[2 * m + 1] : ce ? B[12 + 2 * m + 1] : max(cur[2 * m + 1], B[12 + 2 * m + 1]);
        }
    }
}

__global__ __launch_bounds__(256) void k_alm_box(AlmDev A) { alm_box_wg(A, A.ap_xyz, A.x, A.y, A.z); }

__device__ inline bool box_meets(const int *b, int ilo, int ihi, int jlo, int jhi, int klo, int khi) {
    return b[0] <= b[1] && ilo <= b[1] && ihi >= b[0] && jlo <= b[3] && jhi >= b[2] && klo <= b[5] && khi >= b[4];
}

constexpr int kCullChunk = 256;
constexpr int PJ_X = 32, PJ_Y = 2;  // projection tile: 32 x 2 cells of one plane, one per lane
constexpr int PJ_W = 4;             // waves per workgroup, each summing a share of the tile's points

// The Gaussian projection, walking the tiles of the cells it writes with a
// fixed grid: the union of this and the previous step's force box
// (union_only; ff is zero elsewhere and stays untouched), or every interior
// cell.  Cells outside this step's box get zero.  Per tile, the points whose
// Gaussian reaches the tile within the culling radius are compacted into
// LDS with their per-axis factors; each of the four waves sums a contiguous
// quarter of them for the tile's 64 cells, and the quarters are added in
// wave order.  A box of a few tens of thousands of cells then keeps four
// waves busy per tile instead of one.
template <class T>
__global__ __launch_bounds__(256) void k_alm_project(Box interior, AlmDev A, T *ff0, T *ff1, T *ff2, double cut2,
                                                     int union_only) {
    static_assert(PJ_X * PJ_Y == 64 && PJ_W * 64 == 256, "one tile cell per lane, four waves");
    __shared__ double sf0[kCullChunk], sf1[kCullChunk], sf2[kCullChunk], sez[kCullChunk];
    __shared__ double sey[PJ_Y][kCullChunk];
    __shared__ int sap[kCullChunk];
    __shared__ double part[PJ_W - 1][3][64];
    __shared__ int wcount[PJ_W];
    int wlo[3], wn[3];
    if (union_only) {
        const int *U = A.box + 6;
        if (U[0] > U[1]) return;
        for (int m = 0; m < 3; m++) {
            wlo[m] = U[2 * m];
            wn[m] = U[2 * m + 1] - U[2 * m] + 1;
        }
    } else {
        for (int m = 0; m < 3; m++) {
            wlo[m] = interior.lo[m];
            wn[m] = interior.n[m];
        }
    }
    const int tx = (wn[0] + PJ_X - 1) / PJ_X, ty = (wn[1] + PJ_Y - 1) / PJ_Y;
    const long long ntiles = (long long)tx * ty * wn[2];
    const int *B = A.box;
    const double eps = A.eps;
    const double eps2 = eps * eps;
    const double sp = eps * sqrt(kPi);
    const double eta = 1. / (sp * sp * sp);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int ci = lane % PJ_X, jr = lane / PJ_X;
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int bx = (int)(t % tx), by = (int)((t / tx) % ty), kc = (int)(t / ((long long)tx * ty));
        const int ilo = wlo[0] + bx * PJ_X, ihi = min(ilo + PJ_X, wlo[0] + wn[0]) - 1;
        const int jlo = wlo[1] + by * PJ_Y, jhi = min(jlo + PJ_Y, wlo[1] + wn[1]) - 1;
        const int k = wlo[2] + kc;
        const int i = ilo + ci, j = jlo + jr;
        const bool valid = i <= ihi && j <= jhi;
        const size_t cc = (size_t)i + (size_t)A.nx * j + (size_t)A.nx * A.ny * k;
        if (!box_meets(B, ilo, ihi, jlo, jhi, k, k)) {
            if (valid && wid == 0) {
                ff0[cc] = T(0);
                ff1[cc] = T(0);
                ff2[cc] = T(0);
            }
            continue;  // uniform across the workgroup
        }
        const bool inbox = valid && i >= B[0] && i <= B[1] && j >= B[2] && j <= B[3] && k >= B[4] && k <= B[5];
        // tile bounding box in coordinates (cell centres of the tile)
        const double bx0 = fmin(A.x[ilo], A.x[ihi]), bx1 = fmax(A.x[ilo], A.x[ihi]);
        const double by0 = fmin(A.y[jlo], A.y[jhi]), by1 = fmax(A.y[jlo], A.y[jhi]);
        const double bz = A.z[k];
        const int tjn = jhi - jlo + 1;
        double fx = 0., fy = 0., fz = 0.;
        for (int base = 0; base < A.nap; base += kCullChunk) {
            // keep the point unless every cell of the tile is beyond the cutoff
            const int ap = base + (int)threadIdx.x;
            bool keep = false;
            if (ap < A.nap) {
                const double px = A.ap_xyz[3 * ap], py = A.ap_xyz[3 * ap + 1], pz = A.ap_xyz[3 * ap + 2];
                const double dx = fmax(0.0, fmax(bx0 - px, px - bx1));
                const double dy = fmax(0.0, fmax(by0 - py, py - by1));
                const double dz = fabs(bz - pz);
                const double d2 = (dx * dx + dy * dy + dz * dz) / eps2;
                keep = !(d2 > cut2);
            }
            // order-preserving compaction: ballot prefix within the
            // wavefront, wavefront totals across the block
            const unsigned long long mask = __ballot(keep);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (lane == 0) wcount[wid] = __popcll(mask);
            __syncthreads();
            int off = 0, tot = 0;
            for (int w = 0; w < PJ_W; w++) {
                if (w < wid) off += wcount[w];
                tot += wcount[w];
            }
            if (keep) {
                const int slot = off + before;
                sap[slot] = ap;
                sf0[slot] = A.ap_force[3 * ap] * eta;
                sf1[slot] = A.ap_force[3 * ap + 1] * eta;
                sf2[slot] = A.ap_force[3 * ap + 2] * eta;
                sez[slot] = A.fz[(size_t)ap * A.nz + k];
                for (int q = 0; q < tjn; q++) sey[q][slot] = A.fy[(size_t)ap * A.ny + jlo + q];
            }
            __syncthreads();
            // this wave's quarter of the kept points
            const int q0 = (tot * wid) / PJ_W, q1 = (tot * (wid + 1)) / PJ_W;
            if (valid) {
                for (int q = q0; q < q1; q++) {
                    // exp(-r^2/eps^2) = ex*ey*ez (per-axis factors of the point)
                    const double e = (A.fx[(size_t)sap[q] * A.nx + i] * sey[jr][q]) * sez[q];
                    fx += sf0[q] * e;
                    fy += sf1[q] * e;
                    fz += sf2[q] * e;
                }
            }
            __syncthreads();
        }
        // the quarters in wave order
        if (wid > 0) {
            part[wid - 1][0][lane] = fx;
            part[wid - 1][1][lane] = fy;
            part[wid - 1][2][lane] = fz;
        }
        __syncthreads();
        if (wid == 0 && valid) {
            for (int w = 0; w < PJ_W - 1; w++) {
                fx += part[w][0][lane];
                fy += part[w][1][lane];
                fz += part[w][2][lane];
            }
            // cells outside the box receive nothing (every term there is
            // below the cutoff by more than a cell width)
            ff0[cc] = inbox ? (T)fx : T(0);
            ff1[cc] = inbox ? (T)fy : T(0);
            ff2[cc] = inbox ? (T)fz : T(0);
        }
        __syncthreads();
    }
}

// per-turbine sums in ascending point order; the point values are first
// gathered into LDS by the whole workgroup
__global__ void k_alm_totals(AlmDev A) {
    __shared__ double stq[1024], sth[1024];
    const int tid = blockIdx.x;
    if (tid >= A.nt) return;
    const int napt = A.napb * A.nb;
    double tq = 0.0, th = 0.0;
    for (int a0 = 0; a0 < napt; a0 += 1024) {
        const int m = napt - a0 < 1024 ? napt - a0 : 1024;
        for (int a = threadIdx.x; a < m; a += blockDim.x) {
            stq[a] = A.ap_torque[tid * napt + a0 + a];
            sth[a] = A.ap_thrust[tid * napt + a0 + a];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int a = 0; a < m; a++) {
                tq += stq[a];
                th += sth[a];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        A.torque[tid] = tq;
        A.thrust[tid] = th;
    }
}

// rotor power of the last step, accumulated: acc += sum_t Omega_t torque_t
__global__ void k_power_accum(AlmDev A, double *acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double p = 0.0;
    for (int t = 0; t < A.nt; t++) p += A.tb[t].tip_rate * A.torque[t];
    acc[0] += p;
}

// ------------------------------------------------------------- adjoint
// x = R^T y for the rotation of rot()
__device__ inline void rot_T(const double y[3], const double ang[3], int type, double x[3]) {
    if (type == 1) {
        const double C = cos(ang[0]), S = sin(ang[0]);
        x[0] = y[0]; x[1] = C * y[1] - S * y[2]; x[2] = S * y[1] + C * y[2];
    } else if (type == 2) {
        const double C = cos(ang[1]), S = sin(ang[1]);
        x[0] = C * y[0] + S * y[2]; x[1] = y[1]; x[2] = -S * y[0] + C * y[2];
    } else if (type == 3) {
        const double C = cos(ang[2]), S = sin(ang[2]);
        x[0] = C * y[0] - S * y[1]; x[1] = S * y[0] + C * y[1]; x[2] = y[2];
    } else {
        x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
    }
}

// first index i with seq[i] >= v (seq ascending, size n)
__device__ inline int lower_index(const double *seq, double v, int n) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (seq[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Adjoint of the projection with respect to the point forces: one
// workgroup per point sums ffbar(c) eta exp(-r^2/eps^2) over the cells within
// the culling radius (the cells the forward projection gives its force).
// With Mb, also the moments Mb[9 ap + 3 m + n] = sum_c ffbar_m(c) dG/dX_n of
// the same kernel G = eta exp(-|x - X|^2/eps^2) with respect to the point's
// position X (dG/dX_n = G 2 (x_n - X_n)/eps^2).
__global__ __launch_bounds__(256) void k_alm_adj_force(AlmDev A, const double *ffb0, const double *ffb1,
                                                       const double *ffb2, double *Fb, double *Mb) {
    const int ap = blockIdx.x;
    if (ap >= A.nap) return;
    __shared__ int rng[6];
    __shared__ double lds[12 * 4];
    const double X = A.ap_xyz[3 * ap], Y = A.ap_xyz[3 * ap + 1], Z = A.ap_xyz[3 * ap + 2];
    const double R = A.cutr;
    if (threadIdx.x == 0) {
        const int gc = A.gc;
        rng[0] = max(gc, lower_index(A.x, X - R, A.nx));
        rng[1] = min(A.nx - gc, lower_index(A.x, X + R, A.nx) + 1);
        rng[2] = max(gc, lower_index(A.y, Y - R, A.ny));
        rng[3] = min(A.ny - gc, lower_index(A.y, Y + R, A.ny) + 1);
        rng[4] = max(gc, lower_index(A.z, Z - R, A.nz));
        rng[5] = min(A.nz - gc, lower_index(A.z, Z + R, A.nz) + 1);
    }
    __syncthreads();
    const int ni = max(0, rng[1] - rng[0]), nj = max(0, rng[3] - rng[2]), nk = max(0, rng[5] - rng[4]);
    const long long tot = (long long)ni * nj * nk;
    const double eps = A.eps, eps2 = eps * eps;
    const double sp = eps * sqrt(kPi);
    const double eta = 1. / (sp * sp * sp);
    const int nv = Mb ? 12 : 3;
    double v[12];
    for (int q = 0; q < 12; q++) v[q] = 0.0;
    for (long long q = threadIdx.x; q < tot; q += blockDim.x) {
        const int i = rng[0] + (int)(q % ni);
        const long long r = q / ni;
        const int j = rng[2] + (int)(r % nj), k = rng[4] + (int)(r / nj);
        const double dx = A.x[i] - X, dy = A.y[j] - Y, dz = A.z[k] - Z;
        const double e = eta * exp(-(dx * dx + dy * dy + dz * dz) / eps2);
        const size_t c = (size_t)i + (size_t)A.nx * ((size_t)j + (size_t)A.ny * k);
        const double f[3] = {ffb0[c], ffb1[c], ffb2[c]};
        v[0] += f[0] * e;
        v[1] += f[1] * e;
        v[2] += f[2] * e;
        if (Mb) {
            const double w = 2.0 * e / eps2;
            const double d[3] = {dx * w, dy * w, dz * w};
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++) v[3 + 3 * m + n] += f[m] * d[n];
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int m = 0; m < nv; m++) {
        double x = v[m];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) lds[m * 4 + wid] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double x = 0.0;
        for (int w = 0; w < 4; w++) x += lds[threadIdx.x * 4 + w];
        if (threadIdx.x < 3) Fb[3 * ap + threadIdx.x] = x;
        else Mb[9 * ap + threadIdx.x - 3] = x;
    }
}

// ---------------------------------------------- attitude tangents
// A value with its derivatives with respect to the rotor's yaw (d[0]) and
// the platform angle's static part (d[1]), both in radians.
struct Tan {
    double v, d[2];
};
__device__ inline Tan tc(double v) { return {v, {0.0, 0.0}}; }
__device__ inline Tan operator+(Tan a, Tan b) { return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1]}}; }
__device__ inline Tan operator-(Tan a, Tan b) { return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1]}}; }
__device__ inline Tan operator-(Tan a) { return {-a.v, {-a.d[0], -a.d[1]}}; }
__device__ inline Tan operator*(Tan a, Tan b) {
    return {a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1]}};
}
__device__ inline Tan operator*(double a, Tan b) { return {a * b.v, {a * b.d[0], a * b.d[1]}}; }
__device__ inline Tan operator/(Tan a, Tan b) {
    const double q = a.v / b.v;
    return {q, {(a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v}};
}
__device__ inline Tan tsin(Tan a) {
    const double c = cos(a.v);
    return {sin(a.v), {c * a.d[0], c * a.d[1]}};
}
__device__ inline Tan tcos(Tan a) {
    const double s = -sin(a.v);
    return {cos(a.v), {s * a.d[0], s * a.d[1]}};
}
__device__ inline Tan tatan(Tan a) {
    const double w = 1.0 / (1.0 + a.v * a.v);
    return {atan(a.v), {w * a.d[0], w * a.d[1]}};
}

// rot / rot_dt of the forward kernels on tangent values (the same
// arithmetic; the angle and its rate may carry derivatives)
__device__ inline void rot_t(const Tan x[3], const Tan ang[3], int type, Tan out[3]) {
    if (type == 1) {
        const Tan C = tcos(ang[0]), S = tsin(ang[0]);
        out[0] = x[0]; out[1] = C * x[1] + S * x[2]; out[2] = -S * x[1] + C * x[2];
    } else if (type == 2) {
        const Tan C = tcos(ang[1]), S = tsin(ang[1]);
        out[0] = C * x[0] - S * x[2]; out[1] = x[1]; out[2] = S * x[0] + C * x[2];
    } else if (type == 3) {
        const Tan C = tcos(ang[2]), S = tsin(ang[2]);
        out[0] = C * x[0] + S * x[1]; out[1] = -S * x[0] + C * x[1]; out[2] = x[2];
    } else {
        out[0] = x[0]; out[1] = x[1]; out[2] = x[2];
    }
}
__device__ inline void rot_dt_t(const Tan x[3], const Tan v[3], const Tan ang[3], const Tan om[3], int type,
                                Tan out[3]) {
    if (type == 1) {
        const Tan C = tcos(ang[0]), S = tsin(ang[0]), O = om[0];
        out[0] = v[0];
        out[1] = O * (-S * x[1] + C * x[2]) + (C * v[1] + S * v[2]);
        out[2] = O * (-C * x[1] - S * x[2]) + (-S * v[1] + C * v[2]);
    } else if (type == 2) {
        const Tan C = tcos(ang[1]), S = tsin(ang[1]), O = om[1];
        out[0] = O * (-S * x[0] - C * x[2]) + (C * v[0] - S * v[2]);
        out[1] = v[1];
        out[2] = O * (C * x[0] - S * x[2]) + (S * v[0] + C * v[2]);
    } else if (type == 3) {
        const Tan C = tcos(ang[2]), S = tsin(ang[2]), O = om[2];
        out[0] = O * (-S * x[0] + C * x[1]) + (C * v[0] + S * v[1]);
        out[1] = O * (-C * x[0] - S * x[1]) + (-S * v[0] + C * v[1]);
        out[2] = v[2];
    } else {
        out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    }
}

// Per point: the derivatives of its position X, its force on the fluid F
// and its power Omega |ft| r with respect to the turbine's yaw and platform
// angle offset, the velocity field held (the sampling point moves within
// its cell, the polar segment and the cell are kept), contracted with this
// step's adjoints: the force adjoint Fb, the position moments Mb of
// k_alm_adj_force and the power weight wP.  Accumulates per turbine, per
// radian, into g[2 t] (yaw) and g[2 t + 1] (offset).
__global__ void k_alm_adj_attitude(AlmDev A, const double *u0, const double *u1, const double *u2, double t,
                                   const double *Fb, const double *Mb, double wP, double *g) {
    const int ap = blockIdx.x * blockDim.x + threadIdx.x;
    if (ap >= A.nap) return;
    const int napt = A.napb * A.nb;
    const int tid = ap / napt, bid = (ap % napt) / A.napb, sec = ap % A.napb;
    const TurbDev &T0 = A.tb[tid];
    const int atype = T0.angle_type;
    const Tan yaw = {T0.yaw, {1.0, 0.0}};
    Tan ang[3] = {tc(0), tc(0), tc(0)}, om[3] = {tc(0), tc(0), tc(0)};
    if (atype) {
        const double sa = T0.motion[0] * sin(T0.motion[1] * t + T0.motion[2]);
        ang[atype - 1] = {T0.aoff + sa, {0.0, 1.0}};
        om[atype - 1] = tc(T0.motion[0] * T0.motion[1] * cos(T0.motion[1] * t + T0.motion[2]));
    }
    double base[3] = {T0.base[0], T0.base[1], T0.base[2]}, bvel[3] = {T0.bvel[0], T0.bvel[1], T0.bvel[2]};
    if (T0.surge[0] != 0.0) {
        base[0] = T0.base[0] + T0.surge[0] * sin(T0.surge[1] * t + T0.surge[2]);
        bvel[0] = T0.bvel[0] + T0.surge[0] * T0.surge[1] * cos(T0.surge[1] * t + T0.surge[2]);
    }
    const Tan nang[3] = {-ang[0], -ang[1], -ang[2]};
    const double tip = T0.tip_rate;
    const double theta = tip * t + (2 * kPi / A.nb) * bid;
    const double apr = A.r[ap];
    const Tan c1[3] = {tc(T0.hub[0]), tc(T0.hub[1] + apr * cos(theta)), tc(T0.hub[2] + apr * sin(theta))};
    // to_inertial: turned by -yaw about z, then by -ang about the platform axis
    const Tan nyaw[3] = {tc(0), tc(0), -yaw};
    Tan cy[3], P[3];
    rot_t(c1, nyaw, 3, cy);
    rot_t(cy, nang, atype, P);
    for (int m = 0; m < 3; m++) P[m] = P[m] + tc(base[m]);
    const int I = find_index(A.x, P[0].v, A.nx), J = find_index(A.y, P[1].v, A.ny), K = find_index(A.z, P[2].v, A.nz);
    const int gc = A.gc;
    const bool in = I >= gc - 1 && I < A.nx - gc - 1 && J >= gc - 1 && J < A.ny - gc - 1 && K >= gc - 1 &&
                    K < A.nz - gc - 1;
    if (!in) return;  // no force, no power
    const double dr = (A.radius - A.hub_radius) / A.napb;
    const int i0 = I, j0 = J, k0 = K;
    // P lies at or above the lower cell centre, so |P - x0| = P - x0
    const Tan xd = (P[0] - tc(A.x[i0])) / tc(A.x[i0 + 1] - A.x[i0]);
    const Tan yd = (P[1] - tc(A.y[j0])) / tc(A.y[j0 + 1] - A.y[j0]);
    const Tan zd = (P[2] - tc(A.z[k0])) / tc(A.z[k0 + 1] - A.z[k0]);
    const int nx = A.nx, nxy = A.nx * A.ny;
    const int cb = i0 + nx * j0 + nxy * k0;
    const int id[8] = {cb, cb + 1, cb + nx, cb + nx + 1, cb + nxy, cb + nxy + 1, cb + nxy + nx, cb + nxy + nx + 1};
    const double *uc[3] = {u0, u1, u2};
    const Tan one = tc(1.0);
    Tan vrel[3], rel[3];
    for (int m = 0; m < 3; m++) {
        Tan c[8];
        for (int q = 0; q < 8; q++) c[q] = tc(uc[m][id[q]]);
        // trilinear_interpolation (src/alm/almDevCall.cu:28-41)
        const Tan c01 = c[0] * (one - xd) + c[1] * xd, c23 = c[2] * (one - xd) + c[3] * xd;
        const Tan c45 = c[4] * (one - xd) + c[5] * xd, c67 = c[6] * (one - xd) + c[7] * xd;
        const Tan c0123 = c01 * (one - yd) + c23 * yd, c4567 = c45 * (one - yd) + c67 * yd;
        vrel[m] = c0123 * (one - zd) + c4567 * zd - tc(bvel[m]);
        rel[m] = P[m] - tc(base[m]);
    }
    Tan up[3], utt[3];
    rot_dt_t(rel, vrel, ang, om, atype, up);
    const Tan pyaw[3] = {tc(0), tc(0), yaw};
    rot_t(up, pyaw, 3, utt);  // to_rotor
    const Tan ux = utt[0];
    const Tan ut = tc(tip * apr) + sin(theta) * utt[1] - cos(theta) * utt[2];
    const Tan urel2 = ux * ux + ut * ut;
    const Tan phi = tatan(ux / ut);
    const double twist = A.twist[ap] + T0.pitch_deg[bid < 8 ? bid : 7] + A.twist_off[sec];
    const Tan att = (180. / kPi) * phi - tc(twist);
    const int nat = A.nat;
    Tan cl = tc(0.0), cd = tc(0.0);
    if (att.v < A.attack[0]) {
        cl = tc(A.cl[ap]);
        cd = tc(A.cd[ap]);
    } else if (att.v >= A.attack[nat - 1]) {
        cl = tc(A.cl[ap + (nat - 1) * A.nap]);
        cd = tc(A.cd[ap + (nat - 1) * A.nap]);
    } else if (att.v == att.v) {
        int a = 0, hi = nat - 1;
        while (hi - a > 1) {
            const int mid = (a + hi) >> 1;
            if (A.attack[mid] <= att.v) a = mid;
            else hi = mid;
        }
        const Tan p = (att - tc(A.attack[a])) / tc(A.attack[a + 1] - A.attack[a]);
        cl = (one - p) * tc(A.cl[ap + a * A.nap]) + p * tc(A.cl[ap + (a + 1) * A.nap]);
        cd = (one - p) * tc(A.cd[ap + a * A.nap]) + p * tc(A.cd[ap + (a + 1) * A.nap]);
    }
    const double chord = A.chord[ap];
    const Tan fl = (.5 * chord * dr) * (cl * urel2), fd = (.5 * chord * dr) * (cd * urel2);
    const Tan sphi = tsin(phi), cphi = tcos(phi);
    const Tan fx = fl * cphi + fd * sphi;
    const bool fix = (A.flags & FALM_ALM_FIX_DRAG_TANGENT) != 0;
    const int sg = (tip > 0) - (tip < 0);
    const Tan ft = (double)sg * (fix ? fl * sphi - fd * cphi : fl * sphi - fd * sphi);
    const Tan ftt[3] = {-fx, sin(theta) * ft, (-cos(theta)) * ft};
    Tan fy[3], F[3];
    rot_t(ftt, nyaw, 3, fy);
    rot_t(fy, nang, atype, F);
    // contraction with the step's adjoints
    const double *M = Mb + 9 * ap;
    double Q[3];
    for (int n = 0; n < 3; n++) Q[n] = F[0].v * M[n] + F[1].v * M[3 + n] + F[2].v * M[6 + n];
    const double ps = wP * tip * apr * (double)((ft.v > 0) - (ft.v < 0));  // d(Omega |ft| r)/d ft
    for (int q = 0; q < 2; q++) {
        if (q == 1 && !atype) break;
        double gq = ps * ft.d[q];
        for (int m = 0; m < 3; m++) gq += Fb[3 * ap + m] * F[m].d[q] + Q[m] * P[m].d[q];
        unsafeAtomicAdd(&g[2 * tid + q], gq);
    }
}

__global__ void k_alm_adj_points(AlmDev A, const double *u0, const double *u1, const double *u2, double t,
                                 const double *Fb, double wP, double *ub0, double *ub1, double *ub2,
                                 double *gpitch, double *gtwist) {
    const int ap = blockIdx.x * blockDim.x + threadIdx.x;
    if (ap >= A.nap) return;
    const int napt = A.napb * A.nb;
    const int tid = ap / napt, bid = (ap % napt) / A.napb, sec = ap % A.napb;
    const TurbDev &T0 = A.tb[tid];
    double ang[3] = {0, 0, 0}, om[3] = {0, 0, 0};
    if (T0.angle_type) {
        const double sa = T0.motion[0] * sin(T0.motion[1] * t + T0.motion[2]);
        ang[T0.angle_type - 1] = T0.aoff != 0.0 ? T0.aoff + sa : sa;
        om[T0.angle_type - 1] = T0.motion[0] * T0.motion[1] * cos(T0.motion[1] * t + T0.motion[2]);
    }
    double base[3] = {T0.base[0], T0.base[1], T0.base[2]}, bvel[3] = {T0.bvel[0], T0.bvel[1], T0.bvel[2]};
    if (T0.surge[0] != 0.0) {
        base[0] = T0.base[0] + T0.surge[0] * sin(T0.surge[1] * t + T0.surge[2]);
        bvel[0] = T0.bvel[0] + T0.surge[0] * T0.surge[1] * cos(T0.surge[1] * t + T0.surge[2]);
    }
    const double nang[3] = {-ang[0], -ang[1], -ang[2]};
    const double tip = T0.tip_rate;
    const double theta = tip * t + (2 * kPi / A.nb) * bid;
    const double apr = A.r[ap];
    double c1[3] = {T0.hub[0], T0.hub[1], T0.hub[2]};
    c1[1] += apr * cos(theta);
    c1[2] += apr * sin(theta);
    double P[3];
    to_inertial(c1, T0.yaw, nang, T0.angle_type, P);
    for (int m = 0; m < 3; m++) P[m] += base[m];
    const int I = find_index(A.x, P[0], A.nx), J = find_index(A.y, P[1], A.ny), K = find_index(A.z, P[2], A.nz);
    const int gc = A.gc;
    const bool in = I >= gc - 1 && I < A.nx - gc - 1 && J >= gc - 1 && J < A.ny - gc - 1 && K >= gc - 1 &&
                    K < A.nz - gc - 1;
    if (!in) return;
    const double dr = (A.radius - A.hub_radius) / A.napb;
    const int i0 = I, i1 = I + 1, j0 = J, j1 = J + 1, k0 = K, k1 = K + 1;
    const double xd = fabs(P[0] - A.x[i0]) / fabs(A.x[i1] - A.x[i0]);
    const double yd = fabs(P[1] - A.y[j0]) / fabs(A.y[j1] - A.y[j0]);
    const double zd = fabs(P[2] - A.z[k0]) / fabs(A.z[k1] - A.z[k0]);
    const int nx = A.nx, nxy = A.nx * A.ny;
    const int id[8] = {i0 + nx * j0 + nxy * k0, i1 + nx * j0 + nxy * k0, i0 + nx * j1 + nxy * k0,
                       i1 + nx * j1 + nxy * k0, i0 + nx * j0 + nxy * k1, i1 + nx * j0 + nxy * k1,
                       i0 + nx * j1 + nxy * k1, i1 + nx * j1 + nxy * k1};
    const double wt[8] = {(1 - xd) * (1 - yd) * (1 - zd), xd * (1 - yd) * (1 - zd), (1 - xd) * yd * (1 - zd),
                          xd * yd * (1 - zd), (1 - xd) * (1 - yd) * zd, xd * (1 - yd) * zd,
                          (1 - xd) * yd * zd, xd * yd * zd};
    const double *uc[3] = {u0, u1, u2};
    double uap[3];
    for (int m = 0; m < 3; m++) {
        double c[8];
        for (int q = 0; q < 8; q++) c[q] = uc[m][id[q]];
        uap[m] = trilinear(xd, yd, zd, c);
    }
    double rel[3], vrel[3], utt[3];
    for (int m = 0; m < 3; m++) {
        rel[m] = P[m] - base[m];
        vrel[m] = uap[m] - bvel[m];
    }
    rot_dt(rel, vrel, ang, om, T0.angle_type, utt);
    to_rotor(utt, T0.yaw);
    const double ux = utt[0];
    const double ut = tip * apr + utt[1] * sin(theta) - utt[2] * cos(theta);
    const double urel2 = ux * ux + ut * ut;
    const double phi = atan(ux / ut);
    const double chord = A.chord[ap];
    const double twist = A.twist[ap] + T0.pitch_deg[bid < 8 ? bid : 7] + A.twist_off[sec];
    const double att = phi * 180. / kPi - twist;
    const int nat = A.nat;
    double cl = 0.0, cd = 0.0, scl = 0.0, scd = 0.0;
    if (att < A.attack[0]) {
        cl = A.cl[ap];
        cd = A.cd[ap];
    } else if (att >= A.attack[nat - 1]) {
        cl = A.cl[ap + (nat - 1) * A.nap];
        cd = A.cd[ap + (nat - 1) * A.nap];
    } else if (att == att) {
        int a = 0, hi = nat - 1;
        while (hi - a > 1) {
            const int mid = (a + hi) >> 1;
            if (A.attack[mid] <= att) a = mid;
            else hi = mid;
        }
        const double h = A.attack[a + 1] - A.attack[a];
        const double p = (att - A.attack[a]) / h;
        const double cl0 = A.cl[ap + a * A.nap], cl1 = A.cl[ap + (a + 1) * A.nap];
        const double cd0 = A.cd[ap + a * A.nap], cd1 = A.cd[ap + (a + 1) * A.nap];
        cl = (1. - p) * cl0 + p * cl1;
        cd = (1. - p) * cd0 + p * cd1;
        scl = (cl1 - cl0) / h;
        scd = (cd1 - cd0) / h;
    }
    const double fl = .5 * cl * urel2 * chord * dr;
    const double fd = .5 * cd * urel2 * chord * dr;
    const bool fix = (A.flags & FALM_ALM_FIX_DRAG_TANGENT) != 0;
    const double ft0 = fix ? fl * sin(phi) - fd * cos(phi) : fl * sin(phi) - fd * sin(phi);
    const int sg = (tip > 0) - (tip < 0);
    const double ft = sg * ft0;
    // F = R(nang) (-fx, ft sin(theta), -ft cos(theta))
    const double Fbv[3] = {Fb[3 * ap], Fb[3 * ap + 1], Fb[3 * ap + 2]};
    double fttb[3];
    rot_T(Fbv, nang, T0.angle_type, fttb);
    if (T0.yaw != 0.0) {  // transpose of the yaw turn in to_inertial
        const double ny[3] = {0.0, 0.0, -T0.yaw};
        const double w[3] = {fttb[0], fttb[1], fttb[2]};
        rot_T(w, ny, 3, fttb);
    }
    const double fxb = -fttb[0];
    double ftb = fttb[1] * sin(theta) - fttb[2] * cos(theta);
    ftb += wP * tip * apr * (double)((ft > 0) - (ft < 0));  // power: Omega |ft| r
    const double ft0b = sg * ftb;
    const double flb = cos(phi) * fxb + sin(phi) * ft0b;
    const double fdb = sin(phi) * fxb + (fix ? -cos(phi) : -sin(phi)) * ft0b;
    double phib = (-fl * sin(phi) + fd * cos(phi)) * fxb +
                  (fix ? (fl * cos(phi) + fd * sin(phi)) : (fl * cos(phi) - fd * cos(phi))) * ft0b;
    const double clb = .5 * urel2 * chord * dr * flb, cdb = .5 * urel2 * chord * dr * fdb;
    const double urel2b = .5 * cl * chord * dr * flb + .5 * cd * chord * dr * fdb;
    const double attb = scl * clb + scd * cdb;
    phib += attb * 180. / kPi;
    unsafeAtomicAdd(&gpitch[tid * 8 + (bid < 8 ? bid : 7)], -attb);
    unsafeAtomicAdd(&gtwist[sec], -attb);
    const double den = ux * ux + ut * ut;
    const double uxb = phib * ut / den + 2 * ux * urel2b;
    const double utb = -phib * ux / den + 2 * ut * urel2b;
    double uttb[3] = {uxb, sin(theta) * utb, -cos(theta) * utb};
    if (T0.yaw != 0.0) {  // transpose of to_rotor
        const double py[3] = {0.0, 0.0, T0.yaw};
        const double w[3] = {uttb[0], uttb[1], uttb[2]};
        rot_T(w, py, 3, uttb);
    }
    double vrelb[3];
    rot_T(uttb, ang, T0.angle_type, vrelb);
    double *ubc[3] = {ub0, ub1, ub2};
    for (int q = 0; q < 8; q++)
        for (int m = 0; m < 3; m++) unsafeAtomicAdd(&ubc[m][id[q]], wt[q] * vrelb[m]);
}

}  // namespace falm

using namespace falm;

struct falm_alm {
    falm_domain dom;
    AlmDev d;
    std::vector<void *> bufs;
    double cut2;
    int nt, nap, napb;
    bool full_next = true;  // the next owned-ff step writes every cell
    double *Fb = nullptr;   // adjoint: force adjoint per point
    double *Mb = nullptr;   // adjoint: position moments per point (9 each)
};

namespace falm {

template <class V>
static falm_status upload(falm_alm *A, const V *src, size_t n, V **dst) {
    void *p = nullptr;
    size_t bytes = (n ? n : 1) * sizeof(V);
    FALM_HIP(hipMalloc(&p, bytes));
    A->bufs.push_back(p);
    if (src && n) FALM_HIP(hipMemcpy(p, src, n * sizeof(V), hipMemcpyHostToDevice));
    else FALM_HIP(hipMemset(p, 0, bytes));
    *dst = (V *)p;
    return FALM_OK;
}

// One ALM step.  owned_ff: ff belongs to the caller's step driver and is
// only changed by these steps, so after the first full write only the
// union of this and the previous step's force box needs rewriting.
falm_status alm_step(falm_alm *A, int prec, const void *u, void *ff, double t, hipStream_t s, bool owned_ff) {
    const size_t n = ncell_of(&A->dom);
    Geom g = make_geom(&A->dom);
    Tiling tl = make_tiling(interior_box(g));
    TimedLaunch tm(FALM_TIME_ALM, s);
    const int nfac = A->nap * (A->d.nx + A->d.ny + A->d.nz);
    const size_t stage_bytes = (size_t)(A->d.nx + A->d.ny + A->d.nz + A->d.nat) * sizeof(double);
    const int staged = stage_bytes <= (size_t)kPointsLds;
    const size_t lds = staged ? stage_bytes : 0;
    const int union_only = (owned_ff && !A->full_next) ? 1 : 0;
    A->full_next = !owned_ff;
    // one workgroup holds every point: box and totals in the same launch
    const int fused = A->nap <= kFusedPoints && A->nt <= 256 ? 1 : 0;
    if (prec == FALM_F32)
        hipLaunchKernelGGL((k_alm_points<float>), dim3((A->nap + 255) / 256), dim3(256), lds, s, A->d,
                           col<float>(u, n, 0), col<float>(u, n, 1), col<float>(u, n, 2), t, staged, fused);
    else
        hipLaunchKernelGGL((k_alm_points<double>), dim3((A->nap + 255) / 256), dim3(256), lds, s, A->d,
                           col<double>(u, n, 0), col<double>(u, n, 1), col<double>(u, n, 2), t, staged, fused);
    FALM_CHECK_LAUNCH("k_alm_points");
    if (!fused) hipLaunchKernelGGL(k_alm_box, dim3(1), dim3(256), 0, s, A->d);
    hipLaunchKernelGGL(k_alm_factors, dim3((nfac + 255) / 256), dim3(256), 0, s, A->d);
    FALM_CHECK_LAUNCH("k_alm_factors");
    if (tl.tiles > 0) {
        // enough workgroups for a force box; a full write walks its tiles
        const Box ib = interior_box(g);
        const long long full = (long long)((ib.n[0] + PJ_X - 1) / PJ_X) * ((ib.n[1] + PJ_Y - 1) / PJ_Y) * ib.n[2];
        const dim3 grid((unsigned)(full < 2048 ? full : 2048));
        if (prec == FALM_F32)
            hipLaunchKernelGGL((k_alm_project<float>), grid, dim3(256), 0, s, ib, A->d, col<float>(ff, n, 0),
                               col<float>(ff, n, 1), col<float>(ff, n, 2), A->cut2, union_only);
        else
            hipLaunchKernelGGL((k_alm_project<double>), grid, dim3(256), 0, s, ib, A->d, col<double>(ff, n, 0),
                               col<double>(ff, n, 1), col<double>(ff, n, 2), A->cut2, union_only);
        FALM_CHECK_LAUNCH("k_alm_project");
    }
    if (!fused) hipLaunchKernelGGL(k_alm_totals, dim3(A->nt), dim3(256), 0, s, A->d);
    FALM_CHECK_LAUNCH("k_alm_totals");
    // bytes of the point tables: per point the 8 velocity samples, the two
    // polar columns, its state; the axis factors written once and read once
    const double pb = prec_bytes(prec);
    tm.done((double)A->nap * (24.0 * pb + 16.0 * A->d.nat + 64.0) + 16.0 * (double)nfac);
    if (tm.on) {
        // the force cells written: the interior on a full write, otherwise
        // the union box the point kernel chose (read back only when timed)
        int box[18];
        FALM_HIP(hipMemcpyAsync(box, A->d.box, sizeof(box), hipMemcpyDeviceToHost, s));
        FALM_HIP(hipStreamSynchronize(s));
        double cells = (double)tl.box.size();
        if (union_only) {
            const int *u6 = box + 6;
            cells = 1.0;
            for (int m = 0; m < 3; m++) cells *= u6[2 * m + 1] >= u6[2 * m] ? (double)(u6[2 * m + 1] - u6[2 * m] + 1) : 0.0;
        }
        tm.add_bytes(cells * 3.0 * pb);
    }
    return FALM_OK;
}

const int *alm_force_box(const falm_alm *A) { return A->d.box; }
void alm_ff_reset(falm_alm *A) { A->full_next = true; }
int alm_n_turbine(const falm_alm *A) { return A->nt; }
int alm_n_ap_per_blade(const falm_alm *A) { return A->napb; }

falm_status alm_accumulate_power(falm_alm *A, double *acc, hipStream_t s) {
    hipLaunchKernelGGL(k_power_accum, dim3(1), dim3(64), 0, s, A->d, acc);
    FALM_CHECK_LAUNCH("k_power_accum");
    return FALM_OK;
}

falm_status alm_adjoint(falm_alm *A, const double *u, double t, const double *ffbar, double wP, double *ubar,
                        double *gpitch, double *gtwist, double *gatt, hipStream_t s) {
    const size_t n = ncell_of(&A->dom);
    if (!A->Fb) {
        FALM_HIP(hipMalloc(&A->Fb, 3 * (size_t)A->nap * sizeof(double)));
        A->bufs.push_back(A->Fb);
    }
    if (gatt && !A->Mb) {
        FALM_HIP(hipMalloc(&A->Mb, 9 * (size_t)A->nap * sizeof(double)));
        A->bufs.push_back(A->Mb);
    }
    // point positions at time t (the forward's k_alm_points, which also
    // refreshes the point state)
    const size_t stage_bytes = (size_t)(A->d.nx + A->d.ny + A->d.nz + A->d.nat) * sizeof(double);
    const int staged = stage_bytes <= (size_t)kPointsLds;
    hipLaunchKernelGGL((k_alm_points<double>), dim3((A->nap + 255) / 256), dim3(256), staged ? stage_bytes : 0, s,
                       A->d, u, u + n, u + 2 * n, t, staged, 0);  // no box or totals update
    hipLaunchKernelGGL(k_alm_adj_force, dim3(A->nap), dim3(256), 0, s, A->d, ffbar, ffbar + n, ffbar + 2 * n, A->Fb,
                       gatt ? A->Mb : nullptr);
    hipLaunchKernelGGL(k_alm_adj_points, dim3((A->nap + 63) / 64), dim3(64), 0, s, A->d, u, u + n, u + 2 * n, t,
                       (const double *)A->Fb, wP, ubar, ubar + n, ubar + 2 * n, gpitch, gtwist);
    FALM_CHECK_LAUNCH("k_alm_adj_points");
    if (gatt) {
        hipLaunchKernelGGL(k_alm_adj_attitude, dim3((A->nap + 63) / 64), dim3(64), 0, s, A->d, u, u + n, u + 2 * n, t,
                           (const double *)A->Fb, (const double *)A->Mb, wP, gatt);
        FALM_CHECK_LAUNCH("k_alm_adj_attitude");
    }
    return FALM_OK;
}

}  // namespace falm

extern "C" {

falm_status falm_alm_destroy(falm_alm *A) {
    if (!A) return FALM_OK;
    for (void *p : A->bufs) (void)hipFree(p);
    delete A;
    return FALM_OK;
}

falm_status falm_alm_create(const falm_domain *dom, const falm_alm_config *c, const double *x, const double *y,
                            const double *z, falm_alm **out) {
    FALM_TRY(check_domain(dom));
    if (!c || !x || !y || !z || !out || !c->turbines || !c->ap_r || !c->ap_chord || !c->ap_twist || !c->attack ||
        !c->cl || !c->cd) {
        set_error("falm_alm_create: null argument");
        return FALM_ERR_INVALID;
    }
    if (c->n_turbine < 1 || c->n_blade < 1 || c->n_blade > 8 || c->n_ap_per_blade < 1 || c->n_attack < 2 ||
        !(c->eps > 0.0)) {
        set_error("falm_alm_create: bad sizes");
        return FALM_ERR_INVALID;
    }
    *out = nullptr;
    falm_alm *A = new falm_alm();
    A->dom = *dom;
    AlmDev &d = A->d;
    d.nt = c->n_turbine;
    d.nb = c->n_blade;
    d.napb = c->n_ap_per_blade;
    d.nap = d.nt * d.nb * d.napb;
    d.nat = c->n_attack;
    d.radius = c->radius;
    d.hub_radius = c->hub_radius;
    d.eps = c->eps;
    d.flags = c->flags;
    d.nx = dom->shape[0];
    d.ny = dom->shape[1];
    d.nz = dom->shape[2];
    d.gc = dom->gc;
    A->nt = d.nt;
    A->nap = d.nap;
    A->napb = d.napb;
    // exact mode: points whose exp(-r^2/eps^2) is 0 in double at every cell
    // of a tile are skipped; 746 > -log(smallest subnormal double) = 744.44
    A->cut2 = c->cutoff > 0.0 ? c->cutoff * c->cutoff : 746.0;
    std::vector<TurbDev> tb(d.nt);
    for (int t = 0; t < d.nt; t++) {
        const falm_turbine &T = c->turbines[t];
        for (int m = 0; m < 3; m++) {
            tb[t].base[m] = T.base[m];
            tb[t].bvel[m] = 0.0;
            tb[t].hub[m] = T.hub[m];
            tb[t].motion[m] = T.motion[m];
            tb[t].surge[m] = T.surge[m];
        }
        for (int b = 0; b < 8; b++) tb[t].pitch_deg[b] = T.pitch_deg[b];
        tb[t].yaw = T.yaw_deg * kPi / 180.0;
        tb[t].aoff = T.angle_offset_deg * kPi / 180.0;
        tb[t].tip_rate = T.tip_rate;
        tb[t].angle_type = T.angle_type;
        if (T.angle_type < 0 || T.angle_type > 3) {
            falm_alm_destroy(A);
            return FALM_ERR_INVALID;
        }
    }
    std::vector<double> toff(d.napb, 0.0);
    if (c->twist_offset) toff.assign(c->twist_offset, c->twist_offset + d.napb);
    falm_status st = FALM_OK;
    TurbDev *tbd = nullptr;
    double *dr_ = nullptr, *dch = nullptr, *dtw = nullptr, *dat = nullptr, *dcl = nullptr, *dcd = nullptr,
           *dto = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    if ((st = upload(A, tb.data(), tb.size(), &tbd)) || (st = upload(A, c->ap_r, d.nap, &dr_)) ||
        (st = upload(A, c->ap_chord, d.nap, &dch)) || (st = upload(A, c->ap_twist, d.nap, &dtw)) ||
        (st = upload(A, c->attack, d.nat, &dat)) || (st = upload(A, c->cl, (size_t)d.nat * d.nap, &dcl)) ||
        (st = upload(A, c->cd, (size_t)d.nat * d.nap, &dcd)) || (st = upload(A, toff.data(), toff.size(), &dto)) ||
        (st = upload(A, x, d.nx, &dx)) || (st = upload(A, y, d.ny, &dy)) || (st = upload(A, z, d.nz, &dz)) ||
        (st = upload<double>(A, nullptr, 3 * d.nap, &d.ap_xyz)) ||
        (st = upload<double>(A, nullptr, 3 * d.nap, &d.ap_force)) ||
        (st = upload<double>(A, nullptr, d.nap, &d.ap_torque)) ||
        (st = upload<double>(A, nullptr, d.nap, &d.ap_thrust)) ||
        (st = upload<double>(A, nullptr, d.nt, &d.torque)) || (st = upload<double>(A, nullptr, d.nt, &d.thrust)) ||
        (st = upload<int>(A, nullptr, 3 * d.nap, &d.ap_ijk)) || (st = upload<int>(A, nullptr, d.nap, &d.ap_rank)) ||
        (st = upload<double>(A, nullptr, (size_t)d.nap * d.nx, &d.fx)) ||
        (st = upload<double>(A, nullptr, (size_t)d.nap * d.ny, &d.fy)) ||
        (st = upload<double>(A, nullptr, (size_t)d.nap * d.nz, &d.fz)) ||
        (st = upload<int>(A, nullptr, 18, &d.box))) {
        falm_alm_destroy(A);
        return st;
    }
    d.tb = tbd;
    d.r = dr_;
    d.chord = dch;
    d.twist = dtw;
    d.attack = dat;
    d.cl = dcl;
    d.cd = dcd;
    d.twist_off = dto;
    d.x = dx;
    d.y = dy;
    d.z = dz;
    d.cutr = sqrt(A->cut2) * d.eps;
    const int empty[18] = {1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0};
    hipError_t e = hipMemcpy(d.box, empty, sizeof(empty), hipMemcpyHostToDevice);
    // the uploads and zero fills ran on the null stream: finish them before
    // the caller's (possibly non-blocking) stream uses the buffers
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        st = hip_status(e, "hipMemcpy");
        falm_alm_destroy(A);
        return st;
    }
    *out = A;
    return FALM_OK;
}

falm_status falm_alm_set_twist_offset(falm_alm *A, const double *off) {
    if (!A) return FALM_ERR_INVALID;
    std::vector<double> v(A->napb, 0.0);
    if (off) v.assign(off, off + A->napb);
    FALM_HIP(hipMemcpy((void *)A->d.twist_off, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    return FALM_OK;
}

falm_status falm_alm_set_pitch(falm_alm *A, int turbine, const double *pitch_deg) {
    if (!A || turbine < 0 || turbine >= A->nt) return FALM_ERR_INVALID;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (pitch_deg)
        for (int b = 0; b < 8; b++) v[b] = pitch_deg[b];
    TurbDev *dst = const_cast<TurbDev *>(A->d.tb) + turbine;
    FALM_HIP(hipMemcpy(dst->pitch_deg, v, sizeof(v), hipMemcpyHostToDevice));
    return FALM_OK;
}

falm_status falm_alm_set_attitude(falm_alm *A, int turbine, double yaw_deg, double angle_offset_deg) {
    if (!A || turbine < 0 || turbine >= A->nt || !std::isfinite(yaw_deg) || !std::isfinite(angle_offset_deg))
        return FALM_ERR_INVALID;
    const double v[2] = {yaw_deg * kPi / 180.0, angle_offset_deg * kPi / 180.0};
    TurbDev *dst = const_cast<TurbDev *>(A->d.tb) + turbine;
    static_assert(offsetof(TurbDev, aoff) == offsetof(TurbDev, yaw) + sizeof(double), "yaw, aoff adjacent");
    FALM_HIP(hipMemcpy(&dst->yaw, v, sizeof(v), hipMemcpyHostToDevice));
    return FALM_OK;
}

falm_status falm_alm_set_base(falm_alm *A, int turbine, const double *base, const double *velocity, void *stream) {
    if (!A || turbine < 0 || turbine >= A->nt || !base) return FALM_ERR_INVALID;
    double v[6] = {base[0], base[1], base[2], 0.0, 0.0, 0.0};
    if (velocity)
        for (int m = 0; m < 3; m++) v[3 + m] = velocity[m];
    // the step kernels read the turbine record: let the stream drain first
    FALM_HIP(hipStreamSynchronize((hipStream_t)stream));
    TurbDev *dst = const_cast<TurbDev *>(A->d.tb) + turbine;
    static_assert(offsetof(TurbDev, bvel) == offsetof(TurbDev, base) + 3 * sizeof(double), "base, bvel adjacent");
    FALM_HIP(hipMemcpy(dst->base, v, sizeof(v), hipMemcpyHostToDevice));
    return FALM_OK;
}

falm_status falm_alm_vjp(falm_alm *A, const void *u, double t, const void *ffbar, double wP, void *ubar,
                         void *g_pitch, void *g_twist, void *stream) {
    if (!A || !u || !ffbar || !ubar || !g_pitch || !g_twist) return FALM_ERR_INVALID;
    FALM_TRY(falm::vjp_outputs_on_device({ubar, g_pitch, g_twist}));
    return falm::alm_adjoint(A, (const double *)u, t, (const double *)ffbar, wP, (double *)ubar, (double *)g_pitch,
                             (double *)g_twist, nullptr, (hipStream_t)stream);
}

falm_status falm_alm_step(falm_alm *A, int prec, const void *u, void *ff, double t, void *stream) {
    if (!A || !u || !ff) return FALM_ERR_INVALID;
    return falm::alm_step(A, prec, u, ff, t, (hipStream_t)stream, false);
}

falm_status falm_alm_get(falm_alm *A, double *ap_xyz, double *ap_force, double *torque, double *thrust,
                         void *stream) {
    if (!A) return FALM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (ap_xyz) FALM_HIP(hipMemcpyAsync(ap_xyz, A->d.ap_xyz, 3 * A->nap * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ap_force)
        FALM_HIP(hipMemcpyAsync(ap_force, A->d.ap_force, 3 * A->nap * sizeof(double), hipMemcpyDeviceToHost, s));
    if (torque) FALM_HIP(hipMemcpyAsync(torque, A->d.torque, A->nt * sizeof(double), hipMemcpyDeviceToHost, s));
    if (thrust) FALM_HIP(hipMemcpyAsync(thrust, A->d.thrust, A->nt * sizeof(double), hipMemcpyDeviceToHost, s));
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

// BladeHandler::buildAP (src/alm/bladeHandler.h:16-104) without the JSON
// round trip; the station radius column is used as an absolute radius, as
// the reference does.
falm_status falm_build_ap(int nst, int nat, const double *rr, const double *chord, const double *twist,
                          const double *cl, const double *cd, int nt, int nb, int napb, double radius,
                          double hub, double *ap_r, double *ap_chord, double *ap_twist, double *ap_cl,
                          double *ap_cd) {
    if (nst < 1 || nat < 1 || nt < 1 || nb < 1 || napb < 1 || !rr || !chord || !twist || !cl || !cd || !ap_r ||
        !ap_chord || !ap_twist || !ap_cl || !ap_cd)
        return FALM_ERR_INVALID;
    const int nap = nt * nb * napb;
    const double drr = (radius - hub) / napb;
    for (int ap = 0; ap < nap; ap++) {
        const double r = (ap % napb + 0.5) * drr + hub;
        ap_r[ap] = r;
        int i = 0, i1 = 0;
        double p = 0.0;
        if (r < rr[0]) {
            i = i1 = 0;
        } else if (r >= rr[nst - 1]) {
            i = i1 = nst - 1;
        } else {
            for (i = 0; i < nst - 1; i++)
                if (rr[i] <= r && rr[i + 1] > r) {
                    p = (r - rr[i]) / (rr[i + 1] - rr[i]);
                    break;
                }
            i1 = i + 1;
        }
        ap_chord[ap] = i == i1 ? chord[i] : (1. - p) * chord[i] + p * chord[i1];
        ap_twist[ap] = i == i1 ? twist[i] : (1. - p) * twist[i] + p * twist[i1];
        for (int a = 0; a < nat; a++) {
            const double c0 = cl[(size_t)i * nat + a], c1 = cl[(size_t)i1 * nat + a];
            const double d0 = cd[(size_t)i * nat + a], d1 = cd[(size_t)i1 * nat + a];
            ap_cl[ap + (size_t)a * nap] = i == i1 ? c0 : (1. - p) * c0 + p * c1;
            ap_cd[ap + (size_t)a * nap] = i == i1 ? d0 : (1. - p) * d0 + p * d1;
        }
    }
    return FALM_OK;
}

}  // extern "C"
