// cfd.hpp — fractional-step operators shared between the C entry points and
// the step driver.
#pragma once

#include <vector>

#include "common.hpp"

namespace falm {

// SepMesh: device copies of the 1-D cell widths of a tensor-product mesh and
// their reciprocals (1.0/h computed on the host), one array per axis over the
// padded local shape, and per-axis coefficient tables of the step driver's
// stencils, each computed on the host in double from the widths and stored in
// double (td) and float (tf).  On a tensor-product mesh every metric product
// the schemes use
// factors into one value per axis index:
//   SEP_R   r = 1/h                            (kx along the axis)
//   SEP_H   h
//   SEP_CP  0.5 (r_i + r_{i+1}) r_i            ((g_c + g_{c+e}) / 2J)
//   SEP_CM  0.5 (r_i + r_{i-1}) r_i            ((g_c + g_{c-e}) / 2J)
//   SEP_RF  0.5 (r_i + r_{i+1})                (face mean of kx)
//   SEP_H23 cbrt(h)^2                           (ja^(2/3) = product over axes)
// (g = ja kx^2 and ja = hx hy hz, src/falm.h:383-404.)  Entries past the
// array end use the last width.
enum { SEP_R = 0, SEP_H, SEP_CP, SEP_CM, SEP_RF, SEP_H23, SEP_NTAB };

struct SepMesh {
    const double *h[3];
    const double *rh[3];
    const double *td[SEP_NTAB][3];
    const float *tf[SEP_NTAB][3];
};

// host-side values of the tables for one axis: out[tab][i], tab < SEP_NTAB
void sep_tables(const double *h, int len, std::vector<double> (&out)[SEP_NTAB]);

// ffbox: device int[6] cell box outside which ff is zero and not read
falm_status sep_pseudo_u(const falm_domain *dom, int prec, int scheme, double ReI, double dt, const void *un,
                         const void *u, const void *uu, void *ua, const void *nut, const SepMesh &sm, const void *ff,
                         const int *ffbox, hipStream_t s);
falm_status sep_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu, const SepMesh &sm,
                        hipStream_t s);
// sep_u_to_uu then sep_poisson_rhs in one pass over u (no wall-flux
// correction in between); ghost cells of u are read from gu
falm_status sep_u_to_uu_rhs(const falm_domain *dom, int prec, const void *u, const void *gu, void *uu, double *rhs,
                            const SepMesh &sm, double scale, hipStream_t s);
// interior cells only: rhs = div(uu)/ja * scale (the ghost cells stay zero)
falm_status sep_poisson_rhs(const falm_domain *dom, int prec, const void *uu, double *rhs, const SepMesh &sm,
                            double scale, hipStream_t s);
// in place: u -= dt kx grad p (interior), uu -= dt g grad p (face region)
// gate (all four calls below): the kernels do nothing unless *gate is
// nonzero; null for an unconditional launch.  host_out (the two norms): a
// pinned host copy of the norm, written by the kernel (null: none)
falm_status sep_project(const falm_domain *dom, int prec, double dt, void *u, void *uu, const double *p,
                        const SepMesh &sm, hipStream_t s, const int *gate = nullptr);
falm_status sep_sgs(const falm_domain *dom, int prec, double Cs, const void *u, void *nut, const SepMesh &sm,
                    hipStream_t s);
// div = div(uu)/ja on the interior and dev_out[0] = sum div^2; partials holds
// red_blocks(interior tiling) doubles
falm_status sep_div_norm(const falm_domain *dom, int prec, const void *uu, void *div, const SepMesh &sm,
                         double *dev_out, double *partials, hipStream_t s, const int *gate = nullptr,
                         double *host_out = nullptr);
// sep_sgs (Smagorinsky) and sep_div_norm in one pass: u and uu are read
// once per tile, nut, div and the norm are as the two calls give them
falm_status sep_sgs_div_norm(const falm_domain *dom, int prec, double Cs, const void *u, void *nut, const void *uu,
                             void *div, const SepMesh &sm, double *dev_out, double *partials, hipStream_t s,
                             const int *gate = nullptr, double *host_out = nullptr);

}  // namespace falm
