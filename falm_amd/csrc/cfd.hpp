// cfd.hpp — fractional-step operators shared between the C entry points and
// the step driver.
#pragma once

#include "common.hpp"

namespace falm {

// Device copies of the 1-D cell widths of a tensor-product mesh and their
// reciprocals (1.0/h computed on the host), one array per axis over the
// padded local shape.
struct SepMesh {
    const double *h[3];
    const double *rh[3];
};

// ffbox: device int[6] cell box outside which ff is zero and not read
falm_status sep_pseudo_u(const falm_domain *dom, int prec, int scheme, double ReI, double dt, const void *un,
                         const void *u, const void *uu, void *ua, const void *nut, const SepMesh &sm, const void *ff,
                         const int *ffbox, hipStream_t s);
falm_status sep_u_to_uu(const falm_domain *dom, int prec, const void *u, void *uu, const SepMesh &sm,
                        hipStream_t s);
// interior cells only: rhs = div(uu)/ja * scale (the ghost cells stay zero)
falm_status sep_poisson_rhs(const falm_domain *dom, int prec, const void *uu, double *rhs, const SepMesh &sm,
                            double scale, hipStream_t s);
// in place: u -= dt kx grad p (interior), uu -= dt g grad p (face region)
falm_status sep_project(const falm_domain *dom, int prec, double dt, void *u, void *uu, const double *p,
                        const SepMesh &sm, hipStream_t s);
falm_status sep_sgs(const falm_domain *dom, int prec, double Cs, const void *u, void *nut, const SepMesh &sm,
                    hipStream_t s);
// div = div(uu)/ja on the interior and dev_out[0] = sum div^2; partials holds
// red_blocks(interior tiling) doubles
falm_status sep_div_norm(const falm_domain *dom, int prec, const void *uu, void *div, const SepMesh &sm,
                         double *dev_out, double *partials, hipStream_t s);

}  // namespace falm
