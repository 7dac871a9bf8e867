// alm.hpp — actuator-line entry points used by the step driver.
#pragma once

#include "common.hpp"

namespace falm {

// One step; owned_ff: ff is written only by these steps (the step driver's
// own field), so cells outside the force boxes are left as they are.
falm_status alm_step(falm_alm *A, int prec, const void *u, void *ff, double t, hipStream_t s, bool owned_ff);
// Device int[6] (ilo, ihi, jlo, jhi, klo, khi) of the cells whose force may be
// non-zero after the last step (ff is exactly zero outside; empty when
// ilo > ihi).
const int *alm_force_box(const falm_alm *A);
// The next owned-ff step rewrites every cell (after ff was set from outside).
void alm_ff_reset(falm_alm *A);

}  // namespace falm
