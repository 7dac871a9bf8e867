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
// acc[0] += sum over turbines of tip_rate * torque of the last step (the
// rotor power, AlmHandler::writePowerThrust, src/alm/alm.h:126-138)
falm_status alm_accumulate_power(falm_alm *A, double *acc, hipStream_t s);
// adjoint of alm_step at (u, t) in double: ffbar is the adjoint of the
// projected force, wP the weight of this step's power in the objective;
// accumulates ubar (3 columns) and the parameter adjoints gpitch
// [n_turbine*8] and gtwist [n_ap_per_blade] (per degree), device arrays;
// gatt (may be NULL) [2*n_turbine]: per turbine the yaw and platform angle
// offset adjoints (per radian)
falm_status alm_adjoint(falm_alm *A, const double *u, double t, const double *ffbar, double wP, double *ubar,
                        double *gpitch, double *gtwist, double *gatt, hipStream_t s);
int alm_n_turbine(const falm_alm *A);
int alm_n_ap_per_blade(const falm_alm *A);

}  // namespace falm
