// Falm C++ host API umbrella header: the reference's class API for the
// single-GPU time step (Matrix, Region, CPM, FalmMV, FalmEq, FalmCFD,
// Alm::AlmHandler) implemented over the falm_amd C ABI (include/falm_amd.h).
// Link with -lfalm_amd; no HIP or CUDA header is needed by the caller.
#ifndef FALM_AMD_FALM_H
#define FALM_AMD_FALM_H

#include "CPM.h"
#include "FalmCFD.h"
#include "FalmEq.h"
#include "MV.h"
#include "alm.h"
#include "matrix.h"
#include "typedef.h"

#endif
