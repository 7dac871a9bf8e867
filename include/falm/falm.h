// Falm C++ host API umbrella header: the reference's class API for the
// single-GPU time step (Matrix, Region, CPM, FalmMV, FalmEq, FalmCFD,
// Alm::AlmHandler) and its set-up and output layer (FalmCore, FalmIO,
// Mesher, Json) implemented over the falm_amd C ABI (include/falm_amd.h).
// Link with -lfalm_amd; no HIP or CUDA header is needed by the caller.
#ifndef FALM_AMD_FALM_H
#define FALM_AMD_FALM_H

#include "CPM.h"
#include "FalmCFD.h"
#include "FalmCore.h"
#include "FalmEq.h"
#include "FalmIO.h"
#include "FalmMeshBase.h"
#include "MV.h"
#include "alm.h"
#include "json.h"
#include "matrix.h"
#include "mesher.h"
#include "typedef.h"

#endif
