// Falm C++ host API: FalmBaseMesh (src/FalmMeshBase.h:8-37), the 1-D cell
// centres x, y, z and widths hx, hy, hz of a structured Cartesian mesh, one
// value per padded cell along each axis.
#ifndef FALM_AMD_FALMMESHBASE_H
#define FALM_AMD_FALMMESHBASE_H

#include "matrix.h"

namespace Falm {

struct FalmBaseMesh {
    Matrix<Real> x, y, z, hx, hy, hz;

    void alloc(Int imax, Int jmax, Int kmax, Flag hdctype) {
        x.alloc(imax, 1, hdctype);
        y.alloc(jmax, 1, hdctype);
        z.alloc(kmax, 1, hdctype);
        hx.alloc(imax, 1, hdctype);
        hy.alloc(jmax, 1, hdctype);
        hz.alloc(kmax, 1, hdctype);
    }
    void sync(Flag mcptype) {
        for (Matrix<Real> *m : {&x, &y, &z, &hx, &hy, &hz}) m->sync(mcptype);
    }
    void release() {
        for (Matrix<Real> *m : {&x, &y, &z, &hx, &hy, &hz}) m->release();
    }
};

}  // namespace Falm

#endif
