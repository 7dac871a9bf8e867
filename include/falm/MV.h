// Falm C++ host API: FalmMV (src/MV.h:11-138, src/MVDevCall.h:14-65) over
// the C ABI.  Every call runs on the given stream (default: null stream);
// calls returning a scalar synchronise it.
#ifndef FALM_AMD_MV_H
#define FALM_AMD_MV_H

#include "CPM.h"

namespace Falm {

class FalmMV {
public:
    static void init() {}
    static void release() {}

    static void MV(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &ax, CPM &cpm, dim3 = dim3(8, 8, 8),
                   Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        falmErrCheckMacro(falm_mv(&d, FALM_F64, a.dev.ptr, x.dev.ptr, ax.dev.ptr, stream ? stream[0] : nullptr));
        falmErrCheckMacro(falm_stream_sync(stream ? stream[0] : nullptr));
    }
    static Real DotProduct(Matrix<Real> &a, Matrix<Real> &b, CPM &cpm, dim3 = dim3(8, 8, 8)) {
        falm_domain d = cpm.domain();
        double r = 0.0;
        falmErrCheckMacro(falm_dot(&d, FALM_F64, a.dev.ptr, b.dev.ptr, &r, nullptr));
        return r;
    }
    static Real EuclideanNormSq(Matrix<Real> &a, CPM &cpm, dim3 = dim3(8, 8, 8)) {
        falm_domain d = cpm.domain();
        double r = 0.0;
        falmErrCheckMacro(falm_norm2(&d, FALM_F64, a.dev.ptr, a.shape[1], &r, nullptr));
        return r;
    }
    static Real reduce(int op, Matrix<Real> &a, Int col, CPM &cpm) {
        falm_domain d = cpm.domain();
        double r = 0.0;
        falmErrCheckMacro(falm_reduce(&d, FALM_F64, op, a.dev.ptr, a.shape[1], col, &r, nullptr));
        return r;
    }
    static Real MaxDiag(Matrix<Real> &a, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_COL_ABSMAX, a, 0, cpm); }
    static Real MatColMax(Matrix<Real> &a, Int col, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_COL_MAX, a, col, cpm); }
    static Real MatColMin(Matrix<Real> &a, Int col, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_COL_MIN, a, col, cpm); }
    static Real MatColAbsMax(Matrix<Real> &a, Int col, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_COL_ABSMAX, a, col, cpm); }
    static Real MatColAbsMin(Matrix<Real> &a, Int col, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_COL_ABSMIN, a, col, cpm); }
    static Real VecMax(Matrix<Real> &a, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_VEC_MAX, a, 0, cpm); }
    static Real VecMin(Matrix<Real> &a, CPM &cpm, dim3 = dim3(8, 8, 8)) { return reduce(FALM_RED_VEC_MIN, a, 0, cpm); }

    static void ScaleMatrix(Matrix<Real> &a, Real scale, dim3 = dim3(8, 8, 8)) {
        falmErrCheckMacro(falm_scale(FALM_F64, a.dev.ptr, (size_t)a.size, scale, nullptr));
        falmErrCheckMacro(falm_stream_sync(nullptr));
    }
    static void MatrixAdd(Matrix<Real> &a, Matrix<Real> &b, dim3 = dim3(8, 8, 8)) {
        falmErrCheckMacro(falm_matrix_add(FALM_F64, a.dev.ptr, b.dev.ptr, (size_t)a.size, nullptr));
        falmErrCheckMacro(falm_stream_sync(nullptr));
    }
    static void Vecaxby(Real a, Matrix<Real> &x, Real b, Matrix<Real> &y, Matrix<Real> &result, dim3 = dim3(8, 8, 8)) {
        falmErrCheckMacro(falm_vecaxby(FALM_F64, a, x.dev.ptr, b, y.dev.ptr, result.dev.ptr, (size_t)x.size, nullptr));
        falmErrCheckMacro(falm_stream_sync(nullptr));
    }
};

}  // namespace Falm

#endif
