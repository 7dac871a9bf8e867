// Falm C++ host API: FalmEq (src/FalmEq.h:14-104, src/FalmEqDevCall.h:25-95)
// over the C ABI's device-resident solver.
//
// The library owns the work vectors (created on the first Solve with the
// sub-domain of the CPM passed in).  The first Solve with a given matrix
// declares it to the solver (falm_solver_set_matrix), which may keep a
// compressed copy; call resetMatrix() after changing the matrix in place.
#ifndef FALM_AMD_FALMEQ_H
#define FALM_AMD_FALMEQ_H

#include <string>

#include "MV.h"

namespace Falm {

class FalmEqDevCall {
public:
    Int maxit = 1000;
    Real tol = 1e-6;
    Int it = 0;
    Real err = 0.0;
    Real relax_factor = 1.0;
    Int pc_maxit = 5;
    Real pc_relax_factor = 1.0;
    // falm_solver_config extras: FALM_SOLVE_F64 / FALM_SOLVE_MIXED
    int precision = FALM_SOLVE_F64;

    class SolverType {
    public:
        static const Flag Empty = FALM_LS_EMPTY;
        static const Flag Jacobi = FALM_LS_JACOBI;
        static const Flag SOR = FALM_LS_SOR;
        static const Flag PBiCGStab = FALM_LS_PBICGSTAB;
    };

    Flag type = SolverType::Empty;
    Flag pc_type = SolverType::Empty;

    FalmEqDevCall() {}
    FalmEqDevCall(Flag _type, Int _maxit, Real _tol, Real _relax, Flag _pc_type = SolverType::Empty,
                  Int _pc_maxit = 5, Real _pc_relax = 1.0)
        : maxit(_maxit), tol(_tol), relax_factor(_relax), pc_maxit(_pc_maxit), pc_relax_factor(_pc_relax),
          type(_type), pc_type(_pc_type) {}
    void init(Flag _type, Int _maxit, Real _tol, Real _relax, Flag _pc_type = SolverType::Empty, Int _pc_maxit = 5,
              Real _pc_relax = 1.0) {
        type = _type;
        maxit = _maxit;
        tol = _tol;
        relax_factor = _relax;
        pc_type = _pc_type;
        pc_maxit = _pc_maxit;
        pc_relax_factor = _pc_relax;
    }
};

typedef FalmEqDevCall::SolverType LSType;

class FalmEq : public FalmEqDevCall {
public:
    falm_solver *solver = nullptr;
    const void *declared = nullptr;
    falm_domain dom_cached{};

    using FalmEqDevCall::FalmEqDevCall;
    FalmEq(const FalmEq &) = delete;
    FalmEq &operator=(const FalmEq &) = delete;
    ~FalmEq() { release(); }

    // FalmEq::alloc (src/FalmEq.h:27-40): the work vectors are created with
    // the first Solve, which knows the sub-domain
    void alloc(Int3) {}
    void release() {
        if (solver) falm_solver_destroy(solver);
        solver = nullptr;
        declared = nullptr;
    }
    void resetMatrix() { declared = nullptr; }

    void Solve(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Matrix<Real> &r, CPM &cpm, dim3 = dim3(8, 8, 8),
               Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        if (!solver || memcmp(&d, &dom_cached, sizeof(d)) != 0) {
            release();
            falm_solver_config c{};
            c.type = (int)type;
            c.maxit = maxit;
            c.tol = tol;
            c.relax = relax_factor;
            c.pc_type = (int)pc_type;
            c.pc_maxit = pc_maxit;
            c.pc_relax = pc_relax_factor;
            c.precision = precision;
            if (falmErrCheckMacro(falm_solver_create(&d, &c, &solver)) != FALM_OK) return;
            dom_cached = d;
        }
        Stream s = stream ? stream[0] : nullptr;
        if (declared != a.dev.ptr) {
            if (falmErrCheckMacro(falm_solver_set_matrix(solver, a.dev.ptr, s)) != FALM_OK) return;
            declared = a.dev.ptr;
        }
        falm_solve_info info{};
        falmErrCheckMacro(falm_solver_solve(solver, a.dev.ptr, x.dev.ptr, b.dev.ptr, r.dev.ptr, s, &info));
        it = info.it;
        err = info.err;
    }
    void PBiCGStab(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Matrix<Real> &r, CPM &cpm, dim3 bd = dim3(8, 8, 8),
                   Stream *stream = nullptr) {
        Flag t = type;
        type = LSType::PBiCGStab;
        Solve(a, x, b, r, cpm, bd, stream);
        type = t;
    }
    void SOR(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Matrix<Real> &r, CPM &cpm, dim3 bd = dim3(8, 8, 8),
             Stream *stream = nullptr) {
        Flag t = type;
        type = LSType::SOR;
        Solve(a, x, b, r, cpm, bd, stream);
        type = t;
    }

    // FalmEq::Res (src/FalmEq.cpp:45-78)
    static void Res(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Matrix<Real> &r, CPM &cpm,
                    dim3 = dim3(8, 8, 8), Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        Stream s = stream ? stream[0] : nullptr;
        falmErrCheckMacro(falm_residual(&d, FALM_F64, a.dev.ptr, x.dev.ptr, b.dev.ptr, r.dev.ptr, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }
    // FalmEqDevCall::SORSweep / JacobiSweep (src/dev/FalmEqDevCall.cu:134-221)
    void SORSweep(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Real omega, Int color, CPM &cpm,
                  Stream s = nullptr) {
        falm_domain d = cpm.domain();
        falmErrCheckMacro(falm_sor_sweep(&d, FALM_F64, a.dev.ptr, x.dev.ptr, b.dev.ptr, omega, color, s));
    }
    void JacobiSweep(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &xp, Matrix<Real> &b, CPM &cpm,
                     Stream s = nullptr) {
        falm_domain d = cpm.domain();
        falmErrCheckMacro(falm_jacobi_sweep(&d, FALM_F64, a.dev.ptr, x.dev.ptr, xp.dev.ptr, b.dev.ptr, s));
    }

    // src/FalmEq.h:79-103
    static Flag str2type(std::string str) {
        if (str == "PBiCGStab") return LSType::PBiCGStab;
        if (str == "SOR") return LSType::SOR;
        if (str == "Jacobi") return LSType::Jacobi;
        return LSType::Empty;
    }
    static std::string type2str(Flag lst) {
        if (lst == LSType::Empty) return "Empty";
        if (lst == LSType::PBiCGStab) return "PBiCGStab";
        if (lst == LSType::SOR) return "SOR";
        if (lst == LSType::Jacobi) return "Jacobi";
        return "Not defined";
    }
};

}  // namespace Falm

#endif
