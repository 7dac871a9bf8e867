// Falm C++ host API: FalmEq (src/FalmEq.h:14-104, src/FalmEqDevCall.h:25-95)
// over the C ABI's device-resident solver.
//
// The library owns the work vectors.  The reference reads type, maxit, tol,
// relax and the preconditioner settings on every Solve (src/FalmEq.h:56-64),
// so they may change between calls, through init() or the public members:
// each Solve builds the configuration from the current members and uses the
// library solver created for that configuration and sub-domain (created on
// first use, kept until release(), so alternating PBiCGStab() and SOR()
// allocates nothing after the first call of each; at most 8 are kept, the
// oldest is dropped first).  The first Solve of a
// solver with a given matrix declares it (falm_solver_set_matrix), which may
// keep a compressed copy; call resetMatrix() after changing the matrix in
// place.
#ifndef FALM_AMD_FALMEQ_H
#define FALM_AMD_FALMEQ_H

#include <string>
#include <vector>

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
    // one library solver per (configuration, sub-domain) used so far
    struct Entry {
        falm_solver_config cfg;
        falm_domain dom;
        falm_solver *solver;
        const void *declared;
    };
    std::vector<Entry> solvers;
    falm_solver *solver = nullptr;  // the one the last Solve used

    using FalmEqDevCall::FalmEqDevCall;
    FalmEq(const FalmEq &) = delete;
    FalmEq &operator=(const FalmEq &) = delete;
    ~FalmEq() { release(); }

    // FalmEq::alloc (src/FalmEq.h:27-40): the work vectors are created with
    // the first Solve, which knows the sub-domain
    void alloc(Int3) {}
    void release() {
        for (Entry &e : solvers) falm_solver_destroy(e.solver);
        solvers.clear();
        solver = nullptr;
    }
    void resetMatrix() {
        for (Entry &e : solvers) e.declared = nullptr;
    }

    // the configuration the reference's Solve would run with now
    falm_solver_config current_config() const {
        falm_solver_config c{};
        c.type = (int)type;
        c.maxit = maxit;
        c.tol = tol;
        c.relax = relax_factor;
        c.pc_type = (int)pc_type;
        c.pc_maxit = pc_maxit;
        c.pc_relax = pc_relax_factor;
        c.precision = precision;
        return c;
    }
    static bool same(const falm_solver_config &a, const falm_solver_config &b) {
        return a.type == b.type && a.maxit == b.maxit && a.tol == b.tol && a.relax == b.relax &&
               a.pc_type == b.pc_type && a.pc_maxit == b.pc_maxit && a.pc_relax == b.pc_relax &&
               a.precision == b.precision && a.inner_maxit == b.inner_maxit && a.inner_rtol == b.inner_rtol &&
               a.check_every == b.check_every;
    }
    static bool same(const falm_domain &a, const falm_domain &b) {
        for (int m = 0; m < 3; m++)
            if (a.shape[m] != b.shape[m] || a.offset[m] != b.offset[m] || a.gshape[m] != b.gshape[m]) return false;
        return a.gc == b.gc;
    }

    void Solve(Matrix<Real> &a, Matrix<Real> &x, Matrix<Real> &b, Matrix<Real> &r, CPM &cpm, dim3 = dim3(8, 8, 8),
               Stream *stream = nullptr) {
        const falm_domain d = cpm.domain();
        const falm_solver_config c = current_config();
        Entry *e = nullptr;
        for (Entry &q : solvers)
            if (same(q.cfg, c) && same(q.dom, d)) e = &q;
        if (!e) {
            if (solvers.size() >= 8) {  // bound the cache: drop the oldest
                falm_solver_destroy(solvers.front().solver);
                solvers.erase(solvers.begin());
            }
            falm_solver *sv = nullptr;
            if (falmErrCheckMacro(falm_solver_create(&d, &c, &sv)) != FALM_OK) return;
            solvers.push_back(Entry{c, d, sv, nullptr});
            e = &solvers.back();
        }
        solver = e->solver;
        Stream s = stream ? stream[0] : nullptr;
        if (e->declared != a.dev.ptr) {
            if (falmErrCheckMacro(falm_solver_set_matrix(solver, a.dev.ptr, s)) != FALM_OK) return;
            e->declared = a.dev.ptr;
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
