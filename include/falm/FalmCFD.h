// Falm C++ host API: FalmCFD (src/FalmCFD.h:11-166, src/FalmCFDDevCall.h)
// over the C ABI.  Each operation returns after its work is complete, as the
// reference's falmWaitStream() at the end of every op.
//
// FSPseudoU: the reference apps pass ua == u (app/windtunnel/main.cpp:140),
// which races in the reference kernel.  Here an aliased call first copies u
// into the object's own snapshot buffer (the role FalmCFD::uc plays in the
// reference) and computes out of place from it.
#ifndef FALM_AMD_FALMCFD_H
#define FALM_AMD_FALMCFD_H

#include <string>

#include "FalmEq.h"

namespace Falm {

class AdvectionSchemeType {
public:
    static const Flag Upwind1 = FALM_ADV_UPWIND1;
    static const Flag Upwind3 = FALM_ADV_UPWIND3;
    static const Flag QUICK = FALM_ADV_QUICK;
    static const Flag UTOPIA = FALM_ADV_UTOPIA;
    static const Flag KK = FALM_ADV_KK;
};

class SGSType {
public:
    static const Flag Empty = FALM_SGS_EMPTY;
    static const Flag Smagorinsky = FALM_SGS_SMAGORINSKY;
    static const Flag CSM = FALM_SGS_CSM;
};

class FalmCFDDevCall {
public:
    Real Re = 1.0;
    Real ReI = 1.0;
    Flag AdvScheme = AdvectionSchemeType::Upwind3;
    Flag SGSModel = SGSType::Smagorinsky;
    Real CSmagorinsky = 0.1;

    FalmCFDDevCall() {}
    FalmCFDDevCall(Real _Re, Flag _AdvScheme, Flag _SGSModel = SGSType::Smagorinsky, Real _Cs = 0.1) {
        init(_Re, _AdvScheme, _SGSModel, _Cs);
    }
    void init(Real _Re, Flag _AdvScheme, Flag _SGSModel = SGSType::Smagorinsky, Real _Cs = 0.1) {
        Re = _Re;
        ReI = 1.0 / _Re;
        AdvScheme = _AdvScheme;
        SGSModel = _SGSModel;
        CSmagorinsky = _Cs;
    }
};

class FalmCFD : public FalmCFDDevCall {
public:
    Matrix<Real> uc;  // snapshot of u for aliased FSPseudoU calls

    FalmCFD() {}
    FalmCFD(Real _Re, Flag _AdvScheme, Flag _SGSModel = SGSType::Smagorinsky, Real _Cs = 0.1)
        : FalmCFDDevCall(_Re, _AdvScheme, _SGSModel, _Cs) {}

    void alloc(Int3 shape) { uc.alloc(shape, 3, HDC::Device, "snapshot of u"); }
    void release() { uc.release(); }

    void FSPseudoU(Matrix<Real> &un, Matrix<Real> &u, Matrix<Real> &uu, Matrix<Real> &ua, Matrix<Real> &nut,
                   Matrix<Real> &kx, Matrix<Real> &g, Matrix<Real> &ja, Matrix<Real> &ff, Real dt, CPM &cpm,
                   dim3 = dim3(8, 8, 8), Stream *stream = nullptr, Int /*margin*/ = 0) {
        falm_domain d = cpm.domain();
        Stream s = stream ? stream[0] : nullptr;
        const Real *pu = u.dev.ptr, *pun = un.dev.ptr;
        if (ua.dev.ptr == u.dev.ptr || ua.dev.ptr == un.dev.ptr) {
            if (uc.size != u.size || !(uc.hdctype & HDC::Device)) uc.alloc(u.shape[0], 3, HDC::Device, "snapshot of u");
            if (ua.dev.ptr == u.dev.ptr) {
                uc.copy(u, HDC::Device);
                pu = uc.dev.ptr;
                if (un.dev.ptr == u.dev.ptr) pun = uc.dev.ptr;
            } else {
                uc.copy(un, HDC::Device);
                pun = uc.dev.ptr;
            }
        }
        falmErrCheckMacro(falm_pseudo_u(&d, FALM_F64, (int)AdvScheme, ReI, dt, pun, pu, uu.dev.ptr, ua.dev.ptr,
                                        nut.dev.ptr, kx.dev.ptr, g.dev.ptr, ja.dev.ptr, ff.dev.ptr, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }

    void UtoUU(Matrix<Real> &u, Matrix<Real> &uu, Matrix<Real> &kx, Matrix<Real> &ja, CPM &cpm, dim3 = dim3(8, 8, 8),
               Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        Stream s = stream ? stream[0] : nullptr;
        falmErrCheckMacro(falm_u_to_uu(&d, FALM_F64, u.dev.ptr, uu.dev.ptr, kx.dev.ptr, ja.dev.ptr, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }

    void ProjectP(Matrix<Real> &u, Matrix<Real> &ua, Matrix<Real> &uu, Matrix<Real> &uua, Matrix<Real> &p,
                  Matrix<Real> &kx, Matrix<Real> &g, Real dt, CPM &cpm, dim3 = dim3(8, 8, 8), Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        Stream s = stream ? stream[0] : nullptr;
        falmErrCheckMacro(falm_project_p(&d, FALM_F64, dt, u.dev.ptr, ua.dev.ptr, uu.dev.ptr, uua.dev.ptr, p.dev.ptr,
                                         kx.dev.ptr, g.dev.ptr, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }

    void SGS(Matrix<Real> &u, Matrix<Real> &nut, Matrix<Real> & /*x*/, Matrix<Real> &kx, Matrix<Real> &ja, CPM &cpm,
             dim3 = dim3(8, 8, 8), Stream *stream = nullptr) {
        falm_domain d = cpm.domain();
        Stream s = stream ? stream[0] : nullptr;
        falmErrCheckMacro(falm_sgs(&d, FALM_F64, (int)SGSModel, CSmagorinsky, u.dev.ptr, nut.dev.ptr, kx.dev.ptr,
                                   ja.dev.ptr, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }

    void Divergence(Matrix<Real> &uu, Matrix<Real> &dvr, Matrix<Real> &ja, CPM &cpm, dim3 = dim3(8, 8, 8)) {
        falm_domain d = cpm.domain();
        falmErrCheckMacro(falm_divergence(&d, FALM_F64, uu.dev.ptr, dvr.dev.ptr, ja.dev.ptr, nullptr));
        falmErrCheckMacro(falm_stream_sync(nullptr));
    }

    void MACCalcPoissonRHS(Matrix<Real> &uu, Matrix<Real> &rhs, Matrix<Real> &ja, Real dt, CPM &cpm,
                           dim3 = dim3(8, 8, 8), Real maxdiag = 1.0) {
        falm_domain d = cpm.domain();
        falmErrCheckMacro(falm_poisson_rhs(&d, FALM_F64, uu.dev.ptr, rhs.dev.ptr, ja.dev.ptr, dt, maxdiag, nullptr));
        falmErrCheckMacro(falm_stream_sync(nullptr));
    }

    // src/FalmCFD.h:111-165
    static Flag str2advscheme(std::string str) {
        if (str == "Upwind1") return AdvectionSchemeType::Upwind1;
        if (str == "Upwind3") return AdvectionSchemeType::Upwind3;
        if (str == "QUICK") return AdvectionSchemeType::QUICK;
        if (str == "UTOPIA") return AdvectionSchemeType::UTOPIA;
        if (str == "KK") return AdvectionSchemeType::KK;
        return AdvectionSchemeType::Upwind1;
    }
    static Flag str2sgs(std::string str) {
        if (str == "Empty") return SGSType::Empty;
        if (str == "Smagorinsky") return SGSType::Smagorinsky;
        if (str == "CSM") return SGSType::CSM;
        return SGSType::Empty;
    }
    static std::string advscheme2str(Flag adv) {
        if (adv == AdvectionSchemeType::Upwind1) return "Upwind1";
        if (adv == AdvectionSchemeType::Upwind3) return "Upwind3";
        if (adv == AdvectionSchemeType::QUICK) return "QUICK";
        if (adv == AdvectionSchemeType::UTOPIA) return "UTOPIA";
        if (adv == AdvectionSchemeType::KK) return "KK";
        return "Not defined";
    }
    static std::string sgs2str(Flag sgs) {
        if (sgs == SGSType::Empty) return "Empty";
        if (sgs == SGSType::Smagorinsky) return "Smagorinsky";
        if (sgs == SGSType::CSM) return "CSM";
        return "Not defined";
    }
};

}  // namespace Falm

#endif
