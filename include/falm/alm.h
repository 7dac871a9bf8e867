// Falm C++ host API: Alm::AlmHandler (src/alm/alm.h:12-140) over the C ABI.
//
// init() takes the turbine description as a falm_alm_config (the reference
// reads it from its turbine/AP JSON files, src/alm/alm.h:20-57) plus the
// 1-D cell-centre coordinates of the sub-domain.  Alm() runs the whole
// actuator-line update on the device; torque and thrust stay on the device
// until writePowerThrust() or getPowerThrust() needs them.
#ifndef FALM_AMD_ALM_H
#define FALM_AMD_ALM_H

#include <fstream>
#include <string>
#include <vector>

#include "CPM.h"

namespace Falm {
namespace Alm {

class AlmHandler {
public:
    falm_alm *handle = nullptr;
    int n_turbine = 0;
    std::vector<Real> tip_rate, torque, thrust;
    std::ofstream cpctOut;
    std::string cpctPath;

    AlmHandler() {}
    AlmHandler(const AlmHandler &) = delete;
    ~AlmHandler() { finalize(); }

    int init(const falm_alm_config &cfg, const CPM &cpm, const Real *x, const Real *y, const Real *z,
             const std::string &cpct_path = "") {
        finalize();
        falm_domain d = cpm.domain();
        int st = falmErrCheckMacro(falm_alm_create(&d, &cfg, x, y, z, &handle));
        if (st != FALM_OK) return st;
        n_turbine = cfg.n_turbine;
        tip_rate.resize(n_turbine);
        torque.assign(n_turbine, 0.0);
        thrust.assign(n_turbine, 0.0);
        for (int t = 0; t < n_turbine; t++) tip_rate[t] = cfg.turbines[t].tip_rate;
        cpctPath = cpct_path;
        if (!cpct_path.empty()) {
            // power/thrust CSV header (src/alm/alm.h:49-56)
            cpctOut.open(cpct_path);
            cpctOut << "t";
            for (int t = 0; t < n_turbine; t++) cpctOut << ",P" << t << ",T" << t;
            cpctOut << std::endl;
        }
        return FALM_OK;
    }

    void finalize() {
        if (handle) falm_alm_destroy(handle);
        handle = nullptr;
        if (cpctOut.is_open()) cpctOut.close();
    }

    // AlmHandler::Alm (src/alm/alm.h:66-77); x, y, z are already held by the
    // library (given to init)
    void Alm(Matrix<Real> &uvw, Matrix<Real> &ff, Real t, dim3 = dim3(8, 8, 8), Stream s = nullptr) {
        falmErrCheckMacro(falm_alm_step(handle, FALM_F64, uvw.dev.ptr, ff.dev.ptr, t, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }
    void Alm(Matrix<Real> &, Matrix<Real> &, Matrix<Real> &, Matrix<Real> &uvw, Matrix<Real> &ff, Real t,
             dim3 b = dim3(8, 8, 8)) {
        Alm(uvw, ff, t, b);
    }

    void getPowerThrust() {
        falmErrCheckMacro(falm_alm_get(handle, nullptr, nullptr, torque.data(), thrust.data(), nullptr));
    }

    // AlmHandler::writePowerThrust (src/alm/alm.h:126-138): P = torque*Omega
    void writePowerThrust(Real t, Real /*U*/ = 1.0) {
        if (!cpctOut.is_open()) return;
        getPowerThrust();
        cpctOut << t;
        for (int tid = 0; tid < n_turbine; tid++) cpctOut << "," << torque[tid] * tip_rate[tid] << "," << thrust[tid];
        cpctOut << std::endl;
    }
};

}  // namespace Alm
}  // namespace Falm

#endif
