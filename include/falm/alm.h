// Falm C++ host API: the actuator-line model, Alm::AlmHandler
// (src/alm/alm.h:12-140), with the turbine and blade handling it builds on
// (src/alm/turbineHandler.h:106-213, src/alm/bladeHandler.h:13-106,
// src/alm/apHandler.h:164-200), over the C ABI.
//
// AlmHandler::init has the reference's form, init(workdir, turbine JSON,
// cpm, power/thrust CSV path): it reads the turbine property file, writes
// the actuator-point file (apFile) from its blade table, reads it back and
// sets up the turbines.  The device handle is created by the first Alm()
// call, which brings the cell-centre coordinates x, y, z.  A second init
// form takes a filled falm_alm_config and the coordinates directly.
//
// Alm() runs the whole actuator-line update on the device; torque and
// thrust stay there until writePowerThrust() or getPowerThrust() needs them.
//
// A baseLocation entry that is a string names a file of base positions over
// time: one "t value" pair per line ('#' starts a comment), interpolated
// linearly on t[q] <= t < t[q+1] and held constant before the first and
// from the last sample on; the slope of the active segment is the base
// velocity.  The reference parses such names but never reads them
// (src/alm/turbineHandler.h:144-150; its print_info lists them,
// src/alm/alm.h:96-100).
#ifndef FALM_AMD_ALM_H
#define FALM_AMD_ALM_H

#include <array>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "CPM.h"
#include "json.h"

namespace Falm {

// src/cpputil.hpp:6-14
static inline std::string glue_path(const std::string &workdir, const std::string &filename) {
    if (!filename.empty() && filename[0] == '/') return filename;
    if (!workdir.empty() && workdir.back() == '/') return workdir + filename;
    return workdir + "/" + filename;
}

static const Real Pi = 3.14159265358979323846;  // src/falmath.h:10
static inline Real deg2rad(Real deg) { return deg * Pi / 180.; }
static inline Real rad2deg(Real rad) { return rad * 180. / Pi; }

namespace Alm {

// host side of the actuator-point table (APFrame, src/alm/apHandler.h)
struct APTable {
    int n_ap = 0, n_attack = 0;
    std::vector<double> r, chord, twist, attack;
    std::vector<double> cl, cd;  // [ap + a*n_ap]
};

class BladeHandler {
public:
    // BladeHandler::buildAP (src/alm/bladeHandler.h:16-104): the blade table
    // {"airfoils": [{"r/R", "chord/R", "twist[deg]", "Cl", "Cd"}],
    // "attacks"} interpolated to apperblade points on every blade, written as
    // {"aps": [{"id", "r", "chord", "twist[deg]", "Cl", "Cd"}], "attacks"}.
    // As in the reference, the "r/R" column is compared with the absolute
    // point radius and "chord/R" is used as the chord.
    static int buildAP(const Json &bladejson, const std::string &apfilepath, int turbinecount, int bladeperturbine,
                       int apperblade, double radius, double hubradius = 0.0) {
        try {
            const Json &af = bladejson["airfoils"], &at = bladejson["attacks"];
            const int nst = (int)af.size(), nat = (int)at.size();
            if (nst < 1 || nat < 1) return fail("blade table without airfoils or attacks");
            std::vector<double> rr(nst), chord(nst), twist(nst), cl((size_t)nst * nat), cd((size_t)nst * nat),
                attack(nat);
            for (int i = 0; i < nst; i++) {
                const Json &a = af[i];
                rr[i] = a["r/R"].get<double>();
                chord[i] = a["chord/R"].get<double>();
                twist[i] = a["twist[deg]"].get<double>();
                const Json &l = a["Cl"], &d = a["Cd"];
                if ((int)l.size() < nat || (int)d.size() < nat) return fail("Cl/Cd shorter than attacks");
                for (int j = 0; j < nat; j++) {
                    cl[(size_t)i * nat + j] = l[j].get<double>();
                    cd[(size_t)i * nat + j] = d[j].get<double>();
                }
            }
            for (int j = 0; j < nat; j++) attack[j] = at[j].get<double>();
            const int nap = turbinecount * bladeperturbine * apperblade;
            std::vector<double> r(nap), c(nap), tw(nap), pcl((size_t)nap * nat), pcd((size_t)nap * nat);
            if (falmErrCheckMacro(falm_build_ap(nst, nat, rr.data(), chord.data(), twist.data(), cl.data(), cd.data(),
                                                turbinecount, bladeperturbine, apperblade, radius, hubradius,
                                                r.data(), c.data(), tw.data(), pcl.data(), pcd.data())) != FALM_OK)
                return FALM_ERR_SETUP;
            OrderedJson aps = OrderedJson::array();
            for (int ap = 0; ap < nap; ap++) {
                OrderedJson one;
                one["id"] = ap;
                one["r"] = r[ap];
                one["chord"] = c[ap];
                one["twist[deg]"] = tw[ap];
                std::vector<double> l(nat), d(nat);
                for (int j = 0; j < nat; j++) {
                    l[j] = pcl[ap + (size_t)j * nap];
                    d[j] = pcd[ap + (size_t)j * nap];
                }
                one["Cl"] = l;
                one["Cd"] = d;
                aps.push_back(one);
            }
            OrderedJson out;
            out["aps"] = aps;
            out["attacks"] = attack;
            std::ofstream f(apfilepath);
            if (!f) return fail("cannot write " + apfilepath);
            f << out.dump(2);
        } catch (const std::exception &e) {
            return fail(std::string("blade properties: ") + e.what());
        }
        return FALM_OK;
    }

    // APHandler::alloc(ap_path), host part (src/alm/apHandler.h:164-190)
    static int loadAP(const std::string &apfilepath, APTable &t) {
        std::ifstream f(apfilepath);
        if (!f) return fail("cannot open " + apfilepath);
        try {
            const Json j = Json::parse(f);
            const Json &aps = j["aps"], &at = j["attacks"];
            t.n_ap = (int)aps.size();
            t.n_attack = (int)at.size();
            const size_t nap = (size_t)t.n_ap, nat = (size_t)t.n_attack;
            t.r.resize(nap), t.chord.resize(nap), t.twist.resize(nap), t.attack.resize(nat);
            t.cl.resize(nap * nat), t.cd.resize(nap * nat);
            for (size_t ap = 0; ap < nap; ap++) {
                const Json &a = aps[ap];
                if (a["id"].get<size_t>() != ap) std::printf("AP ID ERROR: %zu != %zu\n", ap, a["id"].get<size_t>());
                t.r[ap] = a["r"].get<double>();
                t.chord[ap] = a["chord"].get<double>();
                t.twist[ap] = a["twist[deg]"].get<double>();
                for (size_t k = 0; k < nat; k++) {
                    t.cl[ap + k * nap] = a["Cl"][k].get<double>();
                    t.cd[ap + k * nap] = a["Cd"][k].get<double>();
                }
            }
            for (size_t k = 0; k < nat; k++) t.attack[k] = at[k].get<double>();
        } catch (const std::exception &e) {
            return fail(apfilepath + ": " + e.what());
        }
        return FALM_OK;
    }

    static int fail(const std::string &msg) {
        std::fprintf(stderr, "Alm: %s\n", msg.c_str());
        return lastStatus() = FALM_ERR_SETUP;
    }
};

// a base coordinate over time, read from a "t value" text file
struct BaseTrack {
    std::string path;
    std::vector<double> t, v;

    int load(const std::string &file) {
        path = file;
        std::ifstream f(file);
        if (!f) return BladeHandler::fail("cannot open base motion file " + file);
        std::string line;
        while (std::getline(f, line)) {
            const size_t hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            for (char &ch : line)
                if (ch == ',') ch = ' ';
            std::istringstream in(line);
            double a, b;
            if (in >> a >> b) {
                if (!t.empty() && a <= t.back()) return BladeHandler::fail(file + ": times must increase");
                t.push_back(a);
                v.push_back(b);
            }
        }
        if (t.empty()) return BladeHandler::fail(file + ": no samples");
        return FALM_OK;
    }
    // value and slope at time x
    void eval(double x, double &value, double &slope) const {
        slope = 0.0;
        if (x < t.front()) {
            value = v.front();
            return;
        }
        if (x >= t.back()) {
            value = v.back();
            return;
        }
        size_t q = 0;
        while (t[q + 1] <= x) q++;
        slope = (v[q + 1] - v[q]) / (t[q + 1] - t[q]);
        value = v[q] + slope * (x - t[q]);
    }
};

enum class EulerAngle { Empty, Roll, Pitch, Yaw };  // src/falmath.h:33

static inline std::string get_euler_angle_name(int a) {
    return a == 1 ? "Roll" : a == 2 ? "Pitch" : a == 3 ? "Yaw" : "Not defined";
}

// TurbineHandler::alloc (src/alm/turbineHandler.h:121-205)
struct TurbineHandler {
    int n_turbine = 0, n_blade = 0;
    Real radius = 0, hub_radius = 0;
    std::vector<falm_turbine> host;
    std::vector<std::array<BaseTrack, 3>> track;  // path empty: fixed coordinate
    Json turbine_prop, turbine_arr;

    int alloc(const Json &prop, const Json &arr, const std::string &workdir) {
        turbine_prop = prop;
        turbine_arr = arr;
        try {
            radius = prop["radius"].get<Real>();
            hub_radius = prop.contains("hubRadius") ? prop["hubRadius"].get<Real>() : 0.0;
            n_blade = prop["bladeNumber"].get<int>();
            n_turbine = (int)arr.size();
            host.assign((size_t)n_turbine, falm_turbine{});
            track.assign((size_t)n_turbine, std::array<BaseTrack, 3>{});
            for (int tid = 0; tid < n_turbine; tid++) {
                const Json &tj = arr[tid];
                falm_turbine &T = host[tid];
                T.hub[0] = -prop["overhang"].get<Real>();
                T.hub[1] = 0;
                T.hub[2] = prop["tower"].get<Real>();
                T.tip_rate = tj["tipRate"].get<Real>();
                const Json &bl = tj["baseLocation"];
                for (int i = 0; i < 3; i++) {
                    if (bl[i].is_number()) {
                        T.base[i] = bl[i].get<Real>();
                    } else if (bl[i].is_string()) {
                        int st = track[tid][i].load(glue_path(workdir, bl[i].get<std::string>()));
                        if (st != FALM_OK) return st;
                        double slope;
                        track[tid][i].eval(0.0, T.base[i], slope);
                    }
                }
                const Json &an = tj["angle"];
                int sw[3] = {0, 0, 0};
                for (int i = 0; i < 3; i++) {
                    if (an[i].is_number()) {
                        if (an[i] != 0) {
                            T.motion[0] = deg2rad(an[i].get<Real>());
                            T.motion[1] = 0.;
                            T.motion[2] = 0.5 * Pi;
                            sw[i] = 1;
                        }
                    } else {
                        const Json &m = an[i];
                        T.motion[0] = deg2rad(m["amplitude"].get<Real>());
                        T.motion[1] = 2. * Pi / m["period"].get<Real>();
                        T.motion[2] = (Pi / 180.) * m["phase"].get<Real>();
                        sw[i] = 1;
                    }
                }
                if (sw[0] + sw[1] + sw[2] > 1) {
                    std::fprintf(stderr, "Error in turbine %d, only 1 euler angle allowed.\n", tid);
                    T.angle_type = (int)EulerAngle::Empty;
                } else {
                    T.angle_type = sw[0] ? 1 : sw[1] ? 2 : sw[2] ? 3 : 0;
                }
            }
        } catch (const std::exception &e) {
            return BladeHandler::fail(std::string("turbine description: ") + e.what());
        }
        return FALM_OK;
    }

    bool moving_base(int tid) const {
        return !track[tid][0].path.empty() || !track[tid][1].path.empty() || !track[tid][2].path.empty();
    }
};

class AlmHandler {
public:
    falm_alm *handle = nullptr;
    int n_turbine = 0;
    std::vector<Real> tip_rate, torque, thrust;
    std::ofstream cpctOut;
    std::string cpctPath;

    // JSON form (reference members)
    std::string workdir;
    Json turbine_param, turbine_prop;
    Real euler_eps = 0;
    Int n_ap_per_blade = 0;
    TurbineHandler turbines;
    APTable aps;
    int rank = 0;

    AlmHandler() {}
    AlmHandler(const AlmHandler &) = delete;
    ~AlmHandler() { finalize(); }

    // AlmHandler::init (src/alm/alm.h:20-57)
    int init(const std::string &workdir_, const Json &turbine_params, const CPM &cpm, const std::string &cpct_path) {
        finalize();
        workdir = workdir_;
        cpctPath = cpct_path;
        rank = cpm.rank;
        dom_ = cpm.domain();
        turbine_param = turbine_params;
        std::string ap_path;
        try {
            euler_eps = turbine_param["projectionWidth"].get<Real>();
            n_ap_per_blade = turbine_param["bladePointNumber"].get<Int>();
            const std::string prop_path = glue_path(workdir, turbine_param["turbineProperties"].get<std::string>());
            ap_path = glue_path(workdir, turbine_param["apFile"].get<std::string>());
            std::ifstream pf(prop_path);
            if (!pf) return BladeHandler::fail("cannot open " + prop_path);
            turbine_prop = Json::parse(pf);
            int st = turbines.alloc(turbine_prop, turbine_param["turbineArray"], workdir);
            if (st != FALM_OK) return st;
            if (rank == 0) {
                st = BladeHandler::buildAP(turbine_prop["bladeProperties"], ap_path, turbines.n_turbine,
                                           turbines.n_blade, n_ap_per_blade, turbines.radius, turbines.hub_radius);
                if (st != FALM_OK) return st;
            }
            st = BladeHandler::loadAP(ap_path, aps);
            if (st != FALM_OK) return st;
        } catch (const std::exception &e) {
            return BladeHandler::fail(std::string("turbine parameters: ") + e.what());
        }
        if (aps.n_ap != turbines.n_turbine * turbines.n_blade * n_ap_per_blade)
            return BladeHandler::fail(ap_path + " does not match the turbine array");
        cfg_ = falm_alm_config{};
        cfg_.n_turbine = turbines.n_turbine;
        cfg_.n_blade = turbines.n_blade;
        cfg_.n_ap_per_blade = n_ap_per_blade;
        cfg_.n_attack = aps.n_attack;
        cfg_.radius = turbines.radius;
        cfg_.hub_radius = turbines.hub_radius;
        cfg_.eps = euler_eps;
        cfg_.turbines = turbines.host.data();
        cfg_.ap_r = aps.r.data();
        cfg_.ap_chord = aps.chord.data();
        cfg_.ap_twist = aps.twist.data();
        cfg_.attack = aps.attack.data();
        cfg_.cl = aps.cl.data();
        cfg_.cd = aps.cd.data();
        cfg_.cutoff = 0.0;  // every non-zero Gaussian term, as the all-pairs sum
        pending_ = true;
        set_turbine_arrays(cfg_.n_turbine, cfg_.turbines);
        bool write = turbine_param.contains("writePowerThrust") && turbine_param["writePowerThrust"].is_boolean() &&
                     turbine_param["writePowerThrust"].get<bool>();
        if (write && rank == 0) open_csv();
        return FALM_OK;
    }

    // falm_alm_config form: the coordinates are given now
    int init(const falm_alm_config &cfg, const CPM &cpm, const Real *x, const Real *y, const Real *z,
             const std::string &cpct_path = "") {
        finalize();
        dom_ = cpm.domain();
        rank = cpm.rank;
        int st = falmErrCheckMacro(falm_alm_create(&dom_, &cfg, x, y, z, &handle));
        if (st != FALM_OK) return st;
        set_turbine_arrays(cfg.n_turbine, cfg.turbines);
        cpctPath = cpct_path;
        if (!cpct_path.empty()) open_csv();
        return FALM_OK;
    }

    void finalize() {
        if (handle) falm_alm_destroy(handle);
        handle = nullptr;
        pending_ = false;
        if (cpctOut.is_open()) cpctOut.close();
    }

    // AlmHandler::Alm (src/alm/alm.h:66-77).  The coordinates are read on
    // the first call (host side, synchronised from the device if needed);
    // the mesh is fixed after that.
    void Alm(Matrix<Real> &x, Matrix<Real> &y, Matrix<Real> &z, Matrix<Real> &uvw, Matrix<Real> &ff, Real t,
             dim3 b = dim3(8, 8, 8)) {
        if (pending_) {
            for (Matrix<Real> *m : {&x, &y, &z})
                if (!(m->hdctype & HDC::Host)) m->sync(MCP::Dev2Hst);
            pending_ = false;
            if (falmErrCheckMacro(falm_alm_create(&dom_, &cfg_, x.host.ptr, y.host.ptr, z.host.ptr, &handle)) !=
                FALM_OK)
                return;
        }
        Alm(uvw, ff, t, b);
    }
    void Alm(Matrix<Real> &uvw, Matrix<Real> &ff, Real t, dim3 = dim3(8, 8, 8), Stream s = nullptr) {
        if (!handle) {
            std::fprintf(stderr, "AlmHandler::Alm: not initialised\n");
            lastStatus() = FALM_ERR_INVALID;
            return;
        }
        for (int tid = 0; tid < (int)turbines.track.size(); tid++) {
            if (!turbines.moving_base(tid)) continue;
            double base[3], vel[3];
            for (int m = 0; m < 3; m++) {
                base[m] = turbines.host[tid].base[m];
                vel[m] = 0.0;
                if (!turbines.track[tid][m].path.empty()) turbines.track[tid][m].eval(t, base[m], vel[m]);
            }
            falmErrCheckMacro(falm_alm_set_base(handle, tid, base, vel, s));
        }
        falmErrCheckMacro(falm_alm_step(handle, FALM_F64, uvw.dev.ptr, ff.dev.ptr, t, s));
        falmErrCheckMacro(falm_stream_sync(s));
    }

    void getPowerThrust() {
        falmErrCheckMacro(falm_alm_get(handle, nullptr, nullptr, torque.data(), thrust.data(), nullptr));
    }

    // AlmHandler::writePowerThrust (src/alm/alm.h:126-138): P = torque*Omega
    void writePowerThrust(Real t, Real /*U*/ = 1.0) {
        if (!cpctOut.is_open() || !handle) return;
        getPowerThrust();
        cpctOut << t;
        for (int tid = 0; tid < n_turbine; tid++) cpctOut << "," << torque[tid] * tip_rate[tid] << "," << thrust[tid];
        cpctOut << std::endl;
    }

    // AlmHandler::print_info (src/alm/alm.h:83-124)
    void print_info(int output_rank) {
        if (output_rank != rank) return;
        std::printf("TURBINE INFO START\n");
        std::printf("\tRadius %lf\n", turbines.radius);
        std::printf("\tHub radius %lf\n", turbines.hub_radius);
        std::printf("\tBlade number %d\n", turbines.n_blade);
        std::printf("\tBlade point number %d\n", n_ap_per_blade);
        std::printf("\tEuler projection width %lf\n", euler_eps);
        if (turbine_param.contains("turbineProperties"))
            std::printf("\tTurbine property file %s\n\n",
                        glue_path(workdir, turbine_param["turbineProperties"].get<std::string>()).c_str());
        for (int i = 0; i < turbines.n_turbine; i++) {
            const falm_turbine &T = turbines.host[i];
            std::printf("\tTurbine %d\n", i);
            std::printf("\t\tBase (%lf %lf %lf)\n", T.base[0], T.base[1], T.base[2]);
            std::printf("\t\tBase velocity (%lf %lf %lf)\n", 0.0, 0.0, 0.0);
            for (int j = 0; j < 3; j++)
                if (!turbines.track[i][j].path.empty())
                    std::printf("\t\tCoordinate[%d] subject to %s\n", j, turbines.track[i][j].path.c_str());
            std::printf("\t\tAngle type %s\n", get_euler_angle_name(T.angle_type).c_str());
            if (T.motion[1] == 0) std::printf("\t\tAngle value %lf [deg]\n", rad2deg(T.motion[0] * sin(T.motion[2])));
            else std::printf("\t\tAngle motion %lfsin(%lft + %lf) [rad]\n", T.motion[0], T.motion[1], T.motion[2]);
            std::printf("\t\tRotation center (%lf %lf %lf)\n", T.hub[0], T.hub[1], T.hub[2]);
            std::printf("\t\tRotation speed %lf\n", T.tip_rate);
        }
        if (cpctOut.is_open()) std::printf("\tPower and Thrust output to %s\n", cpctPath.c_str());
        std::printf("TURBINE INFO END\n");
    }

private:
    falm_domain dom_{};
    falm_alm_config cfg_{};
    bool pending_ = false;

    void set_turbine_arrays(int nt, const falm_turbine *tb) {
        n_turbine = nt;
        tip_rate.resize(nt);
        torque.assign(nt, 0.0);
        thrust.assign(nt, 0.0);
        for (int t = 0; t < nt; t++) tip_rate[t] = tb[t].tip_rate;
    }
    // power/thrust CSV header (src/alm/alm.h:49-56)
    void open_csv() {
        cpctOut.open(cpctPath);
        cpctOut << "t";
        for (int t = 0; t < n_turbine; t++) cpctOut << ",P" << t << ",T" << t;
        cpctOut << std::endl;
    }
};

}  // namespace Alm
}  // namespace Falm

#endif
