// Test driver for the host-side file formats of include/falm (JSON, Mesher,
// FalmIO, FalmCore::parse_settings, the JSON form of Alm::AlmHandler::init).
// Built and run by tests/test_io_formats.py on the CPU; no call here touches
// a GPU.  The json/ojson/num/mesh commands take the same arguments as
// oracle/ref_tools, the reference's own code, so outputs compare directly.
//
//   io_tool json FILE | ojson FILE | num V...
//   io_tool mesh CVCENTER MESHPATH OUTFILE GC
//   io_tool cv IN OUT                       read + write a control-volume file
//   io_tool index OUT NX NY NZ GC [STEP:TIME:TAVG]...
//   io_tool vec OUT NX NY NZ NCOL GC STEP TIME   field value = its flat index
//   io_tool setup FILE                      FalmCore::parse_settings as JSON
//   io_tool turbine WORKDIR FILE [T...]     AlmHandler::init(workdir, json) as JSON
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/falm/falm.h"

using namespace Falm;

static Json vec3(const double *v) { return Json{v[0], v[1], v[2]}; }

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if ((mode == "json" || mode == "ojson") && argc == 3) {
            std::ifstream f(argv[2]);
            if (mode == "json") std::cout << Json::parse(f).dump(2);
            else std::cout << OrderedJson::parse(f).dump(2);
            return 0;
        }
        if (mode == "num") {
            for (int a = 2; a < argc; a++) std::cout << Json(std::strtod(argv[a], nullptr)).dump() << "\n";
            return 0;
        }
        if (mode == "mesh" && argc == 6) return Mesher::build_mesh(argv[2], argv[3], argv[4], std::atoi(argv[5]));
        if (mode == "cv" && argc == 4) {
            FalmBaseMesh m;
            Int3 idmax;
            Int gc;
            if (!FalmIO::readControlVolumeFile(argv[2], m, idmax, gc)) return 1;
            FalmIO::writeControlVolumeFile(argv[3], m, idmax, gc);
            return 0;
        }
        if (mode == "index" && argc >= 7) {
            CPM cpm;
            const Int gc = std::atoi(argv[6]);
            cpm.initPartition(Int3{{std::atoi(argv[3]) + 2 * gc, std::atoi(argv[4]) + 2 * gc,
                                    std::atoi(argv[5]) + 2 * gc}},
                              gc);
            std::vector<FalmSnapshotInfo> snaps;
            for (int a = 7; a < argc; a++) {
                FalmSnapshotInfo s;
                int tavg = 0;
                unsigned long step = 0;
                if (std::sscanf(argv[a], "%lu:%lf:%d", &step, &s.time, &tavg) != 3) return 2;
                s.step = step;
                s.tavg = tavg != 0;
                snaps.push_back(s);
            }
            FalmIO::writeIndexFile(argv[2], cpm, snaps);
            return 0;
        }
        if (mode == "vec" && argc == 10) {
            CPM cpm;
            const Int3 shape{{std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])}};
            cpm.initPartition(shape, std::atoi(argv[7]));
            Matrix<Real> v(shape, std::atoi(argv[6]), HDC::Host, "field");
            for (Int q = 0; q < v.size; q++) v.host.ptr[q] = q;
            FalmIO::writeVectorFile(argv[2], cpm, v, std::strtoul(argv[8], nullptr, 10), std::strtod(argv[9], nullptr));
            FalmVectorHeader hd;
            std::vector<Real> back;
            if (!FalmIO::readVectorFile(argv[2], hd, back)) return 1;
            for (size_t q = 0; q < back.size(); q++)
                if (back[q] != (Real)q) return 1;
            return 0;
        }
        if (mode == "setup" && argc == 3) {
            FalmCore core;
            const Int st = core.parse_settings(argv[2]);
            Json j;
            j["status"] = st;
            j["workdir"] = core.workdir;
            j["setupFile"] = core.setupFile;
            j["dt"] = core.dt;
            j["startTime"] = core.startTime;
            j["maxIt"] = core.maxIt;
            j["timeAvgStartIt"] = core.timeAvgStartIt;
            j["timeAvgEndIt"] = core.timeAvgEndIt;
            j["outputStartIt"] = core.outputStartIt;
            j["outputEndIt"] = core.outputEndIt;
            j["outputIntervalIt"] = core.outputIntervalIt;
            j["outputPrefix"] = core.outputPrefix;
            j["Re"] = core.falmCfd.Re;
            j["advection"] = FalmCFD::advscheme2str(core.falmCfd.AdvScheme);
            j["sgs"] = FalmCFD::sgs2str(core.falmCfd.SGSModel);
            j["Cs"] = core.falmCfd.CSmagorinsky;
            j["solver"] = FalmEq::type2str(core.falmEq.type);
            j["maxit"] = core.falmEq.maxit;
            j["tol"] = core.falmEq.tol;
            j["relax"] = core.falmEq.relax_factor;
            j["pc"] = FalmEq::type2str(core.falmEq.pc_type);
            j["pc_maxit"] = core.falmEq.pc_maxit;
            j["pc_relax"] = core.falmEq.pc_relax_factor;
            j["cvFile"] = core.falmMeshInfo.cvFile;
            j["cvCenter"] = core.falmMeshInfo.cvCenter;
            j["convert"] = core.falmMeshInfo.convert;
            j["convertSrc"] = core.falmMeshInfo.convertSrc;
            std::cout << j.dump(2);
            return 0;
        }
        if (mode == "turbine" && argc >= 4) {
            std::ifstream f(argv[3]);
            const Json tp = Json::parse(f);
            CPM cpm;
            cpm.initPartition(Int3{{8, 8, 8}}, 2);
            Alm::AlmHandler h;
            const int st = h.init(argv[2], tp, cpm, "");
            Json j;
            j["status"] = st;
            if (st == FALM_OK) {
                j["radius"] = h.turbines.radius;
                j["hub_radius"] = h.turbines.hub_radius;
                j["n_blade"] = h.turbines.n_blade;
                j["n_ap_per_blade"] = h.n_ap_per_blade;
                j["eps"] = h.euler_eps;
                j["n_ap"] = h.aps.n_ap;
                j["n_attack"] = h.aps.n_attack;
                Json ts = Json::array();
                for (int t = 0; t < h.turbines.n_turbine; t++) {
                    const falm_turbine &T = h.turbines.host[t];
                    Json one;
                    one["base"] = vec3(T.base);
                    one["hub"] = vec3(T.hub);
                    one["tip_rate"] = T.tip_rate;
                    one["angle_type"] = T.angle_type;
                    one["motion"] = vec3(T.motion);
                    Json track = Json::array();
                    for (int a = 4; a < argc; a++) {
                        const double tt = std::strtod(argv[a], nullptr);
                        double b[3], v[3];
                        for (int m = 0; m < 3; m++) {
                            b[m] = T.base[m];
                            v[m] = 0.0;
                            if (!h.turbines.track[t][m].path.empty()) h.turbines.track[t][m].eval(tt, b[m], v[m]);
                        }
                        track.push_back(Json{tt, vec3(b), vec3(v)});
                    }
                    one["track"] = track;
                    ts.push_back(one);
                }
                j["turbines"] = ts;
            }
            std::cout << j.dump(2);
            return 0;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "io_tool: %s\n", e.what());
        return 3;
    }
    return 2;
}
