// Falm C++ host API: FalmCore (src/falm.h:16-483), the set-up and output
// driver of the reference's applications, over the C ABI.
//
//   parse_settings(setup.json)  runtime / solver / mesh parameters
//   computation_init(...)       Mesher conversion, control-volume file,
//                               metrics, field allocation, setup copies
//   TAvg() / outputUVWP()       running time average and snapshot files
//   env_finalize()              index file and end message
//
// One process drives one GPU and holds the whole domain (CPM single-rank).
// Setup errors are reported as FalmErr::setUpFileErr (the reference throws
// from its JSON library or asserts).
#ifndef FALM_AMD_FALMCORE_H
#define FALM_AMD_FALMCORE_H

#include <cstdio>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "FalmCFD.h"
#include "FalmEq.h"
#include "FalmIO.h"
#include "FalmMeshBase.h"
#include "MV.h"
#include "alm.h"
#include "mesher.h"

namespace Falm {

struct FalmBasicVar {
    Matrix<Real> xyz, kx, g, ja;
    Matrix<Real> u, uu, uc, p, nut;
    Matrix<Real> poi_a, poi_rhs, poi_res;
    Matrix<Real> ff;
    Matrix<Real> divergence;
    Matrix<Real> utavg, ptavg;
    Matrix<Real> uvwp;

    void release_all() {
        for (Matrix<Real> *m : {&xyz, &kx, &g, &ja, &u, &uu, &uc, &p, &nut, &poi_a, &poi_rhs, &poi_res, &ff,
                                &divergence, &utavg, &ptavg, &uvwp})
            m->release();
    }
};

struct FalmMeshInfo {
    bool convert = false;
    std::string convertSrc;
    std::string cvFile;
    std::string cvCenter;
};

class FalmCore {
public:
    Json params;
    FalmCFD falmCfd;
    FalmEq falmEq;
    FalmMeshInfo falmMeshInfo;
    CPM cpm;
    FalmBasicVar fv;
    FalmBaseMesh gBaseMesh;
    FalmBaseMesh baseMesh;

    Int it = 0;
    Real dt = 0;
    Real startTime = 0;
    Int maxIt = 0;
    Int timeAvgStartIt = 0;
    Int timeAvgEndIt = -1;
    Int timeAvgCount = 0;
    Int outputStartIt = 0;
    Int outputEndIt = -1;
    Int outputIntervalIt = 1;
    std::string workdir = ".";
    std::string outputPrefix;
    std::string setupFile;
    std::string cvFile;
    std::vector<FalmSnapshotInfo> timeSlices;
    int ngpu = 1;
    int gpuid = 0;

    Real gettime() { return dt * it; }

    bool is_time_averaging() { return it >= timeAvgStartIt && it <= timeAvgEndIt; }

    // FalmCore::env_init (src/falm.h:87-94): device 0 of this process
    int env_init(int & /*argc*/, char **& /*argv*/) {
        timeSlices.clear();
        gpuid = 0;
        return falmErrCheckMacro(falm_init(gpuid));
    }

    // FalmCore::env_finalize (src/falm.h:96-113)
    void env_finalize() {
        if (cpm.rank == 0 && !outputPrefix.empty())
            FalmIO::writeIndexFile(wpath(outputPrefix + ".json"), cpm, timeSlices);
        fv.release_all();
        falmEq.release();
        falmCfd.release();
        gBaseMesh.release();
        baseMesh.release();
        // written relative to the current directory, as the reference does
        std::ofstream mfs(outputPrefix + "_endmsg.txt");
        mfs << "Falm/Main ends normally.";
    }

    std::string wpath(const std::string &str) { return glue_path(workdir, str); }

    // FalmCore::print_info (src/falm.h:125-220), single rank
    void print_info(int output_rank = 0) {
        if (cpm.rank != output_rank) return;
        std::printf("SETUP INFO START\n");
        std::printf("Working dir %s\n", workdir.c_str());
        std::printf("Read setup file %s\n", wpath(setupFile).c_str());
        std::printf("Control Params:\n");
        std::printf("\tdt %e\n", dt);
        std::printf("\tstart time %e\n", startTime);
        std::printf("\titeration number %d\n", maxIt);
        std::printf("\ttime avg start iteration %d\n", timeAvgStartIt);
        std::printf("\ttime avg end iteration %d\n", timeAvgEndIt);
        std::printf("\toutput start iteration %d\n", outputStartIt);
        std::printf("\toutput end iteration %d\n", outputEndIt);
        std::printf("\toutput every %d iterations\n", outputIntervalIt);
        std::printf("\toutput to %s\n", wpath(outputPrefix).c_str());
        std::printf("CFD Solver Params\n");
        std::printf("\tRe = %e, 1/Re = %e\n", falmCfd.Re, falmCfd.ReI);
        std::printf("\tAdvection scheme %s\n", FalmCFD::advscheme2str(falmCfd.AdvScheme).c_str());
        std::printf("\tSGS model %s\n", FalmCFD::sgs2str(falmCfd.SGSModel).c_str());
        if (falmCfd.SGSModel == SGSType::Smagorinsky) std::printf("\tCsmagorinsky %e\n", falmCfd.CSmagorinsky);
        std::printf("Poisson Solver Params\n");
        std::printf("\tsolver type %s\n", FalmEq::type2str(falmEq.type).c_str());
        std::printf("\tmax iteration number %d\n", falmEq.maxit);
        std::printf("\ttolerance %e\n", falmEq.tol);
        if (falmEq.type == LSType::SOR) std::printf("\trelaxation factor %e\n", falmEq.relax_factor);
        if (falmEq.type == LSType::PBiCGStab) {
            std::printf("Preconditioner Params\n");
            std::printf("\tsolver type %s\n", FalmEq::type2str(falmEq.pc_type).c_str());
            std::printf("\titeration number %d\n", falmEq.pc_maxit);
            if (falmEq.pc_type == LSType::SOR) std::printf("\trelaxation factor %e\n", falmEq.pc_relax_factor);
        }
        if (params.contains("inflow")) {
            const Json &in = params["inflow"];
            std::printf("Inflow Params\n");
            if (in.contains("type") && in["type"].is_string())
                std::printf("\tProfile %s\n", in["type"].get<std::string>().c_str());
            if (in.contains("velocity") && in["velocity"].is_number())
                std::printf("\tVelocity %e\n", in["velocity"].get<Real>());
        }
        std::printf("SETUP INFO END\n");
        const Int gc = cpm.gc;
        const Region &pdm = cpm.pdm_list[cpm.rank];
        std::printf("MPI INFO START\n");
        std::printf("\tGlobal voxel (%d %d %d)\n", cpm.global.shape[0], cpm.global.shape[1], cpm.global.shape[2]);
        std::printf("\tGlobal inner voxel (%d %d %d)\n", cpm.global.shape[0] - 2 * gc, cpm.global.shape[1] - 2 * gc,
                    cpm.global.shape[2] - 2 * gc);
        std::printf("\tGuide voxel length %d\n", gc);
        std::printf("\tMPI cluster shape (%d %d %d)\n", cpm.shape[0], cpm.shape[1], cpm.shape[2]);
        std::printf("\tRank 0\n");
        std::printf("\t\tVoxel (%d %d %d)\n", pdm.shape[0], pdm.shape[1], pdm.shape[2]);
        std::printf("\t\tInner voxel (%d %d %d)\n", pdm.shape[0] - 2 * gc, pdm.shape[1] - 2 * gc,
                    pdm.shape[2] - 2 * gc);
        std::printf("\t\tOffset (%d %d %d)\n", pdm.offset[0], pdm.offset[1], pdm.offset[2]);
        std::printf("MPI INFO END\n");
        std::printf("GPU INFO START\n\t1 GPU in this process\nGPU INFO END\n");
    }

    // FalmCore::parse_settings (src/falm.h:222-340)
    Int parse_settings(const std::string &setup_file_path) {
        const size_t cut = setup_file_path.find_last_of('/');
        if (cut == std::string::npos) {
            workdir = ".";
            setupFile = setup_file_path;
        } else if (cut == 0) {
            workdir = "/";
            setupFile = setup_file_path.substr(1);
        } else {
            workdir = setup_file_path.substr(0, cut);
            setupFile = setup_file_path.substr(cut + 1);
        }
        std::ifstream f(wpath(setupFile));
        if (!f) return FalmErr::setUpFileErr;
        try {
            params = Json::parse(f);
            const Json &rt = params["runtime"];
            startTime = rt["time"]["start"].get<Real>();
            dt = rt["time"]["dt"].get<Real>();
            maxIt = Int((rt["time"]["end"].get<Real>() - startTime) / dt);
            if (rt.contains("timeAvg")) {
                const Json &ta = rt["timeAvg"];
                if (ta.contains("start")) {
                    Real s = ta["start"].get<Real>();
                    if (s < startTime) s = startTime;
                    timeAvgStartIt = Int((s - startTime) / dt);
                } else {
                    timeAvgStartIt = 0;
                }
                timeAvgEndIt = ta.contains("end") ? Int((ta["end"].get<Real>() - startTime) / dt) : maxIt;
            } else {
                timeAvgStartIt = 0;
                timeAvgEndIt = -1;
            }
            if (rt.contains("output")) {
                const Json &o = rt["output"];
                if (o.contains("start")) {
                    Real s = o["start"].get<Real>();
                    if (s < startTime) s = startTime;
                    outputStartIt = Int((s - startTime) / dt);
                } else {
                    outputStartIt = 0;
                }
                outputEndIt = o.contains("end") ? Int((o["end"].get<Real>() - startTime) / dt) : maxIt;
                outputIntervalIt = Int(o["interval"].get<Real>() / dt);
                outputPrefix = o["prefix"].get<std::string>();
            } else {
                outputStartIt = 0;
                outputEndIt = -1;
            }

            const Json &ls = params["solver"]["linearSolver"];
            FalmEq tmp;
            tmp.type = FalmEq::str2type(ls["type"].get<std::string>());
            tmp.maxit = ls["iteration"].get<Int>();
            tmp.tol = ls["tolerance"].get<Real>();
            if (tmp.type == LSType::SOR) tmp.relax_factor = ls["relaxationFactor"].get<Real>();
            if (tmp.type == LSType::PBiCGStab) {
                const Json &pc = ls["preconditioner"];
                tmp.pc_type = FalmEq::str2type(pc["type"].get<std::string>());
                tmp.pc_maxit = pc["iteration"].get<Int>();
                if (tmp.pc_type == LSType::SOR) tmp.pc_relax_factor = pc["relaxationFactor"].get<Real>();
            } else {
                tmp.pc_type = LSType::Empty;
            }
            falmEq.init(tmp.type, tmp.maxit, tmp.tol, tmp.relax_factor, tmp.pc_type, tmp.pc_maxit,
                        tmp.pc_relax_factor);

            const Json &cfd = params["solver"]["cfd"];
            const Real Re = cfd["Re"].get<Real>();
            const Flag adv = FalmCFD::str2advscheme(cfd["advectionScheme"].get<std::string>());
            const Flag sgs = cfd.contains("SGS") ? FalmCFD::str2sgs(cfd["SGS"].get<std::string>()) : SGSType::Empty;
            const Real cs = sgs == SGSType::Smagorinsky ? cfd["Cs"].get<Real>() : 0.1;
            falmCfd.init(Re, adv, sgs, cs);

            const Json &mesh = params["mesh"];
            falmMeshInfo.cvFile = mesh["controlVolumeFile"].get<std::string>();
            falmMeshInfo.cvCenter = mesh["controlVolumeCenter"].get<std::string>();
            if (mesh.contains("convert")) {
                falmMeshInfo.convert = true;
                falmMeshInfo.convertSrc = mesh["convert"].get<std::string>();
            }
        } catch (const std::exception &e) {
            std::fprintf(stderr, "FalmCore::parse_settings: %s: %s\n", wpath(setupFile).c_str(), e.what());
            return lastStatus() = FalmErr::setUpFileErr;
        }
        cvFile = falmMeshInfo.cvFile;
        return FalmErr::success;
    }

    // FalmCore::computation_init (src/falm.h:342-431): control volumes ->
    // partition -> case description files -> local 1-D mesh -> metrics ->
    // fields.
    int computation_init(const Int3 &division, int gc) {
        Int3 idmax;
        int st = load_control_volumes(gc, idmax);
        if (st != FALM_OK) return st;
        st = cpm.initPartition(idmax, gc, division);
        if (st != FALM_OK) return st;
        const Region &local = cpm.pdm_list[cpm.rank];
        if (cpm.rank == 0 && outputEndIt >= outputStartIt) write_case_description(idmax, gc);
        std::printf("(%d %d %d) (%d %d %d)\n", local.shape[0], local.shape[1], local.shape[2], local.offset[0],
                    local.offset[1], local.offset[2]);
        slice_local_mesh(local);
        st = derive_metrics(local.shape);
        if (st != FALM_OK) return st;
        allocate_fields(local.shape);
        falmCfd.alloc(local.shape);
        falmEq.alloc(local.shape);
        return lastStatus() == FALM_OK ? FALM_OK : lastStatus();
    }

    // FalmCore::TAvg (src/falm.h:433-444): the running mean over the
    // averaging window, mean_n = mean_{n-1} + (v - mean_{n-1})/n written as
    // (1 - 1/n) mean_{n-1} + (1/n) v with one Vecaxby per field
    void TAvg(dim3 block = dim3(8, 8, 8)) {
        if (!is_time_averaging()) return;
        timeAvgCount = it - timeAvgStartIt + 1;
        const Real w_new = 1. / timeAvgCount, w_old = 1. - w_new;
        const std::pair<Matrix<Real> *, Matrix<Real> *> avg_of[2] = {{&fv.utavg, &fv.u}, {&fv.ptavg, &fv.p}};
        for (const auto &pr : avg_of) FalmMV::Vecaxby(w_old, *pr.first, w_new, *pr.second, *pr.first, block);
    }

    // FalmCore::outputUVWP (src/falm.h:446-479): at output steps, the
    // snapshot PREFIX_RRRRRR_IIIIIIIIII of (u, v, w, p) and, inside the
    // averaging window, PREFIX_tavg_... of the running means; the step is
    // recorded for the index file.
    void outputUVWP(dim3 /*block*/ = dim3(8, 8, 8)) {
        const bool due = it >= outputStartIt && it <= outputEndIt && outputIntervalIt > 0 && it % outputIntervalIt == 0;
        if (!due) return;
        write_uvwp(outputPrefix, fv.u, fv.p);
        const bool averaged = is_time_averaging();
        if (averaged) write_uvwp(outputPrefix + "_tavg", fv.utavg, fv.ptavg);
        timeSlices.push_back(FalmSnapshotInfo{(size_t)it, gettime(), averaged});
    }

private:
    // Mesher conversion (rank 0, when the setup asks for it) and the
    // control-volume file; the file's guide-cell count must match gc
    int load_control_volumes(int gc, Int3 &idmax) {
        const std::string cv = wpath(falmMeshInfo.cvFile);
        if (falmMeshInfo.convert && cpm.rank == 0 &&
            Mesher::build_mesh(falmMeshInfo.cvCenter, wpath(falmMeshInfo.convertSrc), cv, gc) != 0)
            return lastStatus() = FalmErr::setUpFileErr;
        Int file_gc = 0;
        if (!FalmIO::readControlVolumeFile(cv, gBaseMesh, idmax, file_gc)) {
            std::fprintf(stderr, "FalmCore: cannot read control-volume file %s\n", cv.c_str());
            return lastStatus() = FalmErr::setUpFileErr;
        }
        if (file_gc != gc) {
            std::fprintf(stderr, "FalmCore: control-volume file has %d guide cells, expected %d\n", file_gc, gc);
            return lastStatus() = FalmErr::setUpFileErr;
        }
        return FALM_OK;
    }

    // index file, copy of the control volumes and of the setup next to the
    // snapshots
    void write_case_description(const Int3 &idmax, int gc) {
        FalmIO::writeIndexFile(wpath(outputPrefix + ".json"), cpm, timeSlices);
        FalmIO::writeControlVolumeFile(wpath(outputPrefix + ".cv"), gBaseMesh, idmax, gc);
        FalmIO::writeSetupFile(wpath(outputPrefix + "_setup.json"), params);
    }

    // this sub-domain's window of the global 1-D centres and widths
    void slice_local_mesh(const Region &local) {
        baseMesh.alloc(local.shape[0], local.shape[1], local.shape[2], HDC::HstDev);
        Matrix<Real> *lc[3] = {&baseMesh.x, &baseMesh.y, &baseMesh.z}, *lw[3] = {&baseMesh.hx, &baseMesh.hy,
                                                                                   &baseMesh.hz};
        Matrix<Real> *gcen[3] = {&gBaseMesh.x, &gBaseMesh.y, &gBaseMesh.z},
                     *gwid[3] = {&gBaseMesh.hx, &gBaseMesh.hy, &gBaseMesh.hz};
        for (int m = 0; m < 3; m++)
            for (Int q = 0; q < local.shape[m]; q++) {
                (*lc[m])(q) = (*gcen[m])(q + local.offset[m]);
                (*lw[m])(q) = (*gwid[m])(q + local.offset[m]);
            }
        baseMesh.sync(MCP::Hst2Dev);
    }

    // metric fields from the 1-D mesh on the device (falm_metrics: volume
    // hx hy hz, kx = 1/h, g = volume kx^2, cell centres), with host copies
    // for the application's host-side matrix assembly
    int derive_metrics(const Int3 &shape) {
        fv.xyz.alloc(shape, 3, HDC::HstDev, "xyz");
        fv.kx.alloc(shape, 3, HDC::HstDev, "kx");
        fv.g.alloc(shape, 3, HDC::HstDev, "g");
        fv.ja.alloc(shape, 1, HDC::HstDev, "ja");
        const falm_domain d = cpm.domain();
        const int st = falmErrCheckMacro(falm_metrics(
            &d, FALM_F64, baseMesh.x.host.ptr, baseMesh.y.host.ptr, baseMesh.z.host.ptr, baseMesh.hx.host.ptr,
            baseMesh.hy.host.ptr, baseMesh.hz.host.ptr, fv.xyz.dev.ptr, fv.kx.dev.ptr, fv.g.dev.ptr, fv.ja.dev.ptr,
            nullptr));
        if (st != FALM_OK) return st;
        falmErrCheckMacro(falm_stream_sync(nullptr));
        for (Matrix<Real> *m : {&fv.xyz, &fv.kx, &fv.g, &fv.ja}) m->sync(MCP::Dev2Hst);
        return FALM_OK;
    }

    // the solution, solver and output fields (zero-filled).  The running
    // means exist whenever the averaging window is not empty: the reference
    // allocates them only for end > start (src/falm.h:421) yet averages at
    // it == start == end too
    void allocate_fields(const Int3 &shape) {
        struct Spec {
            Matrix<Real> *m;
            Int ncol;
            Flag where;
            const char *label;
        };
        const Spec specs[] = {{&fv.u, 3, HDC::HstDev, "u"},          {&fv.uu, 3, HDC::HstDev, "uu"},
                              {&fv.p, 1, HDC::HstDev, "p"},          {&fv.nut, 1, HDC::HstDev, "nut"},
                              {&fv.poi_rhs, 1, HDC::HstDev, "poi_rhs"}, {&fv.poi_res, 1, HDC::HstDev, "poi_res"},
                              {&fv.ff, 3, HDC::HstDev, "ff"},        {&fv.divergence, 1, HDC::HstDev, "divergence"},
                              {&fv.uvwp, 4, HDC::HstDev, "uvwp"}};
        for (const Spec &sp : specs) sp.m->alloc(shape, sp.ncol, sp.where, sp.label);
        fv.poi_a.alloc(shape, 7, HDC::HstDev, "poi_a", StencilMatrix::D3P7);
        if (timeAvgEndIt >= timeAvgStartIt) {
            fv.utavg.alloc(shape, 3, HDC::Device, "utavg");
            fv.ptavg.alloc(shape, 1, HDC::Device, "ptavg");
            fv.utavg.clear(HDC::Device);
            fv.ptavg.clear(HDC::Device);
        }
    }

    // stage (velocity, pressure) of the device fields in the host columns of
    // uvwp and write them as one vector file named for this rank and step
    void write_uvwp(const std::string &prefix, Matrix<Real> &vel, Matrix<Real> &pres) {
        const size_t n = (size_t)fv.uvwp.shape[0];
        falmErrCheckMacro(falm_memcpy(&fv.uvwp(0, 0), vel.dev.ptr, sizeof(Real) * n * 3, MCP::Dev2Hst, nullptr));
        falmErrCheckMacro(falm_memcpy(&fv.uvwp(0, 3), pres.dev.ptr, sizeof(Real) * n, MCP::Dev2Hst, nullptr));
        FalmIO::writeVectorFile(wpath(snapshot_name(prefix)), cpm, fv.uvwp, (size_t)it, gettime());
    }

    std::string snapshot_name(const std::string &prefix) {
        std::vector<char> buf(prefix.size() + 32);
        std::snprintf(buf.data(), buf.size(), "%s_%06d_%010d", prefix.c_str(), cpm.rank, it);
        return std::string(buf.data());
    }
};

}  // namespace Falm

#endif
