// Falm C++ host API: FalmIO (src/FalmIO.h:16-135), the reference's file
// formats.
//
//  - setup copy: the parsed setup JSON written back with 2-space indent;
//  - index file: JSON with "global" (interior cells), "guidePoint",
//    "ranks" [{rank, voxel, offset}], "outputSteps" [{time, step[, timeAvg]}]
//    (absent until a snapshot exists), "dataType" and "variables";
//  - vector file: a header of size_t imax, jmax, kmax, N, gc, step, then
//    double time, then size_t sizeof(Real), followed by the N columns of
//    the padded sub-domain, each with i fastest then j then k (the storage
//    order of Matrix, so the host array is written as one block);
//  - control-volume file: see mesher.h.
//
// readVectorFile has no counterpart in src/FalmIO.h; it reads the format as
// the reference's reconstructor does (src/reconstructor/reconstructor.cpp:
// 141-176).
#ifndef FALM_AMD_FALMIO_H
#define FALM_AMD_FALMIO_H

#include <cstdio>
#include <fstream>
#include <string>
#include <typeinfo>
#include <vector>

#include "CPM.h"
#include "FalmMeshBase.h"
#include "json.h"

namespace Falm {

// src/typedef.h:35-39
struct FalmSnapshotInfo {
    size_t step;
    double time;
    bool tavg;
};

struct FalmVectorHeader {
    size_t imax = 0, jmax = 0, kmax = 0, n = 0, gc = 0, step = 0;
    double time = 0.0;
    size_t dsize = 0;
};

class FalmIO {
public:
    static void writeSetupFile(const std::string &fpath, const Json &params) {
        std::ofstream f(fpath);
        f << params.dump(2);
    }

    static void writeIndexFile(const std::string &fpath, const CPM &cpm, const std::vector<FalmSnapshotInfo> &itlist) {
        const Int gc = cpm.gc;
        Json idx;
        idx["global"] = {cpm.global.shape[0] - 2 * gc, cpm.global.shape[1] - 2 * gc, cpm.global.shape[2] - 2 * gc};
        idx["guidePoint"] = gc;
        Json ranks = Json::array();
        for (int r = 0; r < cpm.size; r++) {
            const Region &pdm = cpm.pdm_list[r];
            Json one;
            one["rank"] = r;
            one["voxel"] = {pdm.shape[0] - 2 * gc, pdm.shape[1] - 2 * gc, pdm.shape[2] - 2 * gc};
            one["offset"] = {pdm.offset[0], pdm.offset[1], pdm.offset[2]};
            ranks.push_back(one);
        }
        idx["ranks"] = ranks;
        for (const FalmSnapshotInfo &s : itlist) {
            Json one;
            one["time"] = s.time;
            one["step"] = s.step;
            if (s.tavg) one["timeAvg"] = 1;
            idx["outputSteps"].push_back(one);
        }
        // the reference spells the float32 key "dateType" (src/FalmIO.h:56)
        if (typeid(Real) == typeid(float)) idx["dateType"] = "float32";
        else idx["dataType"] = "float64";
        idx["variables"] = {"u", "v", "w", "p"};
        std::ofstream f(fpath);
        f << idx.dump(2);
    }

    static void writeVectorFile(const std::string &fpath, const CPM &cpm, const Matrix<Real> &v, size_t step,
                                double time) {
        const Region &pdm = cpm.pdm_list[cpm.rank];
        FalmVectorHeader hd;
        hd.imax = (size_t)pdm.shape[0];
        hd.jmax = (size_t)pdm.shape[1];
        hd.kmax = (size_t)pdm.shape[2];
        hd.n = (size_t)v.shape[1];
        hd.gc = (size_t)cpm.gc;
        hd.step = step;
        hd.time = time;
        hd.dsize = sizeof(Real);
        if ((size_t)v.shape[0] != hd.imax * hd.jmax * hd.kmax || !v.host.ptr) {
            std::fprintf(stderr, "FalmIO::writeVectorFile: %s does not hold the sub-domain on the host\n",
                         v.name.c_str());
            return;
        }
        FILE *f = std::fopen(fpath.c_str(), "wb");
        if (!f) {
            std::fprintf(stderr, "FalmIO::writeVectorFile: cannot open %s\n", fpath.c_str());
            return;
        }
        writeHeader(f, hd);
        std::fwrite(v.host.ptr, sizeof(Real), (size_t)v.size, f);
        std::fclose(f);
    }

    // header + data of a vector file into a host vector (column-major as
    // written); false when the file is missing or short
    static bool readVectorFile(const std::string &fpath, FalmVectorHeader &hd, std::vector<Real> &data) {
        FILE *f = std::fopen(fpath.c_str(), "rb");
        if (!f) return false;
        size_t w[6];
        bool ok = std::fread(w, sizeof(size_t), 6, f) == 6 && std::fread(&hd.time, sizeof(double), 1, f) == 1 &&
                  std::fread(&hd.dsize, sizeof(size_t), 1, f) == 1 && hd.dsize == sizeof(Real);
        if (ok) {
            hd.imax = w[0], hd.jmax = w[1], hd.kmax = w[2], hd.n = w[3], hd.gc = w[4], hd.step = w[5];
            data.resize(hd.imax * hd.jmax * hd.kmax * hd.n);
            ok = std::fread(data.data(), sizeof(Real), data.size(), f) == data.size();
        }
        std::fclose(f);
        return ok;
    }

    // src/FalmIO.h:99-116; returns false when the file is missing or short
    static bool readControlVolumeFile(const std::string &srcpath, FalmBaseMesh &mesh, Int3 &idmax, Int &gc) {
        std::ifstream cv(srcpath);
        if (!(cv >> idmax[0] >> idmax[1] >> idmax[2] >> gc)) return false;
        if (idmax[0] < 1 || idmax[1] < 1 || idmax[2] < 1) return false;
        mesh.alloc(idmax[0], idmax[1], idmax[2], HDC::Host);
        Matrix<Real> *c[3] = {&mesh.x, &mesh.y, &mesh.z}, *h[3] = {&mesh.hx, &mesh.hy, &mesh.hz};
        for (int m = 0; m < 3; m++)
            for (Int i = 0; i < idmax[m]; i++)
                if (!(cv >> (*c[m])(i) >> (*h[m])(i))) return false;
        return true;
    }

    // src/FalmIO.h:118-133
    static void writeControlVolumeFile(const std::string &path, const FalmBaseMesh &mesh, const Int3 &idmax, Int gc) {
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return;
        std::fprintf(f, "%d %d %d %d\n", idmax[0], idmax[1], idmax[2], gc);
        const Matrix<Real> *c[3] = {&mesh.x, &mesh.y, &mesh.z}, *h[3] = {&mesh.hx, &mesh.hy, &mesh.hz};
        for (int m = 0; m < 3; m++)
            for (Int i = 0; i < idmax[m]; i++) std::fprintf(f, "\t%.15e\t%.15e\n", (*c[m])(i), (*h[m])(i));
        std::fclose(f);
    }

private:
    static void writeHeader(FILE *f, const FalmVectorHeader &hd) {
        const size_t w[6] = {hd.imax, hd.jmax, hd.kmax, hd.n, hd.gc, hd.step};
        std::fwrite(w, sizeof(size_t), 6, f);
        std::fwrite(&hd.time, sizeof(double), 1, f);
        std::fwrite(&hd.dsize, sizeof(size_t), 1, f);
    }
};

}  // namespace Falm

#endif
