// Mesher (src/mesher/mesher.hpp:9-262): a node list to a control-volume
// file.  Host code behind falm_mesher_build.
//
// Input: a directory holding mesh.json ({"coordinatePoint": "node"}) and
// x.txt, y.txt, z.txt, each a node count on the first line followed by one
// coordinate per line.  Output: "imax jmax kmax gc" then one
// "\t<centre>\t<width>" line per padded cell of x, then y, then z, printed
// with %.15e (the control-volume file of src/FalmIO.h:99-133).
//
// cvcenter "cell": the nodes are cell faces; n nodes give n-1 cells plus gc
// ghost cells per side whose widths continue the first/last two widths
// linearly (or repeat the single width when there is one cell).
// cvcenter "innerNode": the nodes are cell centres and the outermost node
// on each side is itself the first ghost cell, so n nodes give n-2 interior
// cells plus gc ghost cells per side; a width is half the distance between
// the neighbouring nodes and the ghost widths continue linearly.  The upper
// ghost widths along z repeat the last width instead
// (src/mesher/mesher.hpp:84 subtracts hz[k-1] where the x and y loops
// subtract h[i-2]); this is kept so that control-volume files agree.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/falm/json.h"
#include "../../include/falm_amd.h"

namespace falm {
void set_error(const char *fmt, ...);
}

namespace {

using Falm::Json;

struct Mesher {
    static int build_mesh(const std::string &cvcenter, const std::string &meshpath, const std::string &outputfile,
                          int gc) {
        std::ifstream mj(meshpath + "/mesh.json");
        if (!mj) return fail("cannot open " + meshpath + "/mesh.json");
        std::string point;
        try {
            point = Json::parse(mj)["coordinatePoint"].get<std::string>();
        } catch (const std::exception &e) {
            return fail(meshpath + "/mesh.json: " + e.what());
        }
        if (point != "node") return fail("mesh.json: coordinatePoint \"" + point + "\" is not supported");
        if (gc < 1) return fail("guide cell count must be at least 1");

        std::vector<double> c[3], h[3];
        const char *axis[3] = {"x", "y", "z"};
        for (int m = 0; m < 3; m++) {
            std::vector<double> nodes;
            const std::string path = meshpath + "/" + axis[m] + ".txt";
            if (!read_nodes(path, nodes)) return fail("cannot read node list " + path);
            int st;
            if (cvcenter == "innerNode") st = inner_node(nodes, gc, m == 2, c[m], h[m]);
            else if (cvcenter == "cell") st = cell_faces(nodes, gc, c[m], h[m]);
            else return fail("controlVolumeCenter \"" + cvcenter + "\" is not supported");
            if (st) return fail(path + ": too few nodes for " + cvcenter);
        }

        FILE *f = std::fopen(outputfile.c_str(), "w");
        if (!f) return fail("cannot write " + outputfile);
        std::fprintf(f, "%d %d %d %d\n", (int)c[0].size(), (int)c[1].size(), (int)c[2].size(), gc);
        for (int m = 0; m < 3; m++)
            for (size_t i = 0; i < c[m].size(); i++) std::fprintf(f, "\t%.15e\t%.15e\n", c[m][i], h[m][i]);
        std::fclose(f);
        return 0;
    }

    static int fail(const std::string &msg) {
        falm::set_error("Mesher: %s", msg.c_str());
        return 1;
    }

    static bool read_nodes(const std::string &path, std::vector<double> &v) {
        std::ifstream f(path);
        std::string line;
        if (!f || !std::getline(f, line)) return false;
        try {
            const long n = std::stol(line);
            if (n < 0) return false;
            v.resize((size_t)n);
            for (long q = 0; q < n; q++) {
                if (!std::getline(f, line)) return false;
                v[(size_t)q] = std::stod(line);
            }
        } catch (const std::exception &) {
            return false;
        }
        return true;
    }

    // cell centres at the nodes (src/mesher/mesher.hpp:20-106)
    static int inner_node(const std::vector<double> &nd, int gc, bool z_axis, std::vector<double> &c,
                          std::vector<double> &h) {
        const int g = gc - 1;  // ghost cells beyond the outermost node
        const int n = (int)nd.size(), len = n + 2 * g;
        if (n < 4) return 1;
        c.assign((size_t)len, 0.0);
        h.assign((size_t)len, 0.0);
        for (int q = 0; q < n; q++) c[(size_t)(g + q)] = nd[(size_t)q];
        for (int i = g + 1; i <= len - g - 2; i++) h[i] = 0.5 * (c[i + 1] - c[i - 1]);
        for (int i = g; i >= 0; i--) h[i] = 2 * h[i + 1] - h[i + 2];
        for (int i = len - g - 1; i < len; i++) h[i] = z_axis ? h[i - 1] : 2 * h[i - 1] - h[i - 2];
        for (int i = g - 1; i >= 0; i--) c[i] = c[i + 1] - 0.5 * (h[i] + h[i + 1]);
        for (int i = len - g; i < len; i++) c[i] = c[i - 1] + 0.5 * (h[i] + h[i - 1]);
        return 0;
    }

    // cell faces at the nodes (src/mesher/mesher.hpp:107-237)
    static int cell_faces(const std::vector<double> &nd, int gc, std::vector<double> &c, std::vector<double> &h) {
        const int n = (int)nd.size(), len = n - 1 + 2 * gc;
        if (n < 2) return 1;
        c.assign((size_t)len, 0.0);
        h.assign((size_t)len, 0.0);
        for (int i = gc; i < len - gc; i++) h[i] = nd[i - gc + 1] - nd[i - gc];
        for (int i = gc - 1; i >= 0; i--) h[i] = n >= 3 ? 2 * h[i + 1] - h[i + 2] : h[i + 1];
        for (int i = len - gc; i < len; i++) h[i] = n >= 3 ? 2 * h[i - 1] - h[i - 2] : h[i - 1];
        for (int i = gc; i < len - gc; i++) c[i] = nd[i - gc] + 0.5 * h[i];
        for (int i = gc - 1; i >= 0; i--) c[i] = c[i + 1] - 0.5 * (h[i] + h[i + 1]);
        for (int i = len - gc; i < len; i++) c[i] = c[i - 1] + 0.5 * (h[i] + h[i - 1]);
        return 0;
    }
};

}  // namespace

extern "C" falm_status falm_mesher_build(const char *cvcenter, const char *meshpath, const char *outputfile,
                                         int gc) {
    if (!cvcenter || !meshpath || !outputfile) return FALM_ERR_INVALID;
    return Mesher::build_mesh(cvcenter, meshpath, outputfile, gc) == 0 ? FALM_OK : FALM_ERR_SETUP;
}
