// Falm C++ host API: the partition descriptor CPM (src/CPMBase.h:12-206),
// single-rank.  On one GPU there is one sub-domain covering the global
// domain, every neighbour is -1 and the halo exchange of CPMComm is empty;
// ghost cells are filled by the boundary kernels (falm_bc_*).  Asking for
// more than one rank is an error here.
#ifndef FALM_AMD_CPM_H
#define FALM_AMD_CPM_H

#include <vector>

#include "matrix.h"

namespace Falm {

class CPM {
public:
    static const Int XMINUS = 0;
    static const Int XPLUS = 1;
    static const Int YMINUS = 2;
    static const Int YPLUS = 3;
    static const Int ZMINUS = 4;
    static const Int ZPLUS = 5;
    static const Int NFACE = 6;

    int neighbour[6] = {-1, -1, -1, -1, -1, -1};
    Int3 shape{{1, 1, 1}};
    Int3 idx{{0, 0, 0}};
    int rank = 0;
    int size = 1;
    bool use_cuda_aware_mpi = false;
    Int gc = GuideCell;
    Region global;
    std::vector<Region> pdm_list;

    // CPM::initPartition (src/CPMBase.h:36-112) for mpi_shape {1,1,1}
    int initPartition(Int3 gShape, Int guideCell, Int3 mpi_shape = Int3{{1, 1, 1}}) {
        if (PRODUCT3(mpi_shape) != 1) {
            fprintf(stderr, "CPM::initPartition: this build runs one sub-domain per process\n");
            return lastStatus() = FALM_ERR_UNSUPPORTED;
        }
        shape = mpi_shape;
        gc = guideCell;
        global = Region(gShape, Int3{{0, 0, 0}});
        pdm_list.assign(1, Region(gShape, Int3{{0, 0, 0}}));
        return FALM_OK;
    }

    // CPM::set6Region (src/CPMBase.h:114-150): no neighbours, so the inner
    // region is the whole map
    void set6Region(Int3 &inner_shape, Int3 &inner_offset, Int3 *, Int3 *, Int, const Region &map) const {
        inner_shape = map.shape;
        inner_offset = map.offset;
    }

    bool validNeighbour(Int fid) const { return neighbour[fid] >= 0; }

    // the C ABI's domain descriptor of this rank's sub-domain
    falm_domain domain() const {
        falm_domain d;
        const Region &p = pdm_list[rank];
        for (int m = 0; m < 3; m++) {
            d.shape[m] = p.shape[m];
            d.offset[m] = p.offset[m];
            d.gshape[m] = global.shape[m];
        }
        d.gc = gc;
        return d;
    }
};

}  // namespace Falm

#endif
