// windtunnel.cpp — a wind-tunnel application written against the Falm C++
// class API exactly as the reference's app/windtunnel/main.cpp is written
// against the reference's (host matrix assembly, in-place pseudo-velocity
// call, the same step order), running on the falm_amd library.
//
//   windtunnel NX NY NZ H RE DT STEPS U_IN OUT_FILE
//
// Writes u (3 columns) and p (1 column) of the padded domain as float64 to
// OUT_FILE after STEPS steps, and prints the reference's per-step line
// "it, t, |div|/N, solver it, solver err".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/falm/falm.h"

using namespace Falm;

static CPM cpm;
static FalmCFD cfd;
static FalmEq eq;
static Real maxDiag, dt, uInflow;
static Matrix<Real> u, uu, p, nut, ff, kx, g, ja, xyz, poi_a, poi_rhs, poi_res, divergence, uPrevious;
static const dim3 blockSize{8, 8, 8};

// makePoissonCoefficientMatrix, app/windtunnel/main.cpp:20-92
static void makePoissonCoefficientMatrix() {
    Region &pdm = cpm.pdm_list[cpm.rank];
    Region &global = cpm.global;
    g.sync(MCP::Dev2Hst);
    ja.sync(MCP::Dev2Hst);
    poi_a.alloc(pdm.shape, 7, HDC::HstDev, "poisson coefficients", StencilMatrix::D3P7);
    for (Int k = cpm.gc; k < pdm.shape[2] - cpm.gc; k++)
        for (Int j = cpm.gc; j < pdm.shape[1] - cpm.gc; j++)
            for (Int i = cpm.gc; i < pdm.shape[0] - cpm.gc; i++) {
                Int3 gijk = pdm.offset + Int3{{i, j, k}};
                Real ac = 0, ae = 0, aw = 0, an = 0, as = 0, at = 0, ab = 0;
                Int c = IDX(i, j, k, pdm.shape);
                Real coef, jac = ja(c);
                coef = 0.5 * (g(c, 0) + g(IDX(i + 1, j, k, pdm.shape), 0)) / jac;
                if (gijk[0] < global.shape[0] - cpm.gc) { ac -= coef; ae = coef; }
                coef = 0.5 * (g(c, 0) + g(IDX(i - 1, j, k, pdm.shape), 0)) / jac;
                if (gijk[0] > cpm.gc) { ac -= coef; aw = coef; }
                coef = 0.5 * (g(c, 1) + g(IDX(i, j + 1, k, pdm.shape), 1)) / jac;
                if (gijk[1] < global.shape[1] - cpm.gc - 1) { ac -= coef; an = coef; }
                coef = 0.5 * (g(c, 1) + g(IDX(i, j - 1, k, pdm.shape), 1)) / jac;
                if (gijk[1] > cpm.gc) { ac -= coef; as = coef; }
                coef = 0.5 * (g(c, 2) + g(IDX(i, j, k + 1, pdm.shape), 2)) / jac;
                if (gijk[2] < global.shape[2] - cpm.gc - 1) { ac -= coef; at = coef; }
                coef = 0.5 * (g(c, 2) + g(IDX(i, j, k - 1, pdm.shape), 2)) / jac;
                if (gijk[2] > cpm.gc) { ac -= coef; ab = coef; }
                poi_a(c, 0) = ac; poi_a(c, 1) = aw; poi_a(c, 2) = ae; poi_a(c, 3) = as;
                poi_a(c, 4) = an; poi_a(c, 5) = ab; poi_a(c, 6) = at;
            }
    poi_a.sync(MCP::Hst2Dev);
    maxDiag = FalmMV::MaxDiag(poi_a, cpm, blockSize);
    FalmMV::ScaleMatrix(poi_a, 1.0 / maxDiag, blockSize);
}

// ubc / pbc of app/windtunnel/bc.h, through the library's ghost-cell fill
static void ubc() {
    falm_face_bc f[6] = {};
    f[CPM::XMINUS].type = FALM_BC_INLET;
    f[CPM::XMINUS].value[0] = uInflow;
    f[CPM::XPLUS].type = FALM_BC_CONVECTIVE;
    f[CPM::YMINUS].type = f[CPM::YPLUS].type = f[CPM::ZPLUS].type = FALM_BC_SLIP;
    f[CPM::ZMINUS].type = FALM_BC_WALL0;
    falm_domain d = cpm.domain();
    falmErrCheckMacro(falm_bc_velocity(&d, FALM_F64, f, u.dev.ptr, uPrevious.dev.ptr, xyz.dev.ptr, dt, nullptr, nullptr));
    falmErrCheckMacro(falm_stream_sync(nullptr));
}
static void pbc() {
    int f[6] = {FALM_PBC_NEUMANN, FALM_PBC_ZERO, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN};
    falm_domain d = cpm.domain();
    falmErrCheckMacro(falm_bc_pressure(&d, FALM_F64, f, p.dev.ptr, nullptr));
    falmErrCheckMacro(falm_stream_sync(nullptr));
}

// app/windtunnel/main.cpp:135-165
static Real mainLoop() {
    uPrevious.copy(u, HDC::Device);
    cfd.FSPseudoU(uPrevious, u, uu, u, nut, kx, g, ja, ff, dt, cpm, blockSize);
    cfd.UtoUU(u, uu, kx, ja, cpm, blockSize);
    cfd.MACCalcPoissonRHS(uu, poi_rhs, ja, dt, cpm, blockSize, maxDiag);
    eq.Solve(poi_a, p, poi_rhs, poi_res, cpm, blockSize);
    pbc();
    cfd.ProjectP(u, u, uu, uu, p, kx, g, dt, cpm, blockSize);
    ubc();
    cfd.SGS(u, nut, xyz, kx, ja, cpm, blockSize);
    cfd.Divergence(uu, divergence, ja, cpm, blockSize);
    return FalmMV::EuclideanNormSq(divergence, cpm, blockSize);
}

int main(int argc, char **argv) {
    if (argc < 10) {
        fprintf(stderr, "usage: %s NX NY NZ H RE DT STEPS U_IN OUT_FILE\n", argv[0]);
        return 2;
    }
    const Int n[3] = {atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    const Real h = atof(argv[4]), Re = atof(argv[5]);
    dt = atof(argv[6]);
    const int steps = atoi(argv[7]);
    uInflow = atof(argv[8]);
    if (falmErrCheckMacro(falm_init(0)) != FALM_OK) return 1;

    Int3 shape{{n[0] + 2 * GuideCell, n[1] + 2 * GuideCell, n[2] + 2 * GuideCell}};
    cpm.initPartition(shape, GuideCell);
    // uniform control volumes (FalmCore::computation_init, src/falm.h:342-431)
    std::vector<Real> c[3], w[3];
    for (int m = 0; m < 3; m++)
        for (Int i = 0; i < shape[m]; i++) {
            c[m].push_back((i - GuideCell + 0.5) * h);
            w[m].push_back(h);
        }
    for (Matrix<Real> *f : {&xyz, &kx, &g}) f->alloc(shape, 3, HDC::Device, "");
    ja.alloc(shape, 1, HDC::Device, "jacobian");
    falm_domain d = cpm.domain();
    falmErrCheckMacro(falm_metrics(&d, FALM_F64, c[0].data(), c[1].data(), c[2].data(), w[0].data(), w[1].data(),
                                   w[2].data(), xyz.dev.ptr, kx.dev.ptr, g.dev.ptr, ja.dev.ptr, nullptr));
    u.alloc(shape, 3, HDC::HstDev, "velocity");
    for (Matrix<Real> *f : {&uu, &ff, &uPrevious}) f->alloc(shape, 3, HDC::Device, "");
    for (Matrix<Real> *f : {&p, &nut, &poi_rhs, &poi_res, &divergence}) f->alloc(shape, 1, HDC::Device, "");

    cfd.init(Re, AdvectionSchemeType::Upwind3, SGSType::Smagorinsky, 0.1);
    cfd.alloc(shape);
    eq.init(LSType::PBiCGStab, 1000, 1e-12, 1.0, LSType::SOR, 2, 1.5);

    // init: u = (U, 0, 0) everywhere, UtoUU, SGS, Poisson matrix
    // (app/windtunnel/main.cpp:116-131)
    for (Int i = 0; i < u.shape[0]; i++) {
        u(i, 0) = uInflow;
        u(i, 1) = 0.0;
        u(i, 2) = 0.0;
    }
    u.sync(MCP::Hst2Dev);
    cfd.UtoUU(u, uu, kx, ja, cpm, blockSize);
    cfd.SGS(u, nut, xyz, kx, ja, cpm, blockSize);
    makePoissonCoefficientMatrix();

    const Real nint = (Real)PRODUCT3(cpm.pdm_list[0].shape - Int(2 * cpm.gc));
    for (int it = 1; it <= steps; it++) {
        Real div = sqrt(mainLoop()) / nint;
        printf("%8d %12.5e, %12.5e, %3d, %12.5e\n", it, it * dt, div, eq.it, eq.err);
    }
    u.sync(MCP::Dev2Hst);
    p.sync(MCP::Dev2Hst);  // creates the host side on demand (src/matrix.h:290-298)
    FILE *fo = fopen(argv[9], "wb");
    if (!fo) return 1;
    fwrite(u.host.ptr, sizeof(Real), (size_t)u.size, fo);
    fwrite(p.host.ptr, sizeof(Real), (size_t)p.size, fo);
    fclose(fo);
    return lastStatus() == FALM_OK ? 0 : 1;
}
