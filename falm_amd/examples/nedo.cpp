// nedo.cpp — the reference's set-up-file applications, written against the
// Falm C++ class API the way app/nedo/main.cpp and app/windtunnel/main.cpp
// are written against the reference's: FalmCore reads setup.json (runtime,
// solver, mesh, inflow, turbine), converts the node list into the
// control-volume file, runs the step loop and writes snapshots, time
// averages, the index file and (with a turbine) the power/thrust CSV.
//
//   nedo [SETUP_JSON] [info]
//
// SETUP_JSON defaults to setup.json in the current directory; paths inside
// it are relative to its directory, except the output prefix of the
// power/thrust CSV and end message, which the reference resolves against
// the current directory.  "info" stops after the set-up, as in the
// reference.  Boundary conditions: convective outflow at x+, slip at y+-,
// no-slip wall at z-, slip at z+, pressure 0 beyond x+ and zero gradient
// elsewhere, and at x- the inlet of the app the setup belongs to: with a
// "turbine" block, app/nedo's (1, 0, 0) whatever the inflow velocity
// (app/nedo/bcdevcall.cu:16); without one, app/windtunnel's
// (inflow velocity, 0, 0) (app/windtunnel/bcdevcall.cu:8-22).
#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/falm/falm.h"

using namespace Falm;

static FalmCore falm;
static Real maxdiag;
static Matrix<Real> u_previous;
static Alm::AlmHandler aalm;
static bool with_turbine = false;
static Real inlet_u = 1.0;
static const dim3 block{8, 8, 8};

// make_poisson_coefficient_matrix (app/nedo/main.cpp:25-100): the same
// coefficients, assembled on the device, then scaled by max |a_c|
static int make_poisson_coefficient_matrix() {
    falm_domain d = falm.cpm.domain();
    FalmBasicVar &fv = falm.fv;
    return falmErrCheckMacro(falm_poisson_matrix(&d, FALM_F64, FALM_POISSON_OUTLET_XPLUS, fv.g.dev.ptr,
                                                 fv.ja.dev.ptr, fv.poi_a.dev.ptr, &maxdiag, nullptr));
}

// ubc / pbc (app/nedo/bc.h, app/nedo/bcdevcall.cu)
static void ubc(Matrix<Real> &u, Matrix<Real> &uprev, Matrix<Real> &x, Real dt) {
    falm_face_bc f[6] = {};
    f[CPM::XMINUS].type = FALM_BC_INLET;
    f[CPM::XMINUS].value[0] = inlet_u;
    f[CPM::XPLUS].type = FALM_BC_CONVECTIVE;
    f[CPM::YMINUS].type = f[CPM::YPLUS].type = f[CPM::ZPLUS].type = FALM_BC_SLIP;
    f[CPM::ZMINUS].type = FALM_BC_WALL0;
    falm_domain d = falm.cpm.domain();
    falmErrCheckMacro(falm_bc_velocity(&d, FALM_F64, f, u.dev.ptr, uprev.dev.ptr, x.dev.ptr, dt, nullptr, nullptr));
    falmErrCheckMacro(falm_stream_sync(nullptr));
}
static void pbc(Matrix<Real> &p) {
    int f[6] = {FALM_PBC_NEUMANN, FALM_PBC_ZERO, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN, FALM_PBC_NEUMANN};
    falm_domain d = falm.cpm.domain();
    falmErrCheckMacro(falm_bc_pressure(&d, FALM_F64, f, p.dev.ptr, nullptr));
    falmErrCheckMacro(falm_stream_sync(nullptr));
}

// init (app/nedo/main.cpp:102-148)
static int init(int &argc, char **&argv, const std::string &setup) {
    if (falm.env_init(argc, argv) != FALM_OK) return 1;
    if (falm.parse_settings(setup) != FalmErr::success) {
        fprintf(stderr, "cannot read setup file %s\n", setup.c_str());
        return 1;
    }
    if (falm.computation_init(Int3{{falm.cpm.size, 1, 1}}, GuideCell) != FALM_OK) return 1;
    falm.print_info(0);
    u_previous.alloc(falm.fv.u.shape[0], falm.fv.u.shape[1], HDC::Device, "previous velocity");

    const Real u_inflow = falm.params["inflow"]["velocity"].get<Real>();
    FalmBasicVar &fv = falm.fv;
    Matrix<Real> &u = fv.u;
    for (Int i = 0; i < u.shape[0]; i++) {
        u(i, 0) = u_inflow;
        u(i, 1) = u(i, 2) = 0.0;
    }
    u.sync(MCP::Hst2Dev);
    falm.falmCfd.UtoUU(fv.u, fv.uu, fv.kx, fv.ja, falm.cpm, block);
    falm.falmCfd.SGS(fv.u, fv.nut, fv.xyz, fv.kx, fv.ja, falm.cpm, block);
    if (make_poisson_coefficient_matrix() != FALM_OK) return 1;

    with_turbine = falm.params.contains("turbine");
    inlet_u = with_turbine ? 1.0 : u_inflow;
    if (with_turbine) {
        if (aalm.init(falm.workdir, falm.params["turbine"], falm.cpm, falm.outputPrefix + "_PT.csv") != FALM_OK)
            return 1;
        printf("rank %d\n", falm.cpm.rank);
        aalm.print_info(falm.cpm.rank);
        fflush(stdout);
    }
    return lastStatus() == FALM_OK ? 0 : 1;
}

// main_loop (app/nedo/main.cpp:150-204)
static Real main_loop() {
    FalmBasicVar &fv = falm.fv;
    FalmCFD &fcfd = falm.falmCfd;
    u_previous.copy(fv.u, HDC::Device);
    if (with_turbine) aalm.Alm(falm.baseMesh.x, falm.baseMesh.y, falm.baseMesh.z, fv.u, fv.ff, falm.gettime(), block);
    fcfd.FSPseudoU(u_previous, fv.u, fv.uu, fv.u, fv.nut, fv.kx, fv.g, fv.ja, fv.ff, falm.dt, falm.cpm, block, nullptr, 1);
    fcfd.UtoUU(fv.u, fv.uu, fv.kx, fv.ja, falm.cpm, block);
    fcfd.MACCalcPoissonRHS(fv.uu, fv.poi_rhs, fv.ja, falm.dt, falm.cpm, block, maxdiag);
    falm.falmEq.Solve(fv.poi_a, fv.p, fv.poi_rhs, fv.poi_res, falm.cpm, block);
    pbc(fv.p);
    fcfd.ProjectP(fv.u, fv.u, fv.uu, fv.uu, fv.p, fv.kx, fv.g, falm.dt, falm.cpm, block);
    ubc(fv.u, u_previous, fv.xyz, falm.dt);
    fcfd.SGS(fv.u, fv.nut, fv.xyz, fv.kx, fv.ja, falm.cpm, block);
    fcfd.Divergence(fv.uu, fv.divergence, fv.ja, falm.cpm, block);
    const Real divnorm = FalmMV::EuclideanNormSq(fv.divergence, falm.cpm, block);
    falm.TAvg();
    falm.outputUVWP();
    if (with_turbine) aalm.writePowerThrust(falm.gettime(), falm.params["inflow"]["velocity"].get<Real>());
    return divnorm;
}

int main(int argc, char **argv) {
    std::string setup = "setup.json";
    bool info_only = false;
    for (int a = 1; a < argc; a++) {
        const std::string arg(argv[a]);
        if (arg == "info") info_only = true;
        else setup = arg;
    }
    if (init(argc, argv, setup) != 0) return 1;
    if (!info_only) {
        printf("\n");
        printf("maxdiag = %e %e\n", maxdiag, maxdiag);
        const Real nint = (Real)PRODUCT3(falm.cpm.pdm_list[falm.cpm.rank].shape - Int(2 * falm.cpm.gc));
        for (falm.it = 1; falm.it <= falm.maxIt; falm.it++) {
            const Real divnorm = sqrt(main_loop()) / nint;
            printf("%8d %12.5e, %12.5e, %3d, %12.5e\n", falm.it, falm.gettime(), divnorm, falm.falmEq.it,
                   falm.falmEq.err);
            fflush(stdout);
            if (lastStatus() != FALM_OK) break;
        }
        printf("\n");
    }
    u_previous.release();
    aalm.finalize();
    falm.env_finalize();
    return lastStatus() == FALM_OK ? 0 : 1;
}
