// eq_config.cpp — the FalmEq class API with its settings changed between
// solves, as reference apps may do (the reference reads type, tol, maxit,
// relax and the preconditioner settings on every Solve, src/FalmEq.h:56-64).
//
//   eq_config NX NY NZ A_FILE B_FILE
//
// A_FILE holds the 7-column Poisson matrix and B_FILE the right-hand side of
// the padded domain (float64, Matrix layout).  Each solve starts from x = 0
// and prints "label it err".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/falm/falm.h"

using namespace Falm;

static bool read_file(const char *path, std::vector<double> &v) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    size_t got = fread(v.data(), sizeof(double), v.size(), f);
    fclose(f);
    return got == v.size();
}

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: eq_config NX NY NZ A_FILE B_FILE\n");
        return 2;
    }
    if (falm_init(0) != FALM_OK) return 3;
    CPM cpm;
    Int3 n{{atoi(argv[1]), atoi(argv[2]), atoi(argv[3])}};
    cpm.initPartition(n + Int3{{2 * GuideCell, 2 * GuideCell, 2 * GuideCell}}, GuideCell);
    Region &pdm = cpm.pdm_list[cpm.rank];
    const size_t nc = (size_t)PRODUCT3(pdm.shape);
    std::vector<double> ha(7 * nc), hb(nc);
    if (!read_file(argv[4], ha) || !read_file(argv[5], hb)) {
        fprintf(stderr, "cannot read the inputs\n");
        return 4;
    }
    Matrix<Real> a(pdm.shape, 7, HDC::HstDev, "a"), x(pdm.shape, 1, HDC::Device, "x"),
        b(pdm.shape, 1, HDC::HstDev, "b"), r(pdm.shape, 1, HDC::Device, "r");
    for (size_t q = 0; q < 7 * nc; q++) a.host.ptr[q] = ha[q];
    for (size_t q = 0; q < nc; q++) b.host.ptr[q] = hb[q];
    a.sync(MCP::Hst2Dev);
    b.sync(MCP::Hst2Dev);

    FalmEq eq(LSType::PBiCGStab, 1000, 1e-8, 1.2, LSType::SOR, 2, 1.5);
    auto run = [&](const char *label, auto solve) {
        x.clear(HDC::Device);
        solve();
        printf("%s %d %.17e\n", label, (int)eq.it, eq.err);
    };
    run("bicgstab_sor_1e-8", [&] { eq.Solve(a, x, b, r, cpm); });
    run("sor_1e-8", [&] { eq.SOR(a, x, b, r, cpm); });
    run("bicgstab_sor_1e-8_again", [&] { eq.PBiCGStab(a, x, b, r, cpm); });
    eq.tol = 1e-11;
    run("bicgstab_sor_1e-11", [&] { eq.Solve(a, x, b, r, cpm); });
    eq.init(LSType::PBiCGStab, 1000, 1e-9, 1.2, LSType::Jacobi, 3, 1.0);
    run("bicgstab_jacobi_1e-9", [&] { eq.Solve(a, x, b, r, cpm); });
    eq.maxit = 3;
    run("bicgstab_jacobi_maxit3", [&] { eq.Solve(a, x, b, r, cpm); });
    printf("solvers %zu\n", eq.solvers.size());
    return 0;
}
