// shim_check.cpp -- calls every hot-path entry point of the reference's headers
// (fea.hpp, solver.hpp, multigrid.hpp, filter.hpp) on one seeded problem and
// writes each result as a raw file <name>.bin into the working directory.
//
// Built twice by oracle/Makefile's `shim` target: linked with the reference's
// own CPU sources (oracle/_ref/shim_check_ref) and with the drop-in
// integration/top3d_gpu.cpp on libt3d.so (oracle/_ref/shim_check_gpu);
// tests/test_gpu_shim.py runs both and compares the files.  Test
// infrastructure: nothing of the product uses it.
//
// usage: shim_check nely nelx nelz
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "core.hpp"
#include "fea.hpp"
#include "filter.hpp"
#include "multigrid/multigrid.hpp"
#include "solver.hpp"

using namespace top3d;

namespace {
template <class T>
void dump(const char* name, const std::vector<T>& v)
{
    FILE* f = std::fopen((std::string(name) + ".bin").c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", name);
        std::exit(1);
    }
    std::fclose(f);
}
struct Rng {   // xorshift64*
    uint64_t s;
    double uniform(double lo, double hi)
    {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return lo + (hi - lo) * double((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    }
};
}  // namespace

int main(int argc, char** argv)
{
    const int nely = argc > 3 ? std::atoi(argv[1]) : 8;
    const int nelx = argc > 3 ? std::atoi(argv[2]) : 16;
    const int nelz = argc > 3 ? std::atoi(argv[3]) : 8;
    Problem pb;
    InitialSettings(pb.params);
    CreateVoxelFEAmodel_Cuboid(pb, nely, nelx, nelz);
    ApplyBoundaryConditions(pb);
    const int ne = pb.mesh.numElements;
    const size_t nF = pb.freeDofIndex.size();
    dump("eleMapBack", pb.mesh.eleMapBack);
    dump("eNodMat", pb.mesh.eNodMat);
    dump("eDofMat", pb.mesh.eDofMat);
    dump("elemColor", pb.mesh.coloring.elemColor);
    dump("freeDofIndex", pb.freeDofIndex);
    dump("eDofFreeMat", pb.eDofFreeMat);
    dump("Fz", pb.F.uz);
    const auto Ke = ComputeVoxelKe(0.3, 1.0);
    dump("Ke", std::vector<double>(Ke.begin(), Ke.end()));

    Rng rng{0x9e3779b97f4a7c15ull};
    std::vector<double> E(static_cast<size_t>(ne));
    for (int e = 0; e < ne; ++e) E[size_t(e)] = (e % 7 == 3) ? 1e-9 : rng.uniform(1e-3, 1.0);
    std::vector<double> u(nF);
    for (double& v : u) v = rng.uniform(-1.0, 1.0);

    std::vector<double> y;
    K_times_u_finest(pb, E, u, y);
    dump("Ku", y);
    KTimesUWorkspace kws;
    K_times_u_finest(pb, E, y, u, kws);   // K (K u), the workspace overload
    dump("KKu", u);

    DOFData U;
    U.ux.assign(size_t(pb.mesh.numNodes), 0.5);
    U.uy.assign(size_t(pb.mesh.numNodes), -0.25);
    U.uz.assign(size_t(pb.mesh.numNodes), 0.125);
    scatter_from_free(pb, y, U);   // fixed DOFs keep their values
    dump("scatter_uy", U.uy);
    std::vector<double> back;
    restrict_to_free(pb, U, back);
    dump("restrict", back);
    for (size_t i = 0; i < nF; ++i) back[i] = 0.0;
    scatter_from_free(pb, back, U);   // the fixed-DOF values remain, as ComputeCompliance sees them
    scatter_from_free(pb, y, U);
    std::vector<double> ce;
    std::vector<double> C(1, ComputeCompliance(pb, E, U, ce));
    dump("ce", ce);
    dump("C", C);

    mg::MGPrecondConfig cfg;
    mg::MGHierarchy H;
    std::vector<std::vector<uint8_t>> masks;
    mg::build_static_once(pb, cfg, H, masks);
    Preconditioner M = mg::make_diagonal_preconditioner_from_static(pb, H, masks, E, cfg);
    std::vector<double> z;
    M(y, z);
    dump("vcycle", z);

    std::vector<double> b;
    restrict_to_free(pb, pb.F, b);
    PCGFreeWorkspace ws;
    std::vector<int> iters;
    std::vector<double> x;
    iters.push_back(PCG_free(pb, E, b, x, 1e-3, 200, M, ws));
    dump("x_mg", x);
    x.clear();
    Preconditioner J = make_jacobi_preconditioner(pb, E);
    iters.push_back(PCG_free(pb, E, b, x, 1e-3, 300, J, ws));
    dump("x_jacobi", x);
    x.clear();
    // unpreconditioned CG, 20 iterations: with 1e-9 voids the system is so ill
    // conditioned that longer unconverged runs amplify the dot products'
    // summation order to percent-level differences in x
    iters.push_back(PCG_free(pb, E, b, x, 1e-3, 20, Preconditioner(), ws));
    dump("x_cg", x);
    std::vector<double> zero(nF, 0.0);   // zero right-hand side (solver.cpp:147-149)
    x.clear();
    iters.push_back(PCG_free(pb, E, zero, x, 1e-3, 50, M, ws));
    dump("x_zero", x);
    dump("iters", iters);

    PDEFilter pf = SetupPDEFilter(pb, 1.5f);
    dump("filter_kernel", std::vector<double>(pf.kernel.begin(), pf.kernel.end()));
    dump("filter_diag", pf.diagPrecondNode);
    PDEFilterWorkspace fws;
    std::vector<float> src(static_cast<size_t>(ne)), dst;
    for (float& v : src) v = float(rng.uniform(-1.0, 0.0));
    ApplyPDEFilter(pb, pf, src, dst, fws);
    dump("filter1", dst);
    for (float& v : src) v *= 1.05f;   // a warm start from the first call
    ApplyPDEFilter(pb, pf, src, dst, fws);
    dump("filter2", dst);
    std::printf("shim_check done: %d elements, %zu free DOFs, PCG %d %d %d %d\n", ne, nF, iters[0], iters[1],
                iters[2], iters[3]);
    return 0;
}
