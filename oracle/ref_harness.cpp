// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" glue around the reference C++ port, compiled together with
// the reference's own sources (oracle/Makefile builds them where they lie and
// writes only into oracle/_ref/).  Tests load oracle/_ref/libtop3d_ref.so to
// pin the oracle against the reference itself.  Nothing here re-implements an
// algorithm: every number comes from a reference function; the OC loop below
// strings reference calls together exactly as driver.cpp:112-189 does, with
// parameters taken from the caller instead of InitialSettings.
#include "core.hpp"
#include "fea.hpp"
#include "solver.hpp"
#include "filter.hpp"
#include "io.hpp"
#include "geometry_out/export_stl.hpp"
#include "geometry_out/voxel_surface.hpp"
#include "multigrid/multigrid.hpp"
#include "multigrid/detail/diagonal.hpp"
#include "multigrid/detail/coarsest.hpp"
#include "multigrid/detail/transfers.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

using namespace top3d;

namespace {
struct RefState {
    int nely = -1, nelx = -1, nelz = -1;
    Problem pb;
    mg::MGHierarchy H;
    std::vector<std::vector<uint8_t>> masks;
    mg::MGPrecondConfig cfg;
    PDEFilter pf;
    PDEFilterWorkspace pfws;
    bool haveFilter = false;
};
std::unique_ptr<RefState> g;
// hierarchy of the multigrid functions below (MGPrecondConfig, multigrid.hpp:9-13):
// the reference driver's nonDyadic = true, maxLevels = 5 unless ref_set_hierarchy changes it
bool hNonDyadic = true;
int hMaxLevels = 5;

RefState& state(int nely, int nelx, int nelz)
{
    if (!g || g->nely != nely || g->nelx != nelx || g->nelz != nelz) {
        g.reset(new RefState());
        g->nely = nely; g->nelx = nelx; g->nelz = nelz;
        InitialSettings(g->pb.params);
        CreateVoxelFEAmodel_Cuboid(g->pb, nely, nelx, nelz);
        ApplyBoundaryConditions(g->pb);
        g->cfg.nonDyadic = hNonDyadic; g->cfg.maxLevels = hMaxLevels; g->cfg.weight = 0.6f;
        mg::build_static_once(g->pb, g->cfg, g->H, g->masks);
    }
    return *g;
}
}  // namespace

extern "C" {

// the hierarchy used from here on (the cached problem is rebuilt when it changes)
void ref_set_hierarchy(int nonDyadic, int maxLevels)
{
    if ((nonDyadic != 0) != hNonDyadic || maxLevels != hMaxLevels) g.reset();
    hNonDyadic = nonDyadic != 0;
    hMaxLevels = maxLevels;
}

// sizes: resX resY resZ ne nN nF nlev
void ref_sizes(int nely, int nelx, int nelz, int* out)
{
    RefState& s = state(nely, nelx, nelz);
    out[0] = s.pb.mesh.resX; out[1] = s.pb.mesh.resY; out[2] = s.pb.mesh.resZ;
    out[3] = s.pb.mesh.numElements; out[4] = s.pb.mesh.numNodes;
    out[5] = (int)s.pb.freeDofIndex.size(); out[6] = (int)s.H.levels.size();
}

void ref_model(int nely, int nelx, int nelz, int32_t* eleMapBack, int32_t* eNodMat, int32_t* color,
               double* Ke, double* F, uint8_t* isFree, int32_t* freeDofIndex, int32_t* eDofFree)
{
    RefState& s = state(nely, nelx, nelz);
    const auto& m = s.pb.mesh;
    const int ne = m.numElements, nN = m.numNodes;
    std::memcpy(eleMapBack, m.eleMapBack.data(), sizeof(int32_t) * ne);
    std::memcpy(eNodMat, m.eNodMat.data(), sizeof(int32_t) * 8 * (size_t)ne);
    for (int e = 0; e < ne; ++e) color[e] = m.coloring.elemColor[e];
    std::memcpy(Ke, m.Ke.data(), sizeof(double) * 576);
    for (int n = 0; n < nN; ++n) { F[3 * n] = s.pb.F.ux[n]; F[3 * n + 1] = s.pb.F.uy[n]; F[3 * n + 2] = s.pb.F.uz[n]; }
    std::memcpy(isFree, s.pb.isFreeDOF.data(), (size_t)3 * nN);
    std::memcpy(freeDofIndex, s.pb.freeDofIndex.data(), sizeof(int32_t) * s.pb.freeDofIndex.size());
    std::memcpy(eDofFree, s.pb.eDofFreeMat.data(), sizeof(int32_t) * 24 * (size_t)ne);
}

void ref_bfree(int nely, int nelx, int nelz, double* bFree)
{
    RefState& s = state(nely, nelx, nelz);
    std::vector<double> b;
    restrict_to_free(s.pb, s.pb.F, b);
    std::memcpy(bFree, b.data(), sizeof(double) * b.size());
}

void ref_kxu(int nely, int nelx, int nelz, const double* E, const double* u, double* y)
{
    RefState& s = state(nely, nelx, nelz);
    const int ne = s.pb.mesh.numElements;
    const size_t nF = s.pb.freeDofIndex.size();
    std::vector<double> Ev(E, E + ne), uv(u, u + nF), yv;
    K_times_u_finest(s.pb, Ev, uv, yv);
    std::memcpy(y, yv.data(), sizeof(double) * nF);
}

double ref_compliance(int nely, int nelx, int nelz, const double* E, const double* uFree, double* ce)
{
    RefState& s = state(nely, nelx, nelz);
    const int ne = s.pb.mesh.numElements, nN = s.pb.mesh.numNodes;
    const size_t nF = s.pb.freeDofIndex.size();
    DOFData U;
    U.ux.assign(nN, 0.0); U.uy.assign(nN, 0.0); U.uz.assign(nN, 0.0);
    std::vector<double> uv(uFree, uFree + nF), Ev(E, E + ne), cev;
    scatter_from_free(s.pb, uv, U);
    double C = ComputeCompliance(s.pb, Ev, U, cev);
    std::memcpy(ce, cev.data(), sizeof(double) * ne);
    return C;
}

// level dims: out[4*l..] = resX resY resZ span
int ref_mg_levels(int nely, int nelx, int nelz, int* out)
{
    RefState& s = state(nely, nelx, nelz);
    for (size_t l = 0; l < s.H.levels.size(); ++l) {
        out[4 * l] = s.H.levels[l].resX; out[4 * l + 1] = s.H.levels[l].resY;
        out[4 * l + 2] = s.H.levels[l].resZ; out[4 * l + 3] = s.H.levels[l].spanWidth;
    }
    return (int)s.H.levels.size();
}

void ref_mg_mask(int nely, int nelx, int nelz, int level, uint8_t* out)
{
    RefState& s = state(nely, nelx, nelz);
    std::memcpy(out, s.masks[level].data(), s.masks[level].size());
}

void ref_mg_diag(int nely, int nelx, int nelz, const double* E, int level, double* out)
{
    RefState& s = state(nely, nelx, nelz);
    std::vector<double> Ev(E, E + s.pb.mesh.numElements);
    std::vector<std::vector<double>> diag;
    mg::MG_BuildDiagonals(s.pb, s.H, s.masks, Ev, diag);
    std::memcpy(out, diag[level].data(), sizeof(double) * diag[level].size());
}

// coarsest Galerkin matrix with coarsest BC rows applied (before factorisation);
// Emin sets pb.params.youngsModulusMin. returns N
int ref_mg_coarsest(int nely, int nelx, int nelz, const double* E, float Emin, double* Kc_out, double* L_out, int* ok)
{
    RefState& s = state(nely, nelx, nelz);
    s.pb.params.youngsModulusMin = Emin;
    const auto& Lc = s.H.levels.back();
    const int N = 3 * Lc.numNodes;
    std::vector<double> Ev(E, E + s.pb.mesh.numElements), full, Kc;
    mg::EleMod_CompactToFull_Finest(s.pb, Ev, full);
    mg::MG_AssembleCoarsestDenseK_Galerkin(s.pb, s.H, full, s.masks.front(), Kc);
    for (int n = 0; n < Lc.numNodes; n++)
        for (int c = 0; c < 3; c++)
            if (s.masks.back()[3 * n + c]) {
                int d = 3 * n + c;
                for (int j = 0; j < N; j++) { Kc[(size_t)d * N + j] = 0.0; Kc[(size_t)j * N + d] = 0.0; }
                Kc[(size_t)d * N + d] = 1.0;
            }
    if (Kc_out) std::memcpy(Kc_out, Kc.data(), sizeof(double) * (size_t)N * N);
    bool good = mg::chol_spd_inplace(Kc, N);
    if (ok) *ok = good ? 1 : 0;
    if (L_out) std::memcpy(L_out, Kc.data(), sizeof(double) * (size_t)N * N);
    s.pb.params.youngsModulusMin = 1.0e-6f;
    return N;
}

void ref_chol_solve(const double* L, const double* b, double* x, int N)
{
    std::vector<double> Lv(L, L + (size_t)N * N), bv(b, b + N), xv;
    mg::chol_solve_lower(Lv, bv, xv, N);
    std::memcpy(x, xv.data(), sizeof(double) * N);
}

int ref_chol(double* A, int N)
{
    std::vector<double> Av(A, A + (size_t)N * N);
    bool ok = mg::chol_spd_inplace(Av, N);
    std::memcpy(A, Av.data(), sizeof(double) * (size_t)N * N);
    return ok ? 1 : 0;
}

void ref_restrict_vec3(int nely, int nelx, int nelz, int lc, const double* rf, double* rc)
{
    RefState& s = state(nely, nelx, nelz);
    const auto& Lc = s.H.levels[lc];
    const auto& Lf = s.H.levels[lc - 1];
    std::vector<double> rfv(rf, rf + 3 * (size_t)Lf.numNodes), rcv(3 * (size_t)Lc.numNodes), W(16);
    mg::MG_Restrict_nodes_Vec3_Inner(Lc, Lf, rfv, rcv, W);
    std::memcpy(rc, rcv.data(), sizeof(double) * rcv.size());
}

void ref_prolong_vec3(int nely, int nelx, int nelz, int lc, const double* xc, double* xf)
{
    RefState& s = state(nely, nelx, nelz);
    const auto& Lc = s.H.levels[lc];
    const auto& Lf = s.H.levels[lc - 1];
    std::vector<double> xcv(xc, xc + 3 * (size_t)Lc.numNodes), xfv(3 * (size_t)Lf.numNodes), W(16);
    mg::MG_Prolongate_nodes_Vec3_Inner(Lc, Lf, xcv, xfv, W);
    std::memcpy(xf, xfv.data(), sizeof(double) * xfv.size());
}

// one V-cycle z = M r with M built from E (free-DOF vectors)
void ref_mg_apply(int nely, int nelx, int nelz, const double* E, float Emin, const double* r, double* z)
{
    RefState& s = state(nely, nelx, nelz);
    s.pb.params.youngsModulusMin = Emin;
    const size_t nF = s.pb.freeDofIndex.size();
    std::vector<double> Ev(E, E + s.pb.mesh.numElements), rv(r, r + nF), zv;
    Preconditioner M = mg::make_diagonal_preconditioner_from_static(s.pb, s.H, s.masks, Ev, s.cfg);
    M(rv, zv);
    std::memcpy(z, zv.data(), sizeof(double) * nF);
    s.pb.params.youngsModulusMin = 1.0e-6f;
}

int ref_pcg(int nely, int nelx, int nelz, const double* E, float Emin, const double* b, double* x,
            double tol, int maxIt, int useMG)
{
    RefState& s = state(nely, nelx, nelz);
    s.pb.params.youngsModulusMin = Emin;
    const size_t nF = s.pb.freeDofIndex.size();
    std::vector<double> Ev(E, E + s.pb.mesh.numElements), bv(b, b + nF), xv(x, x + nF);
    Preconditioner M;
    if (useMG) M = mg::make_diagonal_preconditioner_from_static(s.pb, s.H, s.masks, Ev, s.cfg);
    PCGFreeWorkspace ws;
    int it = PCG_free(s.pb, Ev, bv, xv, tol, maxIt, M, ws);
    std::memcpy(x, xv.data(), sizeof(double) * nF);
    s.pb.params.youngsModulusMin = 1.0e-6f;
    return it;
}

void ref_filter_setup(int nely, int nelx, int nelz, float rMin, double* kernel, double* diag)
{
    RefState& s = state(nely, nelx, nelz);
    s.pf = SetupPDEFilter(s.pb, rMin);
    s.haveFilter = true;
    if (kernel) std::memcpy(kernel, s.pf.kernel.data(), sizeof(double) * 64);
    if (diag) std::memcpy(diag, s.pf.diagPrecondNode.data(), sizeof(double) * s.pf.diagPrecondNode.size());
}

void ref_filter_apply(int nely, int nelx, int nelz, const float* src, float* dst)
{
    RefState& s = state(nely, nelx, nelz);
    const int ne = s.pb.mesh.numElements;
    std::vector<float> sv(src, src + ne), dv;
    ApplyPDEFilter(s.pb, s.pf, sv, dv, s.pfws);
    std::memcpy(dst, dv.data(), sizeof(float) * ne);
}

// The OC loop of driver.cpp:23-218 (solid solve + nLoop loops), parameters from
// the caller.  Per loop: rec[6*k..] = Cdisp, C, volFrac, sharpness, change, pcgIters.
int ref_run(int nely, int nelx, int nelz, float V0, int nLoop, float rMin, float Emin, float cgTol,
            int cgMaxIt, double* rec, float* x_out, double* Csolid_out)
{
    g.reset();
    RefState& s = state(nely, nelx, nelz);
    Problem& pb = s.pb;
    pb.params.youngsModulusMin = Emin;
    pb.params.cgTol = cgTol;
    pb.params.cgMaxIt = cgMaxIt;
    PDEFilter pfFilter = SetupPDEFilter(pb, rMin);
    PDEFilterWorkspace filter_ws;
    mg::MGPrecondConfig& mgcfg = s.cfg;
    for (float& xv : pb.density) xv = V0;
    const int ne = pb.mesh.numElements;
    std::vector<float> x = pb.density;
    std::vector<float> xPhys = x;
    std::vector<double> ce(ne, 0.0);
    std::vector<double> eleMod(ne, (double)pb.params.youngsModulus);
    DOFData U;
    U.ux.assign(pb.mesh.numNodes, 0.0); U.uy.assign(pb.mesh.numNodes, 0.0); U.uz.assign(pb.mesh.numNodes, 0.0);
    std::vector<double> bFree;
    restrict_to_free(pb, pb.F, bFree);
    std::vector<double> uFreeWarm;
    PCGFreeWorkspace pcg_ws;
    Preconditioner M;
    double CsolidRef;
    {
        std::vector<double> uFree(bFree.size(), 0.0);
        auto M_solid = mg::make_diagonal_preconditioner_from_static(pb, s.H, s.masks, eleMod, mgcfg);
        PCG_free(pb, eleMod, bFree, uFree, pb.params.cgTol, pb.params.cgMaxIt, M_solid, pcg_ws);
        scatter_from_free(pb, uFree, U);
        CsolidRef = ComputeCompliance(pb, eleMod, U, ce);
        uFreeWarm = uFree;
    }
    if (Csolid_out) *Csolid_out = CsolidRef;
    int loop = 0;
    float change = 1.0f, sharpness = 1.0f;
    while (loop < nLoop && change > 1e-4 && sharpness > 0.01) {
        ++loop;
        for (int e = 0; e < ne; e++) {
            float rho = std::clamp(xPhys[e], 0.0f, 1.0f);
            float powp = static_cast<float>(std::pow(rho, pb.params.simpPenalty));
            eleMod[e] = pb.params.youngsModulusMin + powp * (pb.params.youngsModulus - pb.params.youngsModulusMin);
        }
        if (loop == 1 || (loop % 5 == 0))
            M = mg::make_diagonal_preconditioner_from_static(pb, s.H, s.masks, eleMod, mgcfg);
        int pcgIters = PCG_free(pb, eleMod, bFree, uFreeWarm, pb.params.cgTol, pb.params.cgMaxIt, M, pcg_ws);
        scatter_from_free(pb, uFreeWarm, U);
        double C = ComputeCompliance(pb, eleMod, U, ce);
        double cObjNorm = 0.0;
        for (int e = 0; e < ne; e++) cObjNorm += eleMod[e] * (ce[e] / CsolidRef);
        double Cdisp = cObjNorm * CsolidRef;
        std::vector<float> dc(ne, 0.0f);
        for (int e = 0; e < ne; e++) {
            float rho = std::clamp(xPhys[e], 0.0f, 1.0f);
            float dEdrho = pb.params.simpPenalty * static_cast<float>(std::pow(rho, pb.params.simpPenalty - 1.0f)) *
                           (pb.params.youngsModulus - pb.params.youngsModulusMin);
            float ceNorm = static_cast<float>(ce[e] / CsolidRef);
            dc[e] = -dEdrho * ceNorm;
        }
        {
            std::vector<float> xdc(ne);
            for (int e = 0; e < ne; e++) xdc[e] = x[e] * dc[e];
            std::vector<float> dc_filt;
            ApplyPDEFilter(pb, pfFilter, xdc, dc_filt, filter_ws);
            for (int e = 0; e < ne; e++) dc[e] = dc_filt[e] / std::max(1.0e-3f, x[e]);
        }
        float l1 = 0.0f, l2 = 1e9f, move = 0.2f;
        std::vector<float> xnew(ne, 0.0f);
        while ((l2 - l1) / (l1 + l2) > 1e-6) {
            float lmid = 0.5f * (l1 + l2);
            for (int e = 0; e < ne; e++) {
                float val = std::sqrt(std::max(1e-30f, -dc[e] / lmid));
                float xe = std::clamp(x[e] * val, x[e] - move, x[e] + move);
                xe = std::clamp(xe, 0.0f, 1.0f);
                xnew[e] = std::max(1.0e-3f, xe);
            }
            float vol = std::accumulate(xnew.begin(), xnew.end(), 0.0f, [](float a, float v) { return a + v; }) /
                        static_cast<float>(ne);
            if (vol - V0 > 0) l1 = lmid; else l2 = lmid;
        }
        change = 0.0f;
        for (int e = 0; e < ne; e++) change = std::max(change, std::abs(xnew[e] - x[e]));
        x.swap(xnew);
        xPhys = x;
        sharpness = 4.0f * std::accumulate(xPhys.begin(), xPhys.end(), 0.0f,
                                           [](float a, float v) { return a + v * (1.0f - v); }) / static_cast<float>(ne);
        float volFrac = std::accumulate(xPhys.begin(), xPhys.end(), 0.0f) / static_cast<float>(ne);
        double* r = rec + 6 * (loop - 1);
        r[0] = Cdisp; r[1] = C; r[2] = volFrac; r[3] = sharpness; r[4] = change; r[5] = pcgIters;
    }
    if (x_out) std::memcpy(x_out, x.data(), sizeof(float) * ne);
    return loop;
}

}  // extern "C"

// ---- CPU-baseline timing helpers (reference OpenMP code, timed in-process) ----
#include <chrono>
extern "C" {

static double median_of(std::vector<double> v)
{
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

// median wall time (ms) of `reps` K_times_u_finest applies on the given inputs
double ref_time_kxu(int nely, int nelx, int nelz, const double* E, const double* u, int reps)
{
    RefState& s = state(nely, nelx, nelz);
    const int ne = s.pb.mesh.numElements;
    const size_t nF = s.pb.freeDofIndex.size();
    std::vector<double> Ev(E, E + ne), uv(u, u + nF), yv(nF, 0.0);
    KTimesUWorkspace ws;
    K_times_u_finest(s.pb, Ev, uv, yv, ws);  // warm-up
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        K_times_u_finest(s.pb, Ev, uv, yv, ws);
        t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    return median_of(t);
}

// wall time (ms) of one MG-preconditioned PCG_free solve from x0 = 0 with
// maxIt iterations (tol 0 => runs exactly maxIt); *iters receives the count.
double ref_time_pcg(int nely, int nelx, int nelz, const double* E, float Emin, double tol, int maxIt, int* iters)
{
    RefState& s = state(nely, nelx, nelz);
    s.pb.params.youngsModulusMin = Emin;
    const size_t nF = s.pb.freeDofIndex.size();
    std::vector<double> Ev(E, E + s.pb.mesh.numElements), b, x(nF, 0.0);
    restrict_to_free(s.pb, s.pb.F, b);
    Preconditioner M = mg::make_diagonal_preconditioner_from_static(s.pb, s.H, s.masks, Ev, s.cfg);
    PCGFreeWorkspace ws;
    auto t0 = std::chrono::steady_clock::now();
    int it = PCG_free(s.pb, Ev, b, x, tol, maxIt, M, ws);
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (iters) *iters = it;
    s.pb.params.youngsModulusMin = 1.0e-6f;
    return ms;
}

// export_surface_stl (io.cpp:80-87) of the cuboid model: xPhys[ne] in compact order
int ref_export_stl(int nely, int nelx, int nelz, const float* xPhys, const char* path, float iso)
{
    RefState& s = state(nely, nelx, nelz);
    std::vector<float> x(xPhys, xPhys + s.pb.mesh.numElements);
    export_surface_stl(s.pb, x, path, iso);
    return 0;
}

// extract_faces (voxel_surface.hpp:12-93) + write_stl_binary (export_stl.hpp:14-33) of vol[ny*nx*nz]
int ref_stl_volume(int ny, int nx, int nz, const float* vol, float iso, const char* path)
{
    std::vector<float> v(vol, vol + (size_t)ny * nx * nz);
    std::vector<std::array<float, 3>> verts;
    std::vector<std::array<uint32_t, 3>> faces;
    voxsurf::extract_faces(v, ny, nx, nz, iso, verts, faces);
    return ioexp::write_stl_binary(path, verts, faces) ? 0 : 1;
}

}  // extern "C"
