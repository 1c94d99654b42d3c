// driver.cpp -- the TOP3D_XL optimisation loop on the device.
//
// Reference: TOP3D_XL_GLOBAL (cpp/src/driver.cpp:23-218) for one simulation.
// Everything between the host's per-loop bookkeeping stays on the device:
// SIMP moduli, the lazily rebuilt multigrid preconditioner (loop 1 and every
// `rebuildEvery`-th loop), the warm-started PCG solve, compliance, sensitivities,
// the PDE filter and the OC update.  Per loop the host reads back only the
// handful of scalars the loop condition and the log line need.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "t3d_internal.hpp"

using namespace t3d;

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

extern "C" {

void t3d_params_default(t3d_params* P)
{
    // GlobalParams (core.hpp:26-42), MGPrecondConfig as set by the driver
    // (driver.cpp:54), rMin = sqrt(3) from the CLI (top3d_cli.cpp:33)
    P->E0 = 1.0f; P->Emin = 1.0e-6f; P->nu = 0.3f; P->penal = 3.0f; P->cgTol = 1.0e-3f;
    P->cgMaxIt = 800; P->V0 = 0.12f; P->nLoop = 30; P->rMin = std::sqrt(3.0f);
    P->mgMaxLevels = 5; P->mgWeight = 0.6f; P->nonDyadic = 1; P->NlimitDofs = 1000;
    P->rebuildEvery = 5; P->precision = 0; P->mgMode = 0;
    P->stlPath = nullptr; P->stlIso = 0.3f; P->summaryPath = nullptr;
    P->solid = nullptr; P->passive = nullptr; P->model = nullptr;
}

// the reference's summary text (driver.cpp:225-258); iostream's scientific and
// fixed formats with the precisions it sets, as printf conversions
int t3d_summary_format(const t3d_summary* S, char* buf, size_t cap, size_t* len)
{
    T3D_CHECK(S && buf, T3D_ERR_ARG, "t3d_summary_format: null argument");
    const float avgPerIter = S->iterations > 0 ? S->sumIter / float(S->iterations) : 0.0f;
    const float pctCG = S->sumIter > 0 ? 100.0f * S->sumCG / S->sumIter : 0.0f;
    const float pctOpt = S->sumIter > 0 ? 100.0f * S->sumOpt / S->sumIter : 0.0f;
    const float pctFilt = S->sumIter > 0 ? 100.0f * S->sumFilter / S->sumIter : 0.0f;
    const int n = std::snprintf(
        buf, cap,
        "PARAMETERS\n"
        "original dims: x=%d, y=%d, z=%d\n"
        "padded dims: x=%d, y=%d, z=%d\n"
        "v0: %g\n\n"
        "SOLVING\n"
        "iterations: %d\n"
        "cg iterations: %d\n\n"
        "RESULTS\n"
        "compliance solid: %.6e\n"
        "objective first: %.6e\n"
        "objective last: %.6e\n\n"
        "TIME\n"
        "total solver time: %.6e\n"
        "time per iter: %.4f\n"
        "percentage time spent on cg: %.2f%%\n"
        "percentage time spent on optim: %.2f%%\n"
        "percentage time spent on filtering: %.2f%%\n\n"
        "THREADS\n"
        "%d\n\n"
        "NOTES\n",
        S->origX, S->origY, S->origZ, S->resX, S->resY, S->resZ, double(S->V0), S->iterations, S->cgMaxIt,
        S->Csolid, double(S->objFirst), double(S->objLast), double(S->sumCG), double(avgPerIter), double(pctCG),
        double(pctOpt), double(pctFilt), S->threads);
    T3D_CHECK(n >= 0 && size_t(n) < cap, T3D_ERR_ARG, "t3d_summary_format: buffer too small");
    if (len) *len = size_t(n);
    return T3D_OK;
}

// from filter.hip
int t3d_filter_apply_async(t3d_filter* f, const float* src, float* dst);
int t3d_filter_last_iters(const t3d_filter* f);

int t3d_top3d_run(int nely, int nelx, int nelz, const t3d_params* P, t3d_loop_record* rec, int recCap,
                  float* x_out, double* CsolidOut, int* solidIters, int* loopsRun)
{
    T3D_CHECK(P && nely > 0 && nelx > 0 && nelz > 0, T3D_ERR_ARG, "t3d_top3d_run: bad argument");
    T3D_CHECK(P->penal == 3.0f, T3D_ERR_ARG, "t3d_top3d_run: only penal = 3 is supported (powf(rho,3) rounding)");
    T3D_CHECK(P->precision == 0 || P->precision == 1, T3D_ERR_ARG,
              "t3d_top3d_run: precision must be 0 (FP64 plain path) or 1 (mixed-precision solve)");
    T3D_CHECK(P->mgMode == 0 || P->mgMode == 1, T3D_ERR_ARG,
              "t3d_top3d_run: mgMode must be 0 (the C++ port's V-cycle) or 1 (Galerkin multigrid)");
    T3D_CHECK(P->rebuildEvery > 0, T3D_ERR_ARG, "t3d_top3d_run: rebuildEvery must be positive");
    T3D_CHECK(P->nLoop >= 0 && P->cgMaxIt >= 0, T3D_ERR_ARG, "t3d_top3d_run: negative nLoop or cgMaxIt");
    const t3d_voxel_model* vm = P->model;
    T3D_CHECK(!vm || (vm->nely == nely && vm->nelx == nelx && vm->nelz == nelz), T3D_ERR_ARG,
              "t3d_top3d_run: nely/nelx/nelz differ from the voxel model's");
    // a voxel model brings its domain, boundary conditions and load, and its
    // passive elements unless P->passive is given
    const uint8_t* passiveMask = P->passive ? P->passive : (vm ? vm->passive : nullptr);
    if (loopsRun) *loopsRun = 0;
    t3d_problem* p = nullptr;
    t3d_mg* M = nullptr;
    t3d_filter* F = nullptr;
    t3d_mixed* MX = nullptr;
    int rc = vm ? t3d_problem_create_model(vm, &p) : t3d_problem_create_voxels(nely, nelx, nelz, P->solid, &p);
    if (rc) return rc;
    struct Guard {
        t3d_problem*& p;
        t3d_mg*& M;
        t3d_filter*& F;
        t3d_mixed*& MX;
        ~Guard()
        {
            t3d_mixed_destroy(MX);
            t3d_filter_destroy(F);
            t3d_mg_destroy(M);
            t3d_problem_destroy(p);
        }
    } guard{p, M, F, MX};
    if ((rc = t3d_filter_create(p, P->rMin, &F))) return rc;
    if ((rc = t3d_mg_create(p, P->nonDyadic, P->mgMaxLevels, P->mgWeight, P->Emin, P->NlimitDofs, &M))) return rc;
    if (P->mgMode == 1 && (rc = t3d_mg_set_mode(M, 1))) return rc;
    // residual replacement period of the mixed solver: the Galerkin multigrid's
    // residual falls ~1e3 in 25 iterations, where the FP32 recursion has drifted
    const int trueEvery = P->mgMode == 1 ? 25 : 50;
    // precision 1: the solves run in the mixed-precision PCG (FP32 operator and
    // V-cycle, flexible FP64 outer loop, true-residual stopping); everything
    // else is the plain path
    const bool mixed = P->precision == 1;
    const double* b_for_solve = nullptr;   // the load vector, set below
    auto setup = [&](const double* Ed) {
        int r = t3d_mg_setup(M, Ed);
        if (!r && mixed && !MX) r = t3d_mixed_create(p, M, &MX);   // needs the set-up hierarchy
        if (!r && mixed) r = t3d_mixed_setup(MX);
        return r;
    };
    auto solve = [&](double* ud, int* its) {
        return mixed ? t3d_pcg_mixed(MX, b_for_solve, ud, double(P->cgTol), P->cgMaxIt, trueEvery, 1, its)
                     : t3d_pcg(p, M, b_for_solve, ud, double(P->cgTol), P->cgMaxIt, its);
    };
    const int ne = p->ne;
    const int64_t nD = 3 * p->g.nN;
    hipStream_t st = p->stream;
    DevBuf bx, bxn, bE, bce, bxdc, bdcf, bdc, bu, bb, bsc;
    T3D_HIP(hipMalloc(&bx.p, sizeof(float) * ne));
    T3D_HIP(hipMalloc(&bxn.p, sizeof(float) * ne));
    T3D_HIP(hipMalloc(&bE.p, sizeof(double) * ne));
    T3D_HIP(hipMalloc(&bce.p, sizeof(double) * ne));
    T3D_HIP(hipMalloc(&bxdc.p, sizeof(float) * ne));
    T3D_HIP(hipMalloc(&bdcf.p, sizeof(float) * ne));
    T3D_HIP(hipMalloc(&bdc.p, sizeof(float) * ne));
    T3D_HIP(hipMalloc(&bu.p, sizeof(double) * nD));
    T3D_HIP(hipMalloc(&bb.p, sizeof(double) * nD));
    T3D_HIP(hipMalloc(&bsc.p, sizeof(double) * 16));
    // passive voxels in compact element order (matlab/TOP3D_XL.m:85, 154)
    DevBuf bpas;
    if (passiveMask) {
        const HostModel& m = p->host;
        std::vector<uint8_t> pc(size_t(ne), 0);
        for (int e = 0; e < ne; ++e) {
            const int full = m.eleMapBack[e];
            const int ez = full / (m.resX * m.resY), rem = full % (m.resX * m.resY);
            const int ex = rem / m.resY, ey = m.resY - 1 - rem % m.resY;
            pc[size_t(e)] = passiveMask[(size_t(ez) * nelx + ex) * nely + ey] ? 1 : 0;
        }
        T3D_HIP(hipMalloc(&bpas.p, size_t(ne)));
        T3D_HIP(hipMemcpy(bpas.p, pc.data(), size_t(ne), hipMemcpyHostToDevice));
    }
    const uint8_t* passive = (const uint8_t*)bpas.p;
    float* x = (float*)bx.p;
    float* xn = (float*)bxn.p;
    double* E = (double*)bE.p;
    double* ce = (double*)bce.p;
    float* xdc = (float*)bxdc.p;
    float* dcf = (float*)bdcf.p;
    float* dc = (float*)bdc.p;
    double* u = (double*)bu.p;
    double* b = (double*)bb.p;
    double* sc = (double*)bsc.p;   // [0] Csolid [1] C [2] cObj [3] design checksum; float slots from [8]
    float* fsc = reinterpret_cast<float*>(sc + 8);

    T3D_HIP(hipMemcpyAsync(b, p->bFull, sizeof(double) * nD, hipMemcpyDeviceToDevice, st));
    b_for_solve = b;
    if ((rc = launch_fill_f32(p, ne, P->V0, x))) return rc;
    // solid reference solve (driver.cpp:85-103)
    if ((rc = launch_fill_f64(p, ne, double(P->E0), E))) return rc;
    if ((rc = launch_fill_f64(p, nD, 0.0, u))) return rc;
    if ((rc = setup(E))) return rc;
    int it0 = 0;
    if ((rc = solve(u, &it0))) return rc;
    if ((rc = compliance_dev(p, E, u, ce, sc + 0))) return rc;
    double Csolid = 0.0;
    T3D_HIP(hipMemcpyAsync(&Csolid, sc + 0, sizeof(double), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));
    if (CsolidOut) *CsolidOut = Csolid;
    if (solidIters) *solidIters = it0;

    int loop = 0;
    float change = 1.0f, sharp = 1.0f;
    float objFirst = 0.0f, objLast = 0.0f, sumCG = 0.0f, sumOpt = 0.0f, sumFilter = 0.0f, sumIter = 0.0f;
    while (loop < P->nLoop && double(change) > 1e-4 && double(sharp) > 0.01) {
        const double tIt = now_s();
        ++loop;
        if ((rc = launch_simp(p, x, P->E0, P->Emin, E))) return rc;
        const double tS = now_s();
        // the Galerkin hierarchy encodes the moduli, so it is reassembled every loop, as the
        // MATLAB original does (matlab/TOP3D_XL.m:122); the C++ port's V-cycle every rebuildEvery-th
        if (P->mgMode == 1 || loop == 1 || (loop % P->rebuildEvery) == 0) {
            if ((rc = setup(E))) return rc;
        } else {
            if ((rc = t3d_set_modulus(p, E))) return rc;
        }
        int its = 0;
        if ((rc = solve(u, &its))) return rc;
        const double tSolve = now_s() - tS;
        const double tO = now_s();
        if ((rc = compliance_dev(p, E, u, ce, sc + 1))) return rc;
        if ((rc = cobj_dev(p, E, ce, sc + 0, sc + 2))) return rc;
        if ((rc = launch_sens(p, x, ce, sc + 0, P->E0, P->Emin, P->penal, xdc))) return rc;
        T3D_HIP(hipStreamSynchronize(st));
        const double tF = now_s();
        if ((rc = t3d_filter_apply_async(F, xdc, dcf))) return rc;
        T3D_HIP(hipStreamSynchronize(st));
        const double tFilter = now_s() - tF;
        if ((rc = launch_dc_post(p, dcf, x, dc))) return rc;
        if ((rc = launch_oc(p, x, dc, P->V0, 0.2f, xn, fsc, passive))) return rc;
        if ((rc = x_check_dev(p, ne, xn, reinterpret_cast<unsigned long long*>(sc + 3)))) return rc;
        double hs[4];
        float hf[5];
        T3D_HIP(hipMemcpyAsync(hs, sc, sizeof(double) * 4, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipMemcpyAsync(hf, fsc, sizeof(float) * 5, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        std::swap(x, xn);
        std::memcpy(&change, &hf[2], sizeof(float));
        sharp = hf[3];
        const float volFrac = hf[4];
        const double tOpti = now_s() - tO;
        const double Cdisp = (Csolid > 0) ? hs[2] * Csolid : hs[1];
        // the summary's figures, accumulated as the reference does (driver.cpp:203-210)
        if (loop == 1) objFirst = float(Cdisp);
        objLast = float(Cdisp);
        sumCG += float(tSolve);
        sumOpt += float(tOpti);
        sumFilter += float(tFilter);
        const double tTotal = now_s() - tIt;
        sumIter += float(tTotal);
        if (rec && loop - 1 < recCap) {
            t3d_loop_record& R = rec[loop - 1];
            R.loop = loop;
            R.pcgIters = its;
            R.filterIters = t3d_filter_last_iters(F);
            R.bisectSteps = int(hf[1]);
            R.C = hs[1];
            R.Cdisp = Cdisp;
            R.volFrac = volFrac;
            R.sharpness = sharp;
            R.change = change;
            R.fval = volFrac - P->V0;
            R.tSolve = tSolve;
            R.tOpti = tOpti;
            R.tFilter = tFilter;
            R.tTotal = tTotal;
            std::memcpy(&R.xCheck, &hs[3], sizeof(uint64_t));
        }
    }
    T3D_HIP(hipDeviceSynchronize());
    // the final design's surface (driver.cpp:272-275; xPhys = x in the C++ port, driver.cpp:186)
    if (P->stlPath && (rc = t3d_export_surface_stl(p, x, P->stlPath, P->stlIso))) return rc;
    if (P->summaryPath) {   // driver.cpp:225-270
        t3d_summary S{};
        S.origX = nelx; S.origY = nely; S.origZ = nelz;
        S.resX = p->g.resX; S.resY = p->g.resY; S.resZ = p->g.resZ;
        S.V0 = P->V0; S.iterations = loop; S.cgMaxIt = P->cgMaxIt; S.Csolid = Csolid;
        S.threads = 1;
        S.objFirst = objFirst; S.objLast = objLast;
        S.sumCG = sumCG; S.sumOpt = sumOpt; S.sumFilter = sumFilter; S.sumIter = sumIter;
        char text[2048];
        size_t n = 0;
        if ((rc = t3d_summary_format(&S, text, sizeof(text), &n))) return rc;
        FILE* fp = std::fopen(P->summaryPath, "w");
        T3D_CHECK(fp, T3D_ERR_ARG, std::string("t3d_top3d_run: cannot open ") + P->summaryPath);
        const size_t wr = std::fwrite(text, 1, n, fp);
        const int cl = std::fclose(fp);
        T3D_CHECK(wr == n && cl == 0, T3D_ERR_ARG, std::string("t3d_top3d_run: write failed: ") + P->summaryPath);
    }
    if (x_out) T3D_HIP(hipMemcpy(x_out, x, sizeof(float) * ne, hipMemcpyDeviceToHost));
    if (loopsRun) *loopsRun = loop;
    return T3D_OK;
}

}  // extern "C"
