// batch.hip -- many independent problems of uneven sizes solved together with
// the mixed-precision MG-PCG (the batching form; the single-problem solver is
// mixed.hip).
//
// Every kernel of an iteration is launched once for the whole batch: a
// device-side block map sends block b to (problem, block of that problem),
// and each problem's geometry and buffers come from a device descriptor.  The
// problems iterate in lockstep; a converged problem's blocks return at once
// (its done flag), so one launch covers every active problem.  The
// per-iteration launch sequence is captured once per p-buffer parity into a
// hipGraph and replayed; residual-replacement iterations (every `trueEvery`)
// and the true-residual confirmation of a converged problem run outside the
// graphs with the single-problem kernels.  The Walsh-form tile depth and the
// rows per block of the row kernels are chosen per batch by timing the
// candidates on the batch's own data (autotuner), unless fixed by the caller.
//
// Reference: PCG_free (cpp/src/solver.cpp:87-217) with the V-cycle
// (cpp/src/multigrid/preconditioner.cpp:88-265) per problem; the batch has no
// reference counterpart (the reference runs simulation_count problems one
// after another, cpp/src/driver.cpp:23-218).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "chol.hpp"
#include "mixed_internal.hpp"

namespace t3d {

struct BLevel {
    LevelF L;
    int span;               // restriction span into this level
    const uint8_t* mask;
    const float* Dinv;
    float* r;
    float* x;
};

// device descriptor of one problem of the batch
struct BProb {
    Grid g;
    int nlev;
    BLevel lv[8];
    float weight;
    int norm;   // restriction: 1 the reference's normalised form, 0 P^T (Galerkin multigrid)
    int Nc;
    const double* L;
    const double* LT;
    double* rc64;
    double* xc64;
    const double* Ainv;   // coarsest inverse Li^T Li (packed lower tiles, coarse.hpp)
    double* cP;           // partial products of the coarsest product
    const float* Ecell;
    const uint8_t* nodeFixed;
    float *pA, *pB, *Ap, *z, *r, *dx, *tmp;
    double* part;
    double* S;
    int* flags;
    unsigned kgx, kgy;   // Walsh-form tile grid (x, y) at the current tile depth
    unsigned kwTiles;    // tiles of the Walsh-form operator
    unsigned rowBlocks;  // blocks of the row kernels at the current rows per block
};

struct BMap {
    int prob;
    int local;
};

// ---- batched kernels: one block of a map entry per block ----
template <int NZ>
__global__ void __launch_bounds__(32 * KW_CY, KW_CY >= 16 ? 4 : 5) kb_kxw32(const BProb* __restrict__ P, const BMap* __restrict__ map,
                                               const float* __restrict__ KW, int parity)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    const unsigned loc = unsigned(m.local);
    const unsigned tx = loc % q.kgx, ty = (loc / q.kgx) % q.kgy, tz = loc / (q.kgx * q.kgy);
    const float* pprev = parity ? q.pB : q.pA;
    float* pcur = parity ? q.pA : q.pB;
    kxw32_body<NZ, true, KW_CY>(q.g, KW, q.Ecell, q.nodeFixed, pprev, q.z, pcur, q.Ap, q.part, q.S, q.flags, q.dx,
                                parity, tx, ty, tz, loc, 0);
}

__global__ void __launch_bounds__(1024) kb_fin_alpha(const BProb* __restrict__ P)
{
    const BProb& q = P[blockIdx.x];
    fin_alpha_body(q.part, q.kwTiles, q.S, q.flags, 0, 0, 0, 0, 0);
}

// rn2 ; the convergence test of iteration *itDev when check
__global__ void __launch_bounds__(1024) kb_fin_rn2(const BProb* __restrict__ P, const int* __restrict__ itDev,
                                                   int check)
{
    const BProb& q = P[blockIdx.x];
    fin_rn2_body(q.part, q.rowBlocks, q.S, q.flags, *itDev, check, 1, 0, 0, 0, 0, 0);
}

__global__ void kb_tick(int* __restrict__ itDev) { *itDev += 1; }

template <int R>
__global__ void __launch_bounds__(256) kb_xr_rows(const BProb* __restrict__ P, const BMap* __restrict__ map,
                                                  int parity, int withRestrict)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    xr_rows_body<R>(q.g.nnx, q.g.nny * q.g.nnz, q.Ap, q.r, q.part, q.S, q.lv[1].L.nnx, q.lv[1].span,
                    withRestrict ? q.tmp : nullptr, q.flags, unsigned(m.local), 0, 0, unsigned(m.local), 0);
}

template <int R>
__global__ void __launch_bounds__(256) kb_rx_rows(const BProb* __restrict__ P, const BMap* __restrict__ map)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    rx_rows_body<R>(q.g.nnx, q.g.nny * q.g.nnz, q.r, q.lv[1].L.nnx, q.lv[1].span, q.tmp, q.flags,
                    unsigned(m.local), 0, 0, 0, 0);
}

// coarse-level x-pass of the restriction from level l; local = (fz * nny + fy) * ceil(cnx/64) + bx
__global__ void __launch_bounds__(64) kb_restrict_x(const BProb* __restrict__ P, const BMap* __restrict__ map, int l)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    const LevelF& Lf = q.lv[l].L;
    const int cnx = q.lv[l + 1].L.nnx;
    const unsigned gx = unsigned((cnx + 63) / 64);
    const unsigned loc = unsigned(m.local);
    const unsigned bx = loc % gx, fy = (loc / gx) % unsigned(Lf.nny), fz = loc / (gx * unsigned(Lf.nny));
    restrict_x_body<float>(Lf, cnx, q.lv[l + 1].span, (const float*)q.lv[l].r, q.tmp, q.flags, bx, fy, fz, 0, 0);
}

// y/z pass of the restriction into level l + 1 (+ coarsest diagonal solve or FP64 copy)
__global__ void __launch_bounds__(64) kb_restrict_yz(const BProb* __restrict__ P, const BMap* __restrict__ map, int l)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    const LevelF& Lc = q.lv[l + 1].L;
    const unsigned gx = unsigned((Lc.nnx + 63) / 64);
    const unsigned loc = unsigned(m.local);
    const unsigned bx = loc % gx, cy = (loc / gx) % unsigned(Lc.nny), cz = loc / (gx * unsigned(Lc.nny));
    const int L = q.nlev - 1;
    const bool last = (l + 1 == L);
    const bool chol = q.Nc > 0;
    restrict_yz_body(q.lv[l].L, Lc, q.lv[l + 1].span, q.tmp, q.lv[l + 1].mask, q.lv[l + 1].r,
                     (last && !chol) ? q.lv[L].Dinv : nullptr, q.lv[L].x, (last && chol) ? q.rc64 : nullptr, q.flags,
                     bx, cy, cz, 0, 0, q.norm);
}

// restriction of coarse level l >= 1 into l + 1 in one pass (restrict_c32_t,
// bit-identical to kb_restrict_x + kb_restrict_yz), on the y/z pass's map
__global__ void __launch_bounds__(64) kb_restrict_c32(const BProb* __restrict__ P, const BMap* __restrict__ map, int l)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    if (q.flags && *q.flags) return;
    const LevelF& Lc = q.lv[l + 1].L;
    const unsigned gx = unsigned((Lc.nnx + 63) / 64);
    const unsigned loc = unsigned(m.local);
    const unsigned bx = loc % gx, cy = (loc / gx) % unsigned(Lc.nny), cz = loc / (gx * unsigned(Lc.nny));
    const int L = q.nlev - 1;
    const bool last = (l + 1 == L);
    const bool chol = q.Nc > 0;
    const float* Dc = (last && !chol) ? q.lv[L].Dinv : nullptr;
    double* rcd = (last && chol) ? q.rc64 : nullptr;
    if (q.lv[l + 1].span == 4)
        restrict_c32_t<4>(q.lv[l].L, Lc, (const float*)q.lv[l].r, q.lv[l + 1].mask, q.lv[l + 1].r, Dc, q.lv[L].x, rcd,
                          q.norm, bx, cy, cz);
    else
        restrict_c32_t<2>(q.lv[l].L, Lc, (const float*)q.lv[l].r, q.lv[l + 1].mask, q.lv[l + 1].r, Dc, q.lv[L].x, rcd,
                          q.norm, bx, cy, cz);
}

// coarsest solve x = Ainv rc64 (coarse.hpp): one block per packed tile of
// every listed problem, then one per row block (masked, FP64 and FP32 copies)
__global__ void __launch_bounds__(256) kb_coarse_tiles(const BProb* __restrict__ P, const BMap* __restrict__ map,
                                                       unsigned n)
{
    coarse_tiles_loop(n, [=](unsigned e) {
        const BMap m = map[e];
        const BProb& q = P[m.prob];
        return CoarseTile{q.Nc, q.Ainv, q.rc64, q.cP, unsigned(m.local), *q.flags == 0};
    });
}

__global__ void __launch_bounds__(CT) kb_coarse_sum(const BProb* __restrict__ P, const BMap* __restrict__ map)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    if (*q.flags) return;
    const int L = q.nlev - 1;
    coarse_sum_body(q.Nc, q.cP, q.lv[L].mask, q.xc64, q.lv[L].x, m.local);
}

// coarse-level prolongation + smoothing into level l (l >= 1)
__global__ void __launch_bounds__(64) kb_prolong(const BProb* __restrict__ P, const BMap* __restrict__ map, int l)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    const LevelF& Lf = q.lv[l].L;
    const unsigned gx = unsigned((Lf.nnx + 63) / 64);
    const unsigned loc = unsigned(m.local);
    const unsigned bx = loc % gx, fy = (loc / gx) % unsigned(Lf.nny), fz = loc / (gx * unsigned(Lf.nny));
    prolong_smooth32_body(Lf, q.lv[l + 1].L, q.lv[l + 1].span, q.lv[l + 1].x, q.lv[l].r, q.lv[l].Dinv, q.lv[l].mask,
                          q.weight, q.lv[l].x, q.flags, bx, fy, fz, 0, 0);
}

template <int R>
__global__ void __launch_bounds__(256) kb_pro_rows(const BProb* __restrict__ P, const BMap* __restrict__ map)
{
    const BMap m = map[blockIdx.x];
    const BProb& q = P[m.prob];
    pro_rows_body<R>(q.lv[0].L, q.lv[1].L, q.lv[1].span, q.lv[1].x, q.r,
                     q.lv[0].Dinv, q.lv[0].mask, q.weight, q.z, q.Ap, q.part, q.flags, unsigned(m.local), 0, 0, 0,
                     q.rowBlocks);
}

__global__ void __launch_bounds__(1024) kb_fin_beta(const BProb* __restrict__ P, int flexible, int first)
{
    const BProb& q = P[blockIdx.x];
    fin_beta_body(q.part, q.rowBlocks, q.S, flexible, first, q.flags, 0, 0, 0, 0, 0);
}

// per-problem initial state of a solve: stop, beta, flags
__global__ void kb_init(const BProb* __restrict__ P, double tol)
{
    const BProb& q = P[blockIdx.x];
    if (threadIdx.x == 0) {
        const double nb = sqrt(q.S[S_NORMB2]);
        q.S[S_STOP] = (1.0e-30 < nb ? nb : 1.0e-30) * tol;
        q.S[S_BETA] = 0.0;
        q.S[S_ALPHA_DX] = 0.0;
        q.S[S_PLAST] = 0.0;
        q.flags[0] = 0;
        q.flags[1] = -1;
    }
}

}  // namespace t3d

using namespace t3d;

// every problem's [done, iteration] flags into one array (one copy to the host per chunk)
__global__ void kb_gather_flags(const BProb* __restrict__ P, int n, int* __restrict__ out)
{
    const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        out[2 * i] = P[i].flags[0];
        out[2 * i + 1] = P[i].flags[1];
    }
}

struct t3d_batch {
    int n = 0;
    hipStream_t stream = nullptr;
    std::vector<t3d_problem*> P;
    std::vector<t3d_mg*> M;
    std::vector<t3d_mixed*> X;
    std::vector<hipStream_t> ownStreams;   // the problems' own streams, restored at destroy
    BProb* dP = nullptr;
    std::vector<BProb> hP;
    // block maps
    BMap* mapKW = nullptr;
    unsigned nKW = 0;
    BMap* mapRows = nullptr;
    unsigned nRows = 0;
    BMap* mapRX[8] = {};
    unsigned nRX[8] = {};
    BMap* mapYZ[8] = {};
    unsigned nYZ[8] = {};
    BMap* mapPro[8] = {};
    unsigned nPro[8] = {};
    BMap* mapCT = nullptr;   // coarsest product: (problem, packed tile)
    unsigned nCT = 0;
    BMap* mapCS = nullptr;   // coarsest product: (problem, row block)
    unsigned nCS = 0;
    int maxLev = 0;
    bool oneRestrict = true;   // every coarse level of every problem restricts in one pass (spans 2, 4)
    int* itDev = nullptr;
    int* flagsDev = nullptr;    // [2n] every problem's done flag and iteration, gathered for one copy
    int* flagsHost = nullptr;   // pinned mirror of flagsDev
    // tiles
    int kwNZ = 16, rpb = 4;
    bool tuned = false, fixedTiles = false;
    // graphs per parity
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // timing (t3d_batch_set_timing): HIP events around every kernel of the
    // directly launched lockstep iterations, and around every graph replay
    bool timing = false;
    std::vector<hipEvent_t> kev;         // kev[k], kev[k + 1] bracket kernel k of an iteration
    std::vector<std::string> kname;      // name of kernel k of the current iteration
    std::vector<std::string> tname;      // accumulated names
    std::vector<double> tms;             // accumulated milliseconds per name
    std::vector<long long> tn;           // launches per name
    std::vector<hipEvent_t> gev;         // two per graph replay of a chunk
    int gUsed = 0;
    double graphMs = 0.0;
    long long graphN = 0;
    bool setupDone = false;
    std::vector<char> active;   // problems whose blocks are in the maps
};

namespace t3d {

static void free_maps(t3d_batch* B)
{
    auto fr = [](void* p) {
        if (p) (void)hipFree(p);
    };
    fr(B->mapKW);
    fr(B->mapRows);
    for (int l = 0; l < 8; ++l) {
        fr(B->mapRX[l]);
        fr(B->mapYZ[l]);
        fr(B->mapPro[l]);
        B->mapRX[l] = B->mapYZ[l] = B->mapPro[l] = nullptr;
        B->nRX[l] = B->nYZ[l] = B->nPro[l] = 0;
    }
    fr(B->mapCT);
    fr(B->mapCS);
    B->mapKW = B->mapRows = B->mapCT = B->mapCS = nullptr;
    B->nCT = B->nCS = 0;
    for (int k = 0; k < 2; ++k)
        if (B->graph[k]) {
            (void)hipGraphExecDestroy(B->graph[k]);
            B->graph[k] = nullptr;
        }
}

static int upload_map(const std::vector<BMap>& h, BMap** d, unsigned* n)
{
    *n = unsigned(h.size());
    if (h.empty()) return T3D_OK;
    T3D_HIP(hipMalloc(d, sizeof(BMap) * h.size()));
    T3D_HIP(hipMemcpy(*d, h.data(), sizeof(BMap) * h.size(), hipMemcpyHostToDevice));
    return T3D_OK;
}

// descriptors of every problem, and block maps of the active problems for the
// current tile depth and rows per block
static int batch_build(t3d_batch* B)
{
    // no launch in flight may still read the maps or descriptors replaced here
    T3D_HIP(hipDeviceSynchronize());
    free_maps(B);
    if (int(B->active.size()) != B->n) B->active.assign(B->n, 1);
    B->hP.assign(B->n, BProb{});
    std::vector<BMap> kw, rows, rx[8], yz[8], pro[8], ct, cs;
    B->maxLev = 0;
    B->oneRestrict = true;
    for (int i = 0; i < B->n; ++i) {
        t3d_problem* p = B->P[i];
        t3d_mixed* s = B->X[i];
        MGView v;
        int rc = mg_view(B->M[i], v);
        if (rc) return rc;
        T3D_CHECK(v.nlev >= 2, T3D_ERR_STATE, "t3d_batch: every hierarchy needs at least two levels");
        T3D_CHECK(v.vtype == 0, T3D_ERR_STATE, "t3d_batch: the 'Standard' V-cycle runs in the FP64 path only");
        BProb& q = B->hP[i];
        q.g = p->g;
        q.nlev = v.nlev;
        for (int l = 0; l < v.nlev; ++l) {
            q.lv[l].L = levf(v, l);
            q.lv[l].span = v.span[l];
            q.lv[l].mask = v.mask[l];
            q.lv[l].Dinv = s->D32[l];
            q.lv[l].r = s->r32[l];
            q.lv[l].x = s->x32[l];
        }
        q.weight = float(v.weight);
        q.norm = v.norm;
        q.Nc = v.Nc;
        q.L = v.L;
        q.LT = v.LT;
        q.rc64 = s->rc64;
        q.xc64 = s->xc64;
        q.Ainv = s->Ainv;
        q.cP = s->cP;
        T3D_CHECK(v.Nc == s->NcInv, T3D_ERR_STATE, "t3d_batch: inverse coarsest factor not set up");
        q.Ecell = p->EcellF;
        q.nodeFixed = p->nodeFixed;
        q.pA = s->pA;
        q.pB = s->pB;
        q.Ap = s->Ap;
        q.z = s->z;
        q.r = s->r;
        q.dx = s->dx;
        q.tmp = s->tmp;
        q.part = s->part;
        q.S = s->sc;
        q.flags = s->iflags;
        const Grid& g = p->g;
        q.kgx = unsigned((g.nnx + wk::NX - 1) / wk::NX);
        q.kgy = unsigned((g.nny + KW_CY - 2) / (KW_CY - 1));
        const unsigned kgz = unsigned((g.nnz + B->kwNZ - 1) / B->kwNZ);
        q.kwTiles = q.kgx * q.kgy * kgz;
        q.rowBlocks = unsigned((int64_t(g.nny) * g.nnz + B->rpb - 1) / B->rpb);
        const int L = v.nlev - 1;
        B->maxLev = std::max(B->maxLev, L);
        for (int l = 2; l <= L; ++l)
            if (v.span[l] != 2 && v.span[l] != 4) B->oneRestrict = false;
        if (!B->active[i]) continue;
        for (unsigned t = 0; t < q.kwTiles; ++t) kw.push_back({i, int(t)});
        for (unsigned t = 0; t < q.rowBlocks; ++t) rows.push_back({i, int(t)});
        for (int l = 0; l < L; ++l) {
            const LevelF& Lf = q.lv[l].L;
            const LevelF& Lc = q.lv[l + 1].L;
            if (l >= 1) {
                const unsigned nbx = unsigned(int64_t(Lc.nnx + 63) / 64) * unsigned(Lf.nny) * unsigned(Lf.nnz);
                for (unsigned t = 0; t < nbx; ++t) rx[l].push_back({i, int(t)});
                const unsigned npb = unsigned((Lf.nnx + 63) / 64) * unsigned(Lf.nny) * unsigned(Lf.nnz);
                for (unsigned t = 0; t < npb; ++t) pro[l].push_back({i, int(t)});
            }
            const unsigned nyz = unsigned((Lc.nnx + 63) / 64) * unsigned(Lc.nny) * unsigned(Lc.nnz);
            for (unsigned t = 0; t < nyz; ++t) yz[l].push_back({i, int(t)});
        }
        if (v.Nc > 0) {
            for (unsigned t = 0; t < coarse_tiles(v.Nc); ++t) ct.push_back({i, int(t)});
            for (int t = 0; t < coarse_nt(v.Nc); ++t) cs.push_back({i, t});
        }
    }
    T3D_HIP(hipMemcpy(B->dP, B->hP.data(), sizeof(BProb) * B->n, hipMemcpyHostToDevice));
    int rc;
    if ((rc = upload_map(kw, &B->mapKW, &B->nKW))) return rc;
    if ((rc = upload_map(rows, &B->mapRows, &B->nRows))) return rc;
    for (int l = 0; l < 8; ++l) {
        if ((rc = upload_map(rx[l], &B->mapRX[l], &B->nRX[l]))) return rc;
        if ((rc = upload_map(yz[l], &B->mapYZ[l], &B->nYZ[l]))) return rc;
        if ((rc = upload_map(pro[l], &B->mapPro[l], &B->nPro[l]))) return rc;
    }
    if ((rc = upload_map(ct, &B->mapCT, &B->nCT))) return rc;
    if ((rc = upload_map(cs, &B->mapCS, &B->nCS))) return rc;
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}

static void launch_kw(t3d_batch* B, int parity)
{
    const dim3 grid(B->nKW), blk(32 * KW_CY);
    switch (B->kwNZ) {
    case 8: hipLaunchKernelGGL(kb_kxw32<8>, grid, blk, 0, B->stream, B->dP, B->mapKW, (const float*)B->X[0]->KW, parity); break;
    case 32: hipLaunchKernelGGL(kb_kxw32<32>, grid, blk, 0, B->stream, B->dP, B->mapKW, (const float*)B->X[0]->KW, parity); break;
    default: hipLaunchKernelGGL(kb_kxw32<16>, grid, blk, 0, B->stream, B->dP, B->mapKW, (const float*)B->X[0]->KW, parity); break;
    }
}

static void launch_xr(t3d_batch* B, int parity, int withRestrict)
{
    const dim3 grid(B->nRows), blk(256);
    // dynamic LDS: rows per block x the widest row of the batch
    size_t w = 0;
    for (const BProb& q : B->hP) w = std::max<size_t>(w, size_t(3) * q.g.nnx);
    const size_t sh = sizeof(float) * w * B->rpb;
    switch (B->rpb) {
    case 2: hipLaunchKernelGGL(kb_xr_rows<2>, grid, blk, sh, B->stream, B->dP, B->mapRows, parity, withRestrict); break;
    case 8: hipLaunchKernelGGL(kb_xr_rows<8>, grid, blk, sh, B->stream, B->dP, B->mapRows, parity, withRestrict); break;
    default: hipLaunchKernelGGL(kb_xr_rows<4>, grid, blk, sh, B->stream, B->dP, B->mapRows, parity, withRestrict); break;
    }
}

static void launch_rx(t3d_batch* B)
{
    size_t w = 0;
    for (const BProb& q : B->hP) w = std::max<size_t>(w, size_t(3) * q.g.nnx);
    const size_t sh = sizeof(float) * w * B->rpb;
    const dim3 grid(B->nRows), blk(256);
    switch (B->rpb) {
    case 2: hipLaunchKernelGGL(kb_rx_rows<2>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    case 8: hipLaunchKernelGGL(kb_rx_rows<8>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    default: hipLaunchKernelGGL(kb_rx_rows<4>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    }
}

static void launch_pro(t3d_batch* B)
{
    size_t w = 0;
    for (const BProb& q : B->hP) w = std::max<size_t>(w, size_t(3) * q.lv[1].L.nnx);
    const size_t sh = sizeof(float) * w * B->rpb;
    const dim3 grid(B->nRows), blk(256);
    switch (B->rpb) {
    case 2: hipLaunchKernelGGL(kb_pro_rows<2>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    case 8: hipLaunchKernelGGL(kb_pro_rows<8>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    default: hipLaunchKernelGGL(kb_pro_rows<4>, grid, blk, sh, B->stream, B->dP, B->mapRows); break;
    }
}

// kernel timing of a directly launched iteration (no-op unless `on`): the
// event after kernel k closes its interval, named `name`
static void tmark(t3d_batch* B, bool on, const char* name)
{
    if (!on) return;
    const size_t k = B->kname.size();
    if (B->kev.size() < k + 2) {
        hipEvent_t e;
        while (B->kev.size() < k + 2 && hipEventCreate(&e) == hipSuccess) B->kev.push_back(e);
        if (B->kev.size() < k + 2) return;
    }
    (void)hipEventRecord(B->kev[k + 1], B->stream);
    B->kname.push_back(name);
}

// batched V-cycle: z = M32(r), rz and beta per problem
static void batch_vcycle(t3d_batch* B, int flexible, int first, bool tmpReady, bool tm = false)
{
    hipStream_t st = B->stream;
    if (!tmpReady) launch_rx(B);
    for (int l = 0; l < B->maxLev; ++l) {
        if (l >= 1 && B->oneRestrict) {
            if (B->nYZ[l]) {
                hipLaunchKernelGGL(kb_restrict_c32, dim3(B->nYZ[l]), dim3(64), 0, st, (const BProb*)B->dP,
                                   (const BMap*)B->mapYZ[l], l);
                tmark(B, tm, "kb_restrict_c32");
            }
            continue;
        }
        if (l >= 1 && B->nRX[l]) {
            hipLaunchKernelGGL(kb_restrict_x, dim3(B->nRX[l]), dim3(64), 0, st, (const BProb*)B->dP,
                               (const BMap*)B->mapRX[l], l);
            tmark(B, tm, "kb_restrict_x");
        }
        if (B->nYZ[l]) {
            hipLaunchKernelGGL(kb_restrict_yz, dim3(B->nYZ[l]), dim3(64), 0, st, (const BProb*)B->dP,
                               (const BMap*)B->mapYZ[l], l);
            tmark(B, tm, "kb_restrict_yz");
        }
    }
    if (B->nCT) {
        hipLaunchKernelGGL(kb_coarse_tiles, dim3(coarse_tile_blocks(B->nCT)), dim3(256), 0, st, (const BProb*)B->dP,
                           (const BMap*)B->mapCT, B->nCT);
        tmark(B, tm, "kb_coarse_tiles");
        hipLaunchKernelGGL(kb_coarse_sum, dim3(B->nCS), dim3(CT), 0, st, (const BProb*)B->dP, (const BMap*)B->mapCS);
        tmark(B, tm, "kb_coarse_sum");
    }
    for (int l = B->maxLev - 1; l >= 1; --l)
        if (B->nPro[l]) {
            hipLaunchKernelGGL(kb_prolong, dim3(B->nPro[l]), dim3(64), 0, st, (const BProb*)B->dP,
                               (const BMap*)B->mapPro[l], l);
            tmark(B, tm, "kb_prolong");
        }
    launch_pro(B);
    tmark(B, tm, "kb_pro_rows");
    hipLaunchKernelGGL(kb_fin_beta, dim3(B->n), dim3(1024), 0, st, (const BProb*)B->dP, flexible, first);
    tmark(B, tm, "kb_fin_beta");
}

// one lockstep iteration of every active problem (no residual replacement);
// tm: time each kernel (direct launches only, never under capture)
static void batch_iteration(t3d_batch* B, int parity, int flexible, bool tm = false)
{
    hipStream_t st = B->stream;
    if (tm) {
        B->kname.clear();
        tmark(B, true, "");   // kev[1]: the start
        B->kname.clear();
        std::swap(B->kev[0], B->kev[1]);
    }
    launch_kw(B, parity);
    tmark(B, tm, "kb_kxw32");
    hipLaunchKernelGGL(kb_fin_alpha, dim3(B->n), dim3(1024), 0, st, (const BProb*)B->dP);
    tmark(B, tm, "kb_fin_alpha");
    launch_xr(B, parity, 1);
    tmark(B, tm, "kb_xr_rows");
    hipLaunchKernelGGL(kb_fin_rn2, dim3(B->n), dim3(1024), 0, st, (const BProb*)B->dP, (const int*)B->itDev, 1);
    tmark(B, tm, "kb_fin_rn2");
    batch_vcycle(B, flexible, 0, true, tm);
    hipLaunchKernelGGL(kb_tick, dim3(1), dim3(1), 0, st, B->itDev);
    if (tm && hipEventSynchronize(B->kev[B->kname.size()]) == hipSuccess) {
        for (size_t k = 0; k < B->kname.size(); ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, B->kev[k], B->kev[k + 1]) != hipSuccess) continue;
            size_t j = 0;
            while (j < B->tname.size() && B->tname[j] != B->kname[k]) ++j;
            if (j == B->tname.size()) {
                B->tname.push_back(B->kname[k]);
                B->tms.push_back(0.0);
                B->tn.push_back(0);
            }
            B->tms[j] += ms;
            B->tn[j] += 1;
        }
    }
}

}  // namespace t3d

extern "C" {

int t3d_batch_destroy(t3d_batch* B)
{
    if (!B) return T3D_OK;
    (void)hipDeviceSynchronize();
    free_maps(B);
    if (B->dP) (void)hipFree(B->dP);
    if (B->itDev) (void)hipFree(B->itDev);
    if (B->flagsDev) (void)hipFree(B->flagsDev);
    if (B->flagsHost) (void)hipHostFree(B->flagsHost);
    for (hipEvent_t e : B->kev) (void)hipEventDestroy(e);
    for (hipEvent_t e : B->gev) (void)hipEventDestroy(e);
    for (int i = 0; i < B->n; ++i) {
        if (B->P[i]) B->P[i]->stream = B->ownStreams[i];
        t3d_mixed_destroy(B->X[i]);
        t3d_mg_destroy(B->M[i]);
        t3d_problem_destroy(B->P[i]);
    }
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return T3D_OK;
}

int t3d_batch_create(int n, const int* nely, const int* nelx, const int* nelz, float Emin, t3d_batch** out)
{
    return t3d_batch_create_voxels(n, nely, nelx, nelz, nullptr, Emin, out);
}

int t3d_batch_create_voxels(int n, const int* nely, const int* nelx, const int* nelz, const uint8_t* const* solid,
                            float Emin, t3d_batch** out)
{
    T3D_CHECK(n > 0 && n <= 4096 && nely && nelx && nelz && out, T3D_ERR_ARG, "t3d_batch_create: bad argument");
    *out = nullptr;
    t3d_batch* B = new t3d_batch();
    B->n = n;
    B->P.assign(n, nullptr);
    B->M.assign(n, nullptr);
    B->X.assign(n, nullptr);
    B->ownStreams.assign(n, nullptr);
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        set_error("t3d_batch_create: stream creation failed");
        return T3D_ERR_HIP;
    }
    int rc = T3D_OK;
    for (int i = 0; i < n && !rc; ++i) {
        rc = t3d_problem_create_voxels(nely[i], nelx[i], nelz[i], solid ? solid[i] : nullptr, &B->P[i]);
        if (rc) break;
        // every problem of the batch works on the batch's stream
        B->ownStreams[i] = B->P[i]->stream;
        B->P[i]->stream = B->stream;
        rc = t3d_mg_create(B->P[i], 1, 5, 0.6f, Emin, 1000, &B->M[i]);
    }
    if (!rc && (hipMalloc(&B->dP, sizeof(BProb) * n) != hipSuccess || hipMalloc(&B->itDev, sizeof(int)) != hipSuccess)) {
        set_error("t3d_batch_create: allocation failed");
        rc = T3D_ERR_HIP;
    }
    if (rc) {
        t3d_batch_destroy(B);
        return rc;
    }
    *out = B;
    return T3D_OK;
}

int t3d_batch_problem(t3d_batch* B, int i, t3d_problem** p)
{
    T3D_CHECK(B && p && i >= 0 && i < B->n, T3D_ERR_ARG, "t3d_batch_problem: bad argument");
    *p = B->P[i];
    return T3D_OK;
}

int t3d_batch_mg(t3d_batch* B, int i, t3d_mg** g)
{
    T3D_CHECK(B && g && i >= 0 && i < B->n, T3D_ERR_ARG, "t3d_batch_mg: bad argument");
    *g = B->M[i];
    return T3D_OK;
}

// per-problem element moduli (device pointers, compact order): multigrid set-up
int t3d_batch_setup(t3d_batch* B, const double* const* E_dev)
{
    T3D_CHECK(B && E_dev, T3D_ERR_ARG, "t3d_batch_setup: null argument");
    int rc;
    for (int i = 0; i < B->n; ++i) {
        if ((rc = t3d_mg_setup(B->M[i], E_dev[i]))) return rc;
        if (!B->X[i] && (rc = t3d_mixed_create(B->P[i], B->M[i], &B->X[i]))) return rc;
        if ((rc = t3d_mixed_setup(B->X[i]))) return rc;
    }
    B->setupDone = true;
    return batch_build(B);
}

// kwNZ in {8, 16, 32}, rowsPerBlock in {2, 4, 8}; 0, 0 = autotune on the next solve
int t3d_batch_set_tiles(t3d_batch* B, int kwNZ, int rowsPerBlock)
{
    T3D_CHECK(B, T3D_ERR_ARG, "t3d_batch_set_tiles: null");
    if (kwNZ == 0 && rowsPerBlock == 0) {
        B->fixedTiles = false;
        B->tuned = false;
        return T3D_OK;
    }
    T3D_CHECK((kwNZ == 8 || kwNZ == 16 || kwNZ == 32) && (rowsPerBlock == 2 || rowsPerBlock == 4 || rowsPerBlock == 8),
              T3D_ERR_ARG, "t3d_batch_set_tiles: kwNZ in {8,16,32}, rowsPerBlock in {2,4,8}");
    B->kwNZ = kwNZ;
    B->rpb = rowsPerBlock;
    B->fixedTiles = true;
    B->tuned = true;
    return B->setupDone ? batch_build(B) : T3D_OK;
}

// timing of the next solves: on != 0 zeroes the totals and creates the events
int t3d_batch_set_timing(t3d_batch* B, int on)
{
    T3D_CHECK(B, T3D_ERR_ARG, "t3d_batch_set_timing: null");
    B->timing = on != 0;
    B->tname.clear();
    B->tms.clear();
    B->tn.clear();
    B->graphMs = 0.0;
    B->graphN = 0;
    B->gUsed = 0;
    if (on)
        while (B->gev.size() < 2 * 32) {   // two per replay of a chunk of iterations
            hipEvent_t e;
            T3D_HIP(hipEventCreate(&e));
            B->gev.push_back(e);
        }
    return T3D_OK;
}

int t3d_batch_get_timing(const t3d_batch* B, int k, char* name, int nameLen, double* ms_total, long long* launches)
{
    T3D_CHECK(B && ms_total && launches, T3D_ERR_ARG, "t3d_batch_get_timing: null");
    if (k < 0) {   // the graph replays
        *ms_total = B->graphMs;
        *launches = B->graphN;
        if (name && nameLen > 0) std::snprintf(name, size_t(nameLen), "%s", "graph");
        return int(B->tname.size());
    }
    T3D_CHECK(k < int(B->tname.size()), T3D_ERR_ARG, "t3d_batch_get_timing: no such kernel");
    *ms_total = B->tms[size_t(k)];
    *launches = B->tn[size_t(k)];
    if (name && nameLen > 0) std::snprintf(name, size_t(nameLen), "%s", B->tname[size_t(k)].c_str());
    return int(B->tname.size());
}

int t3d_batch_info(const t3d_batch* B, int* kwNZ, int* rowsPerBlock)
{
    T3D_CHECK(B, T3D_ERR_ARG, "t3d_batch_info: null");
    if (kwNZ) *kwNZ = B->kwNZ;
    if (rowsPerBlock) *rowsPerBlock = B->rpb;
    return T3D_OK;
}

// Solves K_i x_i = b_i for every problem (device pointers, full layout, FP64;
// x_i is the initial guess and the result).  iters[i] = iterations of problem
// i (maxIt when not converged).  useGraph: replay the captured iterations.
int t3d_batch_solve(t3d_batch* B, const double* const* b, double* const* x, double tol, int maxIt, int trueEvery,
                    int flexible, int useGraph, int* iters)
{
    T3D_CHECK(B && b && x && iters && maxIt >= 0, T3D_ERR_ARG, "t3d_batch_solve: bad argument");
    T3D_CHECK(B->setupDone, T3D_ERR_STATE, "t3d_batch_solve: t3d_batch_setup not called");
    hipStream_t st = B->stream;
    const int n = B->n;
    int rc;
    if (std::count(B->active.begin(), B->active.end(), 1) != n) {
        B->active.assign(n, 1);
        if ((rc = batch_build(B))) return rc;
    }
    // per-problem start: r = b - K64 x, |b|, V-cycle of the first iteration
    for (int i = 0; i < n; ++i) {
        t3d_problem* p = B->P[i];
        t3d_mixed* s = B->X[i];
        const int64_t nD = 3 * p->g.nN;
        if ((rc = canon_dot_dev(p, b[i], b[i], nD, s->sc + S_NORMB2))) return rc;
        if ((rc = mixed_true_residual(s, b[i], x[i]))) return rc;
        T3D_HIP(hipMemsetAsync(s->Ap, 0, sizeof(float) * nD, st));
        T3D_HIP(hipMemsetAsync(s->pA, 0, sizeof(float) * nD, st));
        T3D_HIP(hipMemsetAsync(s->dx, 0, sizeof(float) * nD, st));
    }
    hipLaunchKernelGGL(kb_init, dim3(n), dim3(64), 0, st, (const BProb*)B->dP, tol);
    T3D_HIP(hipMemsetAsync(B->itDev, 0, sizeof(int), st));
    batch_vcycle(B, flexible, 1, false);
    T3D_HIP(hipGetLastError());
    // autotuner: time the candidate tiles of the Walsh-form kernel and of the
    // row kernels on this batch (the timed launches rewrite p, Ap, z and the
    // partials with the values the first iteration recomputes)
    if (!B->tuned) {
        hipEvent_t e0, e1;
        T3D_HIP(hipEventCreate(&e0));
        T3D_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        int bestNZ = 16, bestR = 4;
        for (int nz : {8, 16, 32}) {
            B->kwNZ = nz;
            if ((rc = batch_build(B))) return rc;
            launch_kw(B, 0);
            T3D_HIP(hipEventRecord(e0, st));
            for (int k = 0; k < 3; ++k) launch_kw(B, 0);
            T3D_HIP(hipEventRecord(e1, st));
            T3D_HIP(hipEventSynchronize(e1));
            float ms = 0;
            T3D_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bestNZ = nz;
            }
        }
        B->kwNZ = bestNZ;
        best = 1e30f;
        for (int r : {2, 4, 8}) {
            B->rpb = r;
            if ((rc = batch_build(B))) return rc;
            launch_pro(B);
            T3D_HIP(hipEventRecord(e0, st));
            for (int k = 0; k < 3; ++k) launch_pro(B);
            T3D_HIP(hipEventRecord(e1, st));
            T3D_HIP(hipEventSynchronize(e1));
            float ms = 0;
            T3D_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bestR = r;
            }
        }
        B->rpb = bestR;
        if ((rc = batch_build(B))) return rc;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        B->tuned = true;
        // the last candidate's pro_rows left its own partial layout: redo the
        // first V-cycle's finest level with the chosen tiles
        batch_vcycle(B, flexible, 1, false);
    }
    // graphs of one iteration per parity (re-captured after the maps change)
    auto ensure_graphs = [&]() -> int {
        if (!useGraph) return T3D_OK;
        for (int par = 0; par < 2; ++par) {
            if (B->graph[par]) continue;
            hipGraph_t gr;
            T3D_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            batch_iteration(B, par, flexible);
            T3D_HIP(hipStreamEndCapture(st, &gr));
            T3D_HIP(hipGraphInstantiate(&B->graph[par], gr, nullptr, nullptr, 0));
            T3D_HIP(hipGraphDestroy(gr));
        }
        return T3D_OK;
    };
    if ((rc = ensure_graphs())) return rc;
    std::vector<int> result(n, maxIt), skipped(n, 0), finished(n, 0);
    std::vector<int> hf(2 * n);
    const int CHUNK = 32;
    int it0 = 0;
    while (it0 < maxIt) {
        const int it1 = std::min(maxIt, it0 + CHUNK);
        for (int it = it0; it < it1; ++it) {
            const int par = it & 1;
            const bool replace = trueEvery > 0 && ((it + 1) % trueEvery) == 0;
            if (!replace) {
                if (useGraph) {
                    const bool tg = B->timing && B->gUsed + 2 <= int(B->gev.size());
                    if (tg) T3D_HIP(hipEventRecord(B->gev[B->gUsed], st));
                    T3D_HIP(hipGraphLaunch(B->graph[par], st));
                    if (tg) {
                        T3D_HIP(hipEventRecord(B->gev[B->gUsed + 1], st));
                        B->gUsed += 2;
                    }
                } else {
                    batch_iteration(B, par, flexible, B->timing);
                }
            } else {
                launch_kw(B, par);
                hipLaunchKernelGGL(kb_fin_alpha, dim3(n), dim3(1024), 0, st, (const BProb*)B->dP);
                launch_xr(B, par, 0);
                hipLaunchKernelGGL(kb_fin_rn2, dim3(n), dim3(1024), 0, st, (const BProb*)B->dP,
                                   (const int*)B->itDev, 0);
                for (int i = 0; i < n; ++i)
                    if (!finished[i] && (rc = mixed_replace_residual(B->X[i], b[i], x[i], it))) return rc;
                batch_vcycle(B, flexible, 0, false);
                hipLaunchKernelGGL(kb_tick, dim3(1), dim3(1), 0, st, B->itDev);
            }
        }
        if (!B->flagsDev) {
            T3D_HIP(hipMalloc(&B->flagsDev, sizeof(int) * 2 * size_t(n)));
            T3D_HIP(hipHostMalloc(&B->flagsHost, sizeof(int) * 2 * size_t(n)));
        }
        hipLaunchKernelGGL(kb_gather_flags, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, (const BProb*)B->dP,
                           n, B->flagsDev);
        T3D_HIP(hipGetLastError());
        T3D_HIP(hipMemcpyAsync(B->flagsHost, B->flagsDev, sizeof(int) * 2 * size_t(n), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        for (int k = 0; k + 1 < B->gUsed; k += 2) {   // graph replays of this chunk
            float ms = 0.f;
            T3D_HIP(hipEventElapsedTime(&ms, B->gev[k], B->gev[k + 1]));
            B->graphMs += ms;
            B->graphN += 1;
        }
        B->gUsed = 0;
        bool any = false, allDone = true;
        for (int i = 0; i < n; ++i) {
            if (finished[i]) continue;
            hf[2 * i] = B->flagsHost[2 * i];
            hf[2 * i + 1] = B->flagsHost[2 * i + 1];
            if (hf[2 * i]) any = true;
        }
        if (any) {
            for (int i = 0; i < n; ++i) {
                if (finished[i] || !hf[2 * i]) continue;
                const int itDone = hf[2 * i + 1] - 1;   // global iteration index
                double tr = 0, stop = 0;
                if ((rc = mixed_confirm(B->X[i], b[i], x[i], &tr, &stop))) return rc;
                if (std::sqrt(tr) < stop) {
                    finished[i] = 1;
                    result[i] = itDone + 1 - skipped[i];
                } else {
                    // resume: idle iterations until it1 do not count; align p parity
                    skipped[i] += it1 - (itDone + 1);
                    if ((rc = mixed_resume(B->X[i], itDone, it1, flexible))) return rc;
                }
            }
        }
        for (int i = 0; i < n; ++i) allDone = allDone && finished[i];
        if (allDone) break;
        // drop finished problems from the block maps once they hold a quarter
        // of the operator's blocks (their blocks would only return early)
        unsigned activeTiles = 0;
        for (int i = 0; i < n; ++i)
            if (!finished[i]) activeTiles += B->hP[i].kwTiles;
        if (4 * activeTiles <= 3 * B->nKW) {
            for (int i = 0; i < n; ++i) B->active[i] = finished[i] ? 0 : 1;
            if ((rc = batch_build(B))) return rc;
            if ((rc = ensure_graphs())) return rc;
        }
        it0 = it1;
    }
    for (int i = 0; i < n; ++i) {
        // a finished problem has folded its last increment at its confirmation
        if (!finished[i]) {
            result[i] = maxIt - skipped[i];
            if ((rc = mixed_fold(B->X[i], x[i]))) return rc;
        }
        iters[i] = result[i];
    }
    T3D_HIP(hipStreamSynchronize(st));
    return T3D_OK;
}

}  // extern "C"
