// oc.hip -- element-wise SIMP / sensitivity / OC kernels, the compliance
// kernel and the exact sequential float sums of the OC update.
//
// Reference: driver.cpp:117-189 (SIMP, dc, filter pre/post scaling, OC
// bisection, change, sharpness, volume fraction) and ComputeCompliance
// (fea.cpp:587-677, 851-864).  Float semantics follow the reference exactly:
// float arithmetic without contraction, correctly rounded sqrt and division
// (plain sqrtf and '/': clang's HIP default is correctly rounded for both --
// the __fsqrt_rn intrinsic is the native approximation unless OCML basic
// rounded operations are enabled, so it is not used),
// std::clamp / std::max selection rules, and std::accumulate<float> sums
// reproduced bit-for-bit by the transducer scan of seqsum.hpp.  powf(rho,3)
// and powf(rho,2) are evaluated correctly rounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "seqsum.hpp"
#include "t3d_internal.hpp"

namespace t3d {

__device__ __forceinline__ float fmax_r(float a, float b) { return (a < b) ? b : a; }          // std::max
__device__ __forceinline__ float fclamp_r(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// correctly rounded x^3 and x^2 for binary32 x (exact double-double product, one rounding)
__device__ __forceinline__ float next_float(float f, bool up)
{
    unsigned b = __float_as_uint(f);
    b = up ? b + 1u : (b == 0u ? 0x80000001u : b - 1u);
    return __uint_as_float(b);
}
__device__ __forceinline__ float cr_cube(float x)
{
    const double d = double(x);
    const double p2 = d * d;
    const double hi = p2 * d;
    const double lo = __fma_rn(p2, d, -hi);
    const float f = float(hi);
    const double df = double(f);
    if (lo == 0.0 || df == hi) return f;
    const float g = next_float(f, hi > df);
    const double mid = 0.5 * (df + double(g));
    if (hi != mid) return f;
    return ((lo > 0.0) == (double(g) > df)) ? g : f;
}
__device__ __forceinline__ float cr_square(float x)
{
    const double d = double(x);
    return float(d * d);
}

// ---------------------------------------------------------------------------
// SIMP (driver.cpp:117-121): E = (double)(Emin + rho^3 * (E0 - Emin)), float ops
// ---------------------------------------------------------------------------
__global__ void k_simp(int ne, const float* __restrict__ x, float E0, float Emin, double* __restrict__ E)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const float rho = fclamp_r(x[e], 0.0f, 1.0f);
    const float pw = cr_cube(rho);
    const float v = __fadd_rn(Emin, __fmul_rn(pw, __fsub_rn(E0, Emin)));
    E[e] = double(v);
}

__global__ void k_fill_f64(int64_t n, double v, double* __restrict__ a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_f32(int64_t n, float v, float* __restrict__ a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// sensitivities and the filter input (driver.cpp:151-163):
//   dc = -(penal * rho^2 * (E0-Emin)) * (float)(ce/Csolid) ; xdc = x*dc
// ---------------------------------------------------------------------------
__global__ void k_sens_xdc(int ne, const float* __restrict__ x, const double* __restrict__ ce,
                           const double* __restrict__ Csolid, float E0, float Emin, float penal,
                           float* __restrict__ xdc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double Cs = *Csolid;
    const float rho = fclamp_r(x[e], 0.0f, 1.0f);
    const float dE = __fmul_rn(__fmul_rn(penal, cr_square(rho)), __fsub_rn(E0, Emin));
    const float ceN = (Cs > 0) ? float(ce[e] / Cs) : float(ce[e]);
    const float dc = __fmul_rn(-dE, ceN);
    xdc[e] = __fmul_rn(x[e], dc);
}

// dc = filt / max(1e-3, x) (driver.cpp:165)
__global__ void k_dc_post(int ne, const float* __restrict__ filt, const float* __restrict__ x, float* __restrict__ dc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    dc[e] = ((filt[e]) / (fmax_r(1.0e-3f, x[e])));
}

// OC candidate density for multiplier lmid (driver.cpp:175-180)
__device__ __forceinline__ float oc_xnew(float x, float dc, float lmid, float move)
{
    const float val = sqrtf(fmax_r(1e-30f, ((-dc) / (lmid))));
    float xe = fclamp_r(__fmul_rn(x, val), __fsub_rn(x, move), __fadd_rn(x, move));
    xe = fclamp_r(xe, 0.0f, 1.0f);
    return fmax_r(1.0e-3f, xe);
}

// ---------------------------------------------------------------------------
// block-wide exact sequential float sum (1024 threads, see seqsum.hpp)
// ---------------------------------------------------------------------------
constexpr int SS_THREADS = 1024;
constexpr int SS_PER = 8;
constexpr int SS_CHUNK = SS_THREADS * SS_PER;
constexpr int SS_PREFIX = 1024;

struct SeqShared {
    seqsum::Tr waveTot[SS_THREADS / 64];
    unsigned long long cross;
    float s;
    long long i0;
};

__device__ __forceinline__ seqsum::Tr shfl_up_tr(const seqsum::Tr& t, int d)
{
    seqsum::Tr o;
    o.r0 = uint32_t(__shfl_up(int(t.r0), d, 64));
    o.r1 = uint32_t(__shfl_up(int(t.r1), d, 64));
    return o;
}

template <class Gen>
__device__ float block_seq_sum(int64_t n, const Gen& gen, SeqShared& sh)
{
    using namespace seqsum;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        float s = 0.0f;
        int64_t i = 0;
        for (; i < n && (i < SS_PREFIX || !(s >= 1e-30f)); ++i) s = __fadd_rn(s, gen(i));
        sh.s = s;
        sh.i0 = i;
    }
    __syncthreads();
    while (sh.i0 < n) {
        const int64_t i0 = sh.i0;
        const float s = sh.s;
        int64_t m;
        int ue;
        decompose(s, m, ue);
        float v[SS_PER];
        Tr F = tr_identity();
#pragma unroll
        for (int k = 0; k < SS_PER; ++k) {
            const int64_t idx = i0 + int64_t(tid) * SS_PER + k;
            v[k] = (idx < n) ? gen(idx) : 0.0f;
            F = tr_compose(F, tr_elem(v[k], ue));
        }
        // inclusive wave scan
        Tr inc = F;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Tr o = shfl_up_tr(inc, d);
            if (lane >= d) inc = tr_compose(o, inc);
        }
        if (lane == 63) sh.waveTot[wave] = inc;
        if (tid == 0) sh.cross = ~0ull;
        __syncthreads();
        if (tid == 0) {
            Tr acc = tr_identity();
            for (int w = 0; w < SS_THREADS / 64; ++w) {
                const Tr t = sh.waveTot[w];
                sh.waveTot[w] = acc;   // exclusive wave prefix
                acc = tr_compose(acc, t);
            }
            // chunk total kept in the cross-free state update below
            const int64_t mt = m + tr_apply(acc, m);
            sh.s = (mt < M_TOP) ? ldexpf(float(mt), ue) : s;   // provisional; replaced on a crossing
        }
        __syncthreads();
        Tr excl = shfl_up_tr(inc, 1);
        if (lane == 0) excl = tr_identity();
        const Tr P = tr_compose(sh.waveTot[wave], excl);
        int64_t mm = m + tr_apply(P, m);
        unsigned long long myCross = ~0ull;
        int64_t mmc = 0;
        float vc = 0.0f;
#pragma unroll
        for (int k = 0; k < SS_PER; ++k) {
            const int64_t idx = i0 + int64_t(tid) * SS_PER + k;
            if (idx < n && myCross == ~0ull) {
                const int64_t r = tr_apply(tr_elem(v[k], ue), mm);
                if (mm + r >= M_TOP) {
                    myCross = (unsigned long long)idx;
                    mmc = mm;
                    vc = v[k];
                } else {
                    mm += r;
                }
            }
        }
        if (myCross != ~0ull) atomicMin(&sh.cross, myCross);
        __syncthreads();
        const unsigned long long cr = sh.cross;
        if (cr == ~0ull) {
            if (tid == 0) sh.i0 = i0 + SS_CHUNK;
        } else if (myCross == cr) {
            sh.s = __fadd_rn(ldexpf(float(mmc), ue), vc);
            sh.i0 = int64_t(cr) + 1;
        }
        __syncthreads();
    }
    const float res = sh.s;
    __syncthreads();
    return res;
}

struct GenX {
    const float* x;
    __device__ float operator()(int64_t i) const { return x[i]; }
};

// ---------------------------------------------------------------------------
// The OC bisection and the statistics of xnew, each step an exact
// std::accumulate<float> over all elements, spread over the GPU.  One sum:
//   k_ocs_start  adds the first OC_PREFIX elements one by one (more while the
//                sum is below 1e-30), so the running sum is a normal float s0
//                before element i0; the bisection step's multiplier is set here;
//   k_cs_sum     double sum of every chunk of OC_CH elements after i0;
//   k_cs_bounds  prefix of the chunk sums.  Every float addition of
//                non-negative terms rounds by a factor in [1-u, 1+u], u = 2^-24,
//                so before chunk c (n = c OC_CH additions after s0) the running
//                sum lies in [P (1-u)^n, P (1+u)^n], P = s0 + the exact sum of
//                the elements before it; the binades of that interval are the
//                chunk's candidates (at most OC_KMAX, else none).  A bisection
//                step whose interval for the whole sum already decides
//                vol > V0 ends here;
//   k_cs_agg     per chunk and candidate binade, the transducer aggregate of the
//                chunk (seqsum.hpp);
//   k_cs_walk    one workgroup walks the chunks in order, composing the
//                aggregates for the running sum's binade over windows of OC_WIN
//                chunks; the first chunk that leaves the binade, or that has no
//                aggregate for it, is summed element by element (a workgroup
//                scan per binade crossing, each crossing one real float add),
//                and the walk goes on with the next chunk.
// The candidates only decide which aggregates exist: a chunk without the one the
// walk needs is summed element by element, so the sum is exact whatever the
// bounds.  Every kernel returns at once when the bisection has ended (B->done)
// or the sum is complete (B->ssDone), so the host launches OC_STEPS steps
// between checks.
// ---------------------------------------------------------------------------
constexpr int OC_CT = 256;                 // threads of a chunk workgroup
constexpr int OC_CPER = 8;                 // elements per thread
constexpr int OC_CH = OC_CT * OC_CPER;     // elements per chunk
constexpr int OC_KMAX = 8;                 // candidate binades per chunk
constexpr int OC_BT = 1024;                // threads of k_cs_bounds
constexpr int OC_WT = 1024;                // threads of the walk
constexpr int OC_WCPT = 1;                 // chunks per walk thread
constexpr int OC_WIN = OC_WT * OC_WCPT;    // chunks per window of the walk
constexpr int OC_EPT = OC_CH / OC_WT;      // elements per walk thread in a chunk
constexpr int OC_PREFIX = 256;             // elements added one by one first
constexpr int OC_STEPS = 4;                // bisection steps launched between host checks
constexpr int OC_APPLY_GRID = 2048;        // workgroups of k_oc_apply (grid-stride)

struct OcBis {
    float l1, l2, lmid;
    int steps, done, err;
    float s;          // the running float sum
    int ssDone;       // the sum is complete, or the step decided by its bounds
    long long i0;     // first element after the prefix
    int decided;      // 1: vol > V0, 2: vol <= V0, known from the bounds alone
};

// the summed elements: 0 = OC densities at B->lmid (passive elements, when the
// nullable mask is given, held at 1: matlab/TOP3D_XL.m:154), 1 = x, 2 = x (1 - x)
__device__ __forceinline__ float oc_gen(int mode, int64_t i, const float* __restrict__ x, const float* __restrict__ dc,
                                        const uint8_t* __restrict__ passive, float lmid, float move)
{
    if (mode == 1) return x[i];
    if (mode == 2) return __fmul_rn(x[i], __fsub_rn(1.0f, x[i]));
    return (passive && passive[i]) ? 1.0f : oc_xnew(x[i], dc[i], lmid, move);
}

// Wave scans of transducers with DPP row shifts (row_shr:n, GFX9 control
// 0x110 + n) inside each row of 16 lanes and the row totals read with
// readlane: no LDS round trips, unlike __shfl_up (ds_bpermute).
template <int CTRL>
__device__ __forceinline__ seqsum::Tr dpp_tr(const seqsum::Tr& t)
{
    return seqsum::Tr{uint32_t(__builtin_amdgcn_update_dpp(0, int(t.r0), CTRL, 0xf, 0xf, false)),
                      uint32_t(__builtin_amdgcn_update_dpp(0, int(t.r1), CTRL, 0xf, 0xf, false))};
}
__device__ __forceinline__ seqsum::Tr readlane_tr(const seqsum::Tr& t, int l)
{
    return seqsum::Tr{uint32_t(__builtin_amdgcn_readlane(int(t.r0), l)),
                      uint32_t(__builtin_amdgcn_readlane(int(t.r1), l))};
}
// inclusive (inc) and exclusive (ex) scan of T over the 64 lanes in lane order;
// every lane of the wave takes part
__device__ __forceinline__ void wave_scan_tr(const seqsum::Tr& T, seqsum::Tr& inc, seqsum::Tr& ex)
{
    using namespace seqsum;
    const int lane = threadIdx.x & 63, r = lane & 15, row = lane >> 4;
    Tr x = T;
    {
        const Tr y = dpp_tr<0x111>(x);
        if (r >= 1) x = tr_compose(y, x);
    }
    {
        const Tr y = dpp_tr<0x112>(x);
        if (r >= 2) x = tr_compose(y, x);
    }
    {
        const Tr y = dpp_tr<0x114>(x);
        if (r >= 4) x = tr_compose(y, x);
    }
    {
        const Tr y = dpp_tr<0x118>(x);
        if (r >= 8) x = tr_compose(y, x);
    }
    const Tr e = dpp_tr<0x111>(x);   // the row's inclusive scan at lane - 1
    const Tr t0 = readlane_tr(x, 15), t1 = readlane_tr(x, 31), t2 = readlane_tr(x, 47);
    const Tr p2 = tr_compose(t0, t1), p3 = tr_compose(p2, t2);
    Tr pre;   // composition of the rows before this lane's row
    pre.r0 = row == 0 ? 0u : row == 1 ? t0.r0 : row == 2 ? p2.r0 : p3.r0;
    pre.r1 = row == 0 ? 0u : row == 1 ? t0.r1 : row == 2 ? p2.r1 : p3.r1;
    inc = tr_compose(pre, x);
    ex = r == 0 ? pre : tr_compose(pre, e);
}

__device__ __forceinline__ int64_t oc_chunks(int ne, int64_t i0) { return (ne - i0 + OC_CH - 1) / OC_CH; }

__global__ void k_bis_init(OcBis* __restrict__ B)
{
    B->l1 = 0.0f; B->l2 = 1e9f; B->lmid = 0.0f;
    B->steps = 0; B->done = 0; B->err = 0; B->s = 0.0f; B->ssDone = 1; B->i0 = 0; B->decided = 0;
}

// start of a sum: mode 0 also begins the bisection step (driver.cpp:170-174)
__global__ void __launch_bounds__(OC_PREFIX) k_ocs_start(OcBis* __restrict__ B, int mode, int ne,
                                                         const float* __restrict__ x, const float* __restrict__ dc,
                                                         const uint8_t* __restrict__ passive, float move)
{
    __shared__ float v[OC_PREFIX];
    float lmid = 0.0f;
    if (mode == 0) {
        if (B->done) return;
        const float l1 = B->l1, l2 = B->l2;
        if (!(double(((__fsub_rn(l2, l1)) / (__fadd_rn(l1, l2)))) > 1e-6)) {   // driver.cpp:170
            if (threadIdx.x == 0) B->done = 1;
            return;
        }
        lmid = __fmul_rn(0.5f, __fadd_rn(l1, l2));
    }
    const int t = threadIdx.x;
    if (t < ne) v[t] = oc_gen(mode, t, x, dc, passive, lmid, move);
    __syncthreads();
    if (t == 0) {
        const int64_t np = ne < OC_PREFIX ? ne : OC_PREFIX;
        float s = 0.0f;
        int64_t i = 0;
        for (; i < np; ++i) s = __fadd_rn(s, v[i]);
        for (; i < ne && !(s >= 1e-30f); ++i) s = __fadd_rn(s, oc_gen(mode, i, x, dc, passive, lmid, move));
        if (mode == 0) {
            B->lmid = lmid;
        } else {
            B->done = 0;
            B->err = 0;
        }
        B->s = s;
        B->i0 = i;
        B->ssDone = i >= ne ? 1 : 0;
        B->decided = 0;
    }
}

// csum[c] = double sum of chunk c (order irrelevant: it only bounds the float sum)
__global__ void __launch_bounds__(OC_CT) k_cs_sum(const OcBis* __restrict__ B, int mode, int ne,
                                                  const float* __restrict__ x, const float* __restrict__ dc,
                                                  const uint8_t* __restrict__ passive, float move,
                                                  double* __restrict__ csum)
{
    if (B->done || B->ssDone) return;
    const int64_t i0 = B->i0;
    const int64_t c = blockIdx.x;
    if (c >= oc_chunks(ne, i0)) return;
    const float lmid = B->lmid;
    const int64_t base = i0 + c * OC_CH + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < OC_CPER; ++k) {
        const int64_t idx = base + k * OC_CT;
        if (idx < ne) s += double(oc_gen(mode, idx, x, dc, passive, lmid, move));
    }
    for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w, 64);
    __shared__ double ws[OC_CT / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < OC_CT / 64; ++w) t += ws[w];
        csum[c] = t;
    }
}

// einfo[c] = {ulp exponent of the chunk's lowest candidate binade, candidates
// (0: none)}; thread t takes a contiguous run of chunks, one workgroup scan
// gives each run its prefix
__global__ void __launch_bounds__(OC_BT) k_cs_bounds(OcBis* __restrict__ B, int mode, int ne, float V0,
                                                     const double* __restrict__ csum, int2* __restrict__ einfo)
{
    if (B->done || B->ssDone) return;
    const int64_t i0 = B->i0;
    const int64_t nch = oc_chunks(ne, i0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t run = (nch + OC_BT - 1) / OC_BT;
    const int64_t c0 = int64_t(tid) * run, c1 = (c0 + run < nch) ? c0 + run : nch;
    double own = 0.0;
    for (int64_t c = c0; c < c1; ++c) own += csum[c];
    double inc = own;
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    double ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = 0.0;
    __shared__ double ws[OC_BT / 64];
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    const double s0 = double(B->s);
    double P = s0 + ex, tot = s0;
    for (int w = 0; w < OC_BT / 64; ++w) {
        if (w < wave) P += ws[w];
        tot += ws[w];
    }
    // (1 - u)^n >= exp(-n u (1 + 2u)) and (1 + u)^n <= exp(n u); the factor
    // 1 -/+ 1e-6 covers the double rounding of P and of exp
    const double uu = 0x1p-24 * (1.0 + 0x1p-22);
    for (int64_t c = c0; c < c1; ++c) {
        const double t = double(c * OC_CH) * uu;
        const double lo = P * exp(-t) * (1.0 - 1e-6), hi = P * exp(t) * (1.0 + 1e-6);
        int elo = 0, ehi = 0;
        frexp(lo, &elo);
        frexp(hi, &ehi);
        const int K = ehi - elo + 1;
        einfo[c] = make_int2(elo - 24, K <= OC_KMAX ? K : 0);
        P += csum[c];
    }
    // vol = s / (float)ne against V0 (driver.cpp:181-182), from the bounds of
    // the whole sum when they decide it: a float quotient above V0 (1 + 2^-23)
    // rounds above V0, one below V0 rounds to at most V0
    if (mode == 0 && tid == 0) {
        const double t = double(ne - i0) * uu;
        const double nef = double(float(ne)), v0 = double(V0);
        const double lo = tot * exp(-t) * (1.0 - 1e-6) / nef, hi = tot * exp(t) * (1.0 + 1e-6) / nef;
        if (lo > v0 * (1.0 + 0x1p-22)) {
            B->decided = 1;
            B->ssDone = 1;
        } else if (hi < v0 * (1.0 - 0x1p-22)) {
            B->decided = 2;
            B->ssDone = 1;
        }
    }
}

// agg[c OC_KMAX + j] = transducer of chunk c under the ulp exponent einfo[c].x + j
__global__ void __launch_bounds__(OC_CT) k_cs_agg(const OcBis* __restrict__ B, int mode, int ne,
                                                  const float* __restrict__ x, const float* __restrict__ dc,
                                                  const uint8_t* __restrict__ passive, float move,
                                                  const int2* __restrict__ einfo, seqsum::Tr* __restrict__ agg)
{
    using namespace seqsum;
    if (B->done || B->ssDone) return;
    const int64_t i0 = B->i0;
    const int64_t c = blockIdx.x;
    if (c >= oc_chunks(ne, i0)) return;
    const int2 ei = einfo[c];
    const int K = ei.y;
    if (K == 0) return;
    const float lmid = B->lmid;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the chunk's values, generated with coalesced loads into LDS (element j
    // at j + j/8: conflict-free reads of 8 consecutive elements per thread)
    __shared__ float sv[OC_CH + OC_CH / 8];
    const int64_t base = i0 + c * OC_CH;
#pragma unroll
    for (int k = 0; k < OC_CPER; ++k) {
        const int j = k * OC_CT + tid;
        const int64_t idx = base + j;
        sv[j + (j >> 3)] = idx < ne ? oc_gen(mode, idx, x, dc, passive, lmid, move) : 0.0f;   // a zero adds nothing
    }
    __syncthreads();
    float v[OC_CPER];
#pragma unroll
    for (int k = 0; k < OC_CPER; ++k) v[k] = sv[tid * (OC_CPER + 1) + k];
    __shared__ Tr wt[OC_KMAX][OC_CT / 64];
#pragma unroll
    for (int j = 0; j < OC_KMAX; ++j) {
        if (j < K) {   // uniform over the workgroup
            Tr F = tr_identity();
#pragma unroll
            for (int k = 0; k < OC_CPER; ++k) F = tr_compose(F, tr_elem(v[k], ei.x + j));
            Tr inc, ex;
            wave_scan_tr(F, inc, ex);
            if (lane == 63) wt[j][wave] = inc;
        }
    }
    __syncthreads();
    if (tid < K) {
        Tr acc = tr_identity();
        for (int w = 0; w < OC_CT / 64; ++w) acc = tr_compose(acc, wt[tid][w]);
        agg[c * OC_KMAX + tid] = acc;
    }
}

struct WalkShared {
    seqsum::Tr wt[OC_WT / 64 + 1];   // exclusive wave prefixes, then the total
    unsigned long long cross;
    long long m;
    int ue;
};

// scan of one transducer per thread over the walk's workgroup in thread order;
// entered and left by every thread, resets sh.cross
__device__ __forceinline__ void walk_scan(const seqsum::Tr& T, seqsum::Tr& excl, seqsum::Tr& inc, seqsum::Tr& tot,
                                          WalkShared& sh)
{
    using namespace seqsum;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Tr x, ex;
    wave_scan_tr(T, x, ex);
    if (lane == 63) sh.wt[wave] = x;
    __syncthreads();
    if (wave == 0) {   // the wave totals scanned by the first wave
        constexpr int NW = OC_WT / 64;
        const Tr w = lane < NW ? sh.wt[lane] : tr_identity();
        Tr wi, we;
        wave_scan_tr(w, wi, we);
        if (lane < NW) sh.wt[lane] = we;
        if (lane == NW - 1) sh.wt[NW] = wi;
        if (lane == 0) sh.cross = ~0ull;
    }
    __syncthreads();
    const Tr pw = sh.wt[wave];
    excl = tr_compose(pw, ex);
    inc = tr_compose(pw, x);
    tot = sh.wt[OC_WT / 64];
}

// adds the elements [from, end) of one chunk to the running sum (m, ue); thread
// t holds elements base + t OC_EPT + k in v[k]
__device__ __forceinline__ void walk_elems(int64_t base, int64_t from, int64_t end, const float (&v)[OC_EPT], int64_t& m, int& ue,
                           WalkShared& sh)
{
    using namespace seqsum;
    const int64_t own = base + int64_t(threadIdx.x) * OC_EPT;
    for (;;) {
        Tr T = tr_identity();
#pragma unroll
        for (int k = 0; k < OC_EPT; ++k)
            if (own + k >= from && own + k < end) T = tr_compose(T, tr_elem(v[k], ue));
        Tr excl, inc, tot;
        walk_scan(T, excl, inc, tot, sh);
        // the running sum only grows, so exactly one thread holds the first
        // crossing: the one whose prefix stays in the binade and whose own
        // elements leave it; it alone writes sh.cross
        int64_t mm = m + tr_apply(excl, m), mmc = 0;
        float vc = 0.0f;
        unsigned long long my = ~0ull;
        if (mm < M_TOP) {
#pragma unroll
            for (int k = 0; k < OC_EPT; ++k) {
                if (own + k >= from && own + k < end && my == ~0ull) {
                    const int64_t r = tr_apply(tr_elem(v[k], ue), mm);
                    if (mm + r >= M_TOP) {
                        my = (unsigned long long)(own + k);
                        mmc = mm;
                        vc = v[k];
                    } else {
                        mm += r;
                    }
                }
            }
        }
        if (my != ~0ull) sh.cross = my;
        __syncthreads();
        const unsigned long long cr = sh.cross;
        if (cr == ~0ull) {
            m += tr_apply(tot, m);
            return;
        }
        if (my == cr) {   // the crossing element: a real float add, then the new binade
            const float s = __fadd_rn(ldexpf(float(mmc), ue), vc);
            int64_t m2;
            int u2;
            decompose(s, m2, u2);
            sh.m = m2;
            sh.ue = u2;
        }
        __syncthreads();
        m = sh.m;
        ue = sh.ue;
        from = int64_t(cr) + 1;
    }
}

// the aggregate of a chunk for candidate j (none: capped, which reads as a crossing)
__device__ __forceinline__ seqsum::Tr pick_tr(const seqsum::Tr (&a)[OC_KMAX], int j)
{
    seqsum::Tr T{seqsum::R_CAP, seqsum::R_CAP};
#pragma unroll
    for (int jj = 0; jj < OC_KMAX; ++jj)
        if (jj == j) T = a[jj];
    return T;
}

__global__ void __launch_bounds__(OC_WT) k_cs_walk(OcBis* __restrict__ B, int mode, int ne,
                                                   const float* __restrict__ x, const float* __restrict__ dc,
                                                   const uint8_t* __restrict__ passive, float move,
                                                   const int2* __restrict__ einfo, const seqsum::Tr* __restrict__ agg)
{
    using namespace seqsum;
    if (B->done || B->ssDone) return;
    __shared__ WalkShared sh;
    const int tid = threadIdx.x;
    const int64_t i0 = B->i0;
    const int64_t nch = oc_chunks(ne, i0);
    const float lmid = B->lmid;
    int64_t m;
    int ue;
    decompose(B->s, m, ue);
    for (int64_t cw = 0; cw < nch; cw += OC_WIN) {
        const int64_t wend = (cw + OC_WIN < nch) ? cw + OC_WIN : nch;
        const int64_t cb = cw + int64_t(tid) * OC_WCPT;   // this thread's chunks: cb .. cb + OC_WCPT - 1
        int lo[OC_WCPT];
        Tr a[OC_WCPT][OC_KMAX];
#pragma unroll
        for (int q = 0; q < OC_WCPT; ++q) {
            const int64_t c = cb + q;
            int K = 0;
            lo[q] = 0;
            if (c < nch) {
                const int2 ei = einfo[c];
                lo[q] = ei.x;
                K = ei.y;
            }
#pragma unroll
            for (int j = 0; j < OC_KMAX; ++j) a[q][j] = (j < K) ? agg[c * OC_KMAX + j] : Tr{R_CAP, R_CAP};
        }
        for (int64_t cs = cw; cs < wend;) {
            Tr T = tr_identity();
#pragma unroll
            for (int q = 0; q < OC_WCPT; ++q)
                if (cb + q >= cs && cb + q < wend) T = tr_compose(T, pick_tr(a[q], ue - lo[q]));
            Tr excl, inc, tot;
            walk_scan(T, excl, inc, tot, sh);
            // the running sum only grows, so one thread holds the first chunk
            // that leaves the binade: its prefix stays inside, its inclusive
            // prefix does not; it finds the chunk and alone writes sh.cross
            int64_t mr = m + tr_apply(excl, m);
            if (mr < M_TOP && m + tr_apply(inc, m) >= M_TOP) {
                bool found = false;
#pragma unroll
                for (int q = 0; q < OC_WCPT; ++q) {
                    if (!found && cb + q >= cs && cb + q < wend) {
                        const int64_t r = tr_apply(pick_tr(a[q], ue - lo[q]), mr);
                        if (mr + r >= M_TOP) {
                            sh.cross = (unsigned long long)(cb + q);
                            sh.m = mr;
                            found = true;
                        } else {
                            mr += r;
                        }
                    }
                }
            }
            __syncthreads();
            const unsigned long long cc = sh.cross;
            if (cc == ~0ull) {   // no crossing up to the window's end
                m += tr_apply(tot, m);
                cs = wend;
                continue;
            }
            m = sh.m;
            const int64_t base = i0 + int64_t(cc) * OC_CH;
            const int64_t end = (base + OC_CH < ne) ? base + OC_CH : ne;
            float v[OC_EPT];
#pragma unroll
            for (int k = 0; k < OC_EPT; ++k) {
                const int64_t idx = base + int64_t(tid) * OC_EPT + k;
                v[k] = idx < end ? oc_gen(mode, idx, x, dc, passive, lmid, move) : 0.0f;
            }
            walk_elems(base, base, end, v, m, ue, sh);
            cs = int64_t(cc) + 1;
        }
    }
    if (tid == 0) {
        B->s = ldexpf(float(m), ue);
        B->ssDone = 1;
    }
}

// sharpness = 4 sum(x(1-x))/ne (mode 2, out[0]) and volume fraction sum(x)/ne
// (mode 1, out[1]) of xnew (driver.cpp:188-201)
__global__ void k_es_out(const OcBis* __restrict__ B, int mode, int ne, float* __restrict__ out)
{
    if (mode == 2) out[0] = ((__fmul_rn(4.0f, B->s)) / (float(ne)));
    else out[1] = ((B->s) / (float(ne)));
}

__global__ void k_bis_end(OcBis* __restrict__ B, int ne, float V0)
{
    if (B->done) return;
    bool above;
    if (B->decided) {
        above = B->decided == 1;
    } else {
        if (!B->ssDone) {   // the walk always completes its sum
            B->err = 1;
            B->done = 1;
            return;
        }
        const float vol = ((B->s) / (float(ne)));
        above = __fsub_rn(vol, V0) > 0;
    }
    if (above) B->l1 = B->lmid;
    else B->l2 = B->lmid;
    ++B->steps;
}

// out[0] = last lmid, out[1] = bisection steps
__global__ void k_bis_out(const OcBis* __restrict__ B, float* __restrict__ out)
{
    out[0] = B->lmid;
    out[1] = float(B->steps);
}

// xnew for the final multiplier and change = max|xnew - x| (driver.cpp:184);
// the max of non-negative floats is the max of their bit patterns
__global__ void __launch_bounds__(256) k_oc_apply(int ne, const float* __restrict__ x, const float* __restrict__ dc,
                                                  const uint8_t* __restrict__ passive, const float* __restrict__ lm,
                                                  float move, float* __restrict__ xnew,
                                                  unsigned* __restrict__ changeBits)
{
    const float lmid = lm[0];
    unsigned b = 0u;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < ne; e += int64_t(gridDim.x) * 256) {
        const float v = (passive && passive[e]) ? 1.0f : oc_xnew(x[e], dc[e], lmid, move);
        xnew[e] = v;
        b = max(b, __float_as_uint(fabsf(__fsub_rn(v, x[e]))));
    }
    for (int w = 32; w >= 1; w >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, w, 64));
    __shared__ unsigned wb[4];
    if ((threadIdx.x & 63) == 0) wb[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(changeBits, max(max(wb[0], wb[1]), max(wb[2], wb[3])));
}

// generic sequential float sum of an array (for tests)
__global__ void __launch_bounds__(SS_THREADS) k_seq_sum(int64_t n, const float* __restrict__ v, float* __restrict__ out)
{
    __shared__ SeqShared sh;
    GenX g{v};
    const float s = block_seq_sum(n, g, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// ---------------------------------------------------------------------------
// compliance (fea.cpp:587-677): ce[e] = sum_i ue_i * (sum_j K_ij ue_j), 0 when
// E <= 1e-16; canonical tile sum of E*ce in compact element order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_compliance(Grid g, int ne, const int32_t* __restrict__ c2cell,
                                                    const double* __restrict__ Ke, const double* __restrict__ E,
                                                    const double* __restrict__ U, double* __restrict__ ce,
                                                    double* __restrict__ partial)
{
    __shared__ double sh[CANON_LANES];
    __shared__ double sK[576];
    for (int i = threadIdx.x; i < 576; i += blockDim.x) sK[i] = Ke[i];
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t e = base + int64_t(k) * CANON_LANES;
        if (e >= ne) continue;
        const double Ee = E[e];
        double item = 0.0;
        if (Ee <= 1.0e-16) {
            ce[e] = 0.0;
        } else {
            const int cell = c2cell[e];
            const int cx = cell % g.resX, ty = (cell / g.resX) % g.resY, cz = cell / (g.resX * g.resY);
            double ue[24];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int b = a & 3;
                const int ox = (b == 1 || b == 2) ? 1 : 0, oy = (b <= 1) ? 1 : 0, oz = a >> 2;
                const int64_t nn = (int64_t(cz + oz) * g.nny + (ty + oy)) * g.nnx + (cx + ox);
                ue[3 * a] = U[3 * nn];
                ue[3 * a + 1] = U[3 * nn + 1];
                ue[3 * a + 2] = U[3 * nn + 2];
            }
            // the row loop stays rolled (unrolled, the 576 products need more
            // registers than a lane has); ue_i is read again from U (cached)
            double c = 0.0;
#pragma unroll 1
            for (int i = 0; i < 24; ++i) {
                double rd = 0.0;
#pragma unroll
                for (int j = 0; j < 24; ++j) rd = rd + sK[i * 24 + j] * ue[j];
                const int a = i / 3, b = a & 3;
                const int ox = (b == 1 || b == 2) ? 1 : 0, oy = (b <= 1) ? 1 : 0, oz = a >> 2;
                const int64_t nn = (int64_t(cz + oz) * g.nny + (ty + oy)) * g.nnx + (cx + ox);
                c = c + U[3 * nn + (i - 3 * a)] * rd;
            }
            ce[e] = c;
            item = Ee * c;
        }
        s = s + item;
    }
    const int l = threadIdx.x;
    sh[l] = s;
    __syncthreads();
    for (int w = CANON_LANES / 2; w >= 64; w >>= 1) {
        if (l < w) sh[l] = sh[l] + sh[l + w];
        __syncthreads();
    }
    if (l < 64) {
        double v = sh[l];
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1) v = v + __shfl_down(v, w, 64);
        if (l == 0) partial[blockIdx.x] = v;
    }
}

// canonical tile partials of E[e] * (ce[e] / Csolid) (driver.cpp:144-150)
__global__ void __launch_bounds__(256) k_cobj_tiles(int ne, const double* __restrict__ E,
                                                    const double* __restrict__ ce, const double* __restrict__ Csolid,
                                                    double* __restrict__ partial)
{
    __shared__ double sh[CANON_LANES];
    const double Cs = *Csolid;
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t e = base + int64_t(k) * CANON_LANES;
        if (e < ne) s = s + ((Cs > 0) ? E[e] * (ce[e] / Cs) : E[e] * ce[e]);
    }
    const int l = threadIdx.x;
    sh[l] = s;
    __syncthreads();
    for (int w = CANON_LANES / 2; w >= 64; w >>= 1) {
        if (l < w) sh[l] = sh[l] + sh[l + w];
        __syncthreads();
    }
    if (l < 64) {
        double v = sh[l];
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1) v = v + __shfl_down(v, w, 64);
        if (l == 0) partial[blockIdx.x] = v;
    }
}

// o_x_check's position-keyed checksum of a float design (the oracle's
// definition): sum of splitmix64((i << 32) | bits(x[i])) modulo 2^64.  Integer
// addition is associative, so the block partials and the atomic combine give
// the oracle's value in any order.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void k_zero_u64(unsigned long long* __restrict__ out) { *out = 0ull; }
__global__ void __launch_bounds__(256) k_x_check(int64_t n, const float* __restrict__ x,
                                                 unsigned long long* __restrict__ out)
{
    unsigned long long h = 0ull;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        h += splitmix64((static_cast<unsigned long long>(i) << 32) | __float_as_uint(x[i]));
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) h += __shfl_down(h, w, 64);
    __shared__ unsigned long long sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, sh[0] + sh[1] + sh[2] + sh[3]);
}

int x_check_dev(t3d_problem* p, int64_t n, const float* x, unsigned long long* dst)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, p->stream, dst);
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_x_check, dim3(unsigned(std::max<int64_t>(blocks, 1))), dim3(256), 0, p->stream, n, x, dst);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

int compliance_dev(t3d_problem* p, const double* E, const double* U, double* ce, double* Cdst)
{
    int rc = ensure_red(p, p->ne);
    if (rc) return rc;
    const int64_t tiles = (p->ne + CANON_TILE - 1) / CANON_TILE;
    hipLaunchKernelGGL(k_compliance, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, p->stream, p->g, p->ne,
                       p->compactToCell, p->Ke, E, U, ce, p->red);
    T3D_HIP(hipGetLastError());
    return canon_finish(p, tiles, Cdst);
}

int cobj_dev(t3d_problem* p, const double* E, const double* ce, const double* Csolid, double* dst)
{
    int rc = ensure_red(p, p->ne);
    if (rc) return rc;
    const int64_t tiles = (p->ne + CANON_TILE - 1) / CANON_TILE;
    hipLaunchKernelGGL(k_cobj_tiles, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, p->stream, p->ne, E, ce, Csolid,
                       p->red);
    T3D_HIP(hipGetLastError());
    return canon_finish(p, tiles, dst);
}

int launch_simp(t3d_problem* p, const float* x, float E0, float Emin, double* E)
{
    hipLaunchKernelGGL(k_simp, dim3(unsigned((p->ne + 255) / 256)), dim3(256), 0, p->stream, p->ne, x, E0, Emin, E);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}
int launch_fill_f64(t3d_problem* p, int64_t n, double v, double* a)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(unsigned((n + 255) / 256)), dim3(256), 0, p->stream, n, v, a);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}
int launch_fill_f32(t3d_problem* p, int64_t n, float v, float* a)
{
    hipLaunchKernelGGL(k_fill_f32, dim3(unsigned((n + 255) / 256)), dim3(256), 0, p->stream, n, v, a);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}
int launch_sens(t3d_problem* p, const float* x, const double* ce, const double* Csolid, float E0, float Emin,
                float penal, float* xdc)
{
    hipLaunchKernelGGL(k_sens_xdc, dim3(unsigned((p->ne + 255) / 256)), dim3(256), 0, p->stream, p->ne, x, ce, Csolid,
                       E0, Emin, penal, xdc);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}
int launch_dc_post(t3d_problem* p, const float* filt, const float* x, float* dc)
{
    hipLaunchKernelGGL(k_dc_post, dim3(unsigned((p->ne + 255) / 256)), dim3(256), 0, p->stream, p->ne, filt, x, dc);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}
// scal: [0] lmid [1] steps [2] changeBits [3] sharpness [4] volFrac (float slots)
int launch_oc(t3d_problem* p, const float* x, const float* dc, float V0, float move, float* xnew, float* scal,
              const uint8_t* passive)
{
    hipStream_t st = p->stream;
    const int ne = p->ne;
    const int64_t nchMax = (int64_t(ne) + OC_CH - 1) / OC_CH + 1;
    if (!p->ocBis) T3D_HIP(hipMalloc(&p->ocBis, sizeof(OcBis)));
    // per chunk: OC_KMAX aggregates, the double sum, the candidate binades
    const size_t perChunk = OC_KMAX * sizeof(seqsum::Tr) + sizeof(double) + sizeof(int2);
    if (p->ocAggCap < nchMax) {
        if (p->ocAgg) T3D_HIP(hipFree(p->ocAgg));
        p->ocAgg = nullptr;
        T3D_HIP(hipMalloc(&p->ocAgg, perChunk * size_t(nchMax)));
        p->ocAggCap = nchMax;
    }
    OcBis* B = static_cast<OcBis*>(p->ocBis);
    seqsum::Tr* agg = static_cast<seqsum::Tr*>(p->ocAgg);
    double* csum = reinterpret_cast<double*>(agg + nchMax * OC_KMAX);
    int2* einfo = reinterpret_cast<int2*>(csum + nchMax);
    const dim3 grid = dim3(unsigned(nchMax));
    auto exact_sum = [&](int mode, const float* xs, const float* dcs, const uint8_t* pas, float mv) {
        hipLaunchKernelGGL(k_ocs_start, dim3(1), dim3(OC_PREFIX), 0, st, B, mode, ne, xs, dcs, pas, mv);
        hipLaunchKernelGGL(k_cs_sum, grid, dim3(OC_CT), 0, st, (const OcBis*)B, mode, ne, xs, dcs, pas, mv, csum);
        hipLaunchKernelGGL(k_cs_bounds, dim3(1), dim3(OC_BT), 0, st, B, mode, ne, V0, (const double*)csum, einfo);
        hipLaunchKernelGGL(k_cs_agg, grid, dim3(OC_CT), 0, st, (const OcBis*)B, mode, ne, xs, dcs, pas, mv,
                           (const int2*)einfo, agg);
        hipLaunchKernelGGL(k_cs_walk, dim3(1), dim3(OC_WT), 0, st, B, mode, ne, xs, dcs, pas, mv,
                           (const int2*)einfo, (const seqsum::Tr*)agg);
    };
    hipLaunchKernelGGL(k_bis_init, dim3(1), dim3(1), 0, st, B);
    // bisection steps in groups of OC_STEPS, the end flag read back between
    // groups; the float bisection from [0, 1e9] ends within ~180 steps even
    // when l1 stays 0
    int hostDone[2] = {0, 0};   // done, err
    for (int step = 0; step < 256 && !hostDone[0]; step += OC_STEPS) {
        for (int k = 0; k < OC_STEPS; ++k) {
            exact_sum(0, x, dc, passive, move);
            hipLaunchKernelGGL(k_bis_end, dim3(1), dim3(1), 0, st, B, ne, V0);
        }
        T3D_HIP(hipGetLastError());
        T3D_HIP(hipMemcpyAsync(hostDone, &B->done, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
    }
    T3D_CHECK(hostDone[0] && !hostDone[1], T3D_ERR_STATE, "OC bisection: exact volume sum did not complete");
    hipLaunchKernelGGL(k_bis_out, dim3(1), dim3(1), 0, st, (const OcBis*)B, scal);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipMemsetAsync(scal + 2, 0, sizeof(float), st));
    const unsigned applyGrid = unsigned(std::min<int64_t>((int64_t(ne) + 255) / 256, OC_APPLY_GRID));
    hipLaunchKernelGGL(k_oc_apply, dim3(applyGrid), dim3(256), 0, st, p->ne, x, dc, passive, scal, move, xnew,
                       reinterpret_cast<unsigned*>(scal + 2));
    T3D_HIP(hipGetLastError());
    // sharpness and volume fraction of xnew, exact, spread over the GPU like the bisection's sums
    for (int mode = 2; mode >= 1; --mode) {
        exact_sum(mode, xnew, nullptr, nullptr, 0.0f);
        hipLaunchKernelGGL(k_es_out, dim3(1), dim3(1), 0, st, (const OcBis*)B, mode, ne, scal + 3);
        T3D_HIP(hipGetLastError());
    }
    return T3D_OK;
}

}  // namespace t3d

// ---------------------------------------------------------------------------
// Compliance without allocation: the canonical tree's tile partials and the
// result live in one file-scope device array instead of the problem's lazily
// allocated reduction buffer.  Each device holds its own copy of the array
// (its address is looked up once per device with hipGetSymbolAddress and
// kept); every call on a device shares it, so calls that overlap on one
// device race.  Every slot a call reads is written by that call first, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t NOALLOC_TILES = (int64_t(T3D_NOALLOC_MAX_ELEMENTS) + t3d::CANON_TILE - 1) / t3d::CANON_TILE;
constexpr int64_t NOALLOC_RESULT = 2 * NOALLOC_TILES;       // slot of the total
constexpr int64_t NOALLOC_DOUBLES = NOALLOC_RESULT + 32;    // 1 MiB + 256 B
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace t3d {
// device address of `scratch` on device dev (looked up with dev current, once)
static int scratch_ptr(int dev, double** out)
{
    T3D_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, T3D_ERR_ARG, "scratch: device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        int cur = 0;
        T3D_HIP(hipGetDevice(&cur));
        if (cur != dev) T3D_HIP(hipSetDevice(dev));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) T3D_HIP(hipSetDevice(cur));
        T3D_CHECK(e == hipSuccess && ptr, T3D_ERR_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_dev[dev];
    return T3D_OK;
}
}  // namespace t3d

using namespace t3d;

extern "C" {

int t3d_compliance(t3d_problem* p, const double* E, const double* U, double* ce, double* C_host)
{
    T3D_CHECK(p && E && U && ce, T3D_ERR_ARG, "t3d_compliance: null argument");
    int rc = compliance_dev(p, E, U, ce, p->scal);
    if (rc) return rc;
    if (C_host) T3D_HIP(hipMemcpyAsync(C_host, p->scal, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_compliance_noalloc(t3d_problem* p, const double* E, const double* U, double* ce, double* C_host)
{
    T3D_CHECK(p && E && U && ce, T3D_ERR_ARG, "t3d_compliance_noalloc: null argument");
    T3D_CHECK(p->ne <= T3D_NOALLOC_MAX_ELEMENTS, T3D_ERR_ARG,
              "t3d_compliance_noalloc: more elements than T3D_NOALLOC_MAX_ELEMENTS");
    double* buf = nullptr;
    int rc = scratch_ptr(p->device, &buf);
    if (rc) return rc;
    const int64_t tiles = (p->ne + CANON_TILE - 1) / CANON_TILE;
    hipLaunchKernelGGL(k_compliance, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, p->stream, p->g, p->ne,
                       p->compactToCell, p->Ke, E, U, ce, buf);
    T3D_HIP(hipGetLastError());
    if ((rc = canon_finish_buf(p, buf, tiles, buf + NOALLOC_RESULT))) return rc;
    if (C_host) T3D_HIP(hipMemcpyAsync(C_host, buf + NOALLOC_RESULT, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_seq_sum_f32(t3d_problem* p, const float* v, int64_t n, float* out_host)
{
    T3D_CHECK(p && v && out_host && n >= 0, T3D_ERR_ARG, "t3d_seq_sum_f32: bad argument");
    float* d = reinterpret_cast<float*>(p->scal + 8);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(SS_THREADS), 0, p->stream, n, v, d);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipMemcpyAsync(out_host, d, sizeof(float), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_simp(t3d_problem* p, const float* x, float E0, float Emin, double* E)
{
    T3D_CHECK(p && x && E, T3D_ERR_ARG, "t3d_simp: null argument");
    int rc = launch_simp(p, x, E0, Emin, E);
    if (rc) return rc;
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

// OC update: x, dc (post-filter sensitivities) -> xnew; out[0] change, out[1]
// sharpness, out[2] volume fraction, out[3] bisection steps (host)
int t3d_oc_update(t3d_problem* p, const float* x, const float* dc, float V0, float move, float* xnew, float* out)
{
    return t3d_oc_update_passive(p, x, dc, nullptr, V0, move, xnew, out);
}

int t3d_oc_update_passive(t3d_problem* p, const float* x, const float* dc, const uint8_t* passive, float V0,
                          float move, float* xnew, float* out)
{
    T3D_CHECK(p && x && dc && xnew, T3D_ERR_ARG, "t3d_oc_update: null argument");
    float* sc = reinterpret_cast<float*>(p->scal + 16);
    int rc = launch_oc(p, x, dc, V0, move, xnew, sc, passive);
    if (rc) return rc;
    float h[5];
    T3D_HIP(hipMemcpyAsync(h, sc, sizeof(float) * 5, hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    if (out) {
        out[0] = h[2];   // bit pattern of a non-negative float
        out[1] = h[3];
        out[2] = h[4];
        out[3] = h[1];
    }
    return T3D_OK;
}

}  // extern "C"

namespace t3d {
// device floating-point semantics probe: op 0 sqrtf(a), 1 a/b (float), 2 sqrt(a) (double)
__global__ void k_fp_probe(int op, int64_t n, const void* __restrict__ a, const void* __restrict__ b,
                           void* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) ((float*)out)[i] = sqrtf(((const float*)a)[i]);
    else if (op == 1) ((float*)out)[i] = ((const float*)a)[i] / ((const float*)b)[i];
    else ((double*)out)[i] = sqrt(((const double*)a)[i]);
}
}  // namespace t3d

extern "C" int t3d_fp_probe(int op, int64_t n, const void* a, const void* b, void* out)
{
    T3D_CHECK(op >= 0 && op <= 2 && n >= 0 && a && out && (op != 1 || b), T3D_ERR_ARG, "t3d_fp_probe: bad argument");
    hipLaunchKernelGGL(t3d::k_fp_probe, dim3(unsigned((n + 255) / 256)), dim3(256), 0, 0, op, n, a, b, out);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}
