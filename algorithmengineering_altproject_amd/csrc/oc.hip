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
    o.r0 = __shfl_up((long long)t.r0, d, 64);
    o.r1 = __shfl_up((long long)t.r1, d, 64);
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

// passive (nullable): elements held at density 1 (matlab/TOP3D_XL.m:154)
struct GenXnew {
    const float* x;
    const float* dc;
    const uint8_t* passive;
    float lmid, move;
    __device__ float operator()(int64_t i) const
    {
        return (passive && passive[i]) ? 1.0f : oc_xnew(x[i], dc[i], lmid, move);
    }
};
struct GenX {
    const float* x;
    __device__ float operator()(int64_t i) const { return x[i]; }
};

// ---------------------------------------------------------------------------
// The OC bisection with its exact sums spread over the GPU.  Per bisection
// step: k_bis_start takes lmid and sums the first elements sequentially; then
// rounds of k_ss_agg (every workgroup forms the transducer aggregate of one
// chunk of OC densities in a window after the sum's position, under the
// current binade) and k_ss_carry (one wavefront composes the window's
// aggregates in chunk order, finds the first binade crossing, locates its
// element and restarts after it in the new binade; without a crossing it moves
// past the window and doubles it) until the sum is complete; k_bis_end takes
// the bisection decision.  Every kernel returns at once when its flags say
// there is nothing to do.  OC densities are >= 1e-3 (or 1 for passive
// elements), so after the first 1024 elements the sum is >= 1 and crosses at
// most log2(n) + 1 binades; with at most log2(n / 2^15) window doublings,
// OC_ROUNDS rounds always complete it (k_bis_end flags an error otherwise).
// ---------------------------------------------------------------------------
constexpr int OC_CT = 256;                 // threads of an aggregate workgroup
constexpr int OC_CPER = 8;                 // elements per thread
constexpr int OC_CH = OC_CT * OC_CPER;     // elements per chunk
constexpr int OC_ROUNDS = 64;
constexpr int OC_GROUP = 8;               // rounds launched between host checks
constexpr int OC_PREFIX = 64;             // elements summed one by one before the rounds (any count is exact)
constexpr int OC_W0 = 16;                 // first window, chunks
constexpr int OC_AGG_GRID = 1024;         // workgroups of k_ss_agg (grid-stride over the window)

struct OcBis {
    float l1, l2, lmid;
    int steps, done, err;
    float s;          // running float sum
    int ssDone;
    long long i0;     // next element of the sum
    long long W;      // window of the next round, chunks
};

// the summed elements: 0 = OC densities at B->lmid, 1 = x, 2 = x (1 - x)
__device__ __forceinline__ float oc_gen(int mode, int64_t i, const float* __restrict__ x, const float* __restrict__ dc,
                                        const uint8_t* __restrict__ passive, float lmid, float move)
{
    if (mode == 1) return x[i];
    if (mode == 2) return __fmul_rn(x[i], __fsub_rn(1.0f, x[i]));
    return (passive && passive[i]) ? 1.0f : oc_xnew(x[i], dc[i], lmid, move);
}

__device__ __forceinline__ seqsum::Tr shfl_down_tr(const seqsum::Tr& t, int d)
{
    seqsum::Tr o;
    o.r0 = __shfl_down((long long)t.r0, d, 64);
    o.r1 = __shfl_down((long long)t.r1, d, 64);
    return o;
}

__global__ void k_bis_init(OcBis* __restrict__ B)
{
    B->l1 = 0.0f; B->l2 = 1e9f; B->lmid = 0.0f;
    B->steps = 0; B->done = 0; B->err = 0; B->s = 0.0f; B->ssDone = 1; B->i0 = 0; B->W = OC_W0;
}

__global__ void k_bis_start(OcBis* __restrict__ B, int ne, const float* __restrict__ x, const float* __restrict__ dc,
                            const uint8_t* __restrict__ passive, float move)
{
    if (B->done) return;
    const float l1 = B->l1, l2 = B->l2;
    if (!(double(((__fsub_rn(l2, l1)) / (__fadd_rn(l1, l2)))) > 1e-6)) {   // driver.cpp:170
        B->done = 1;
        return;
    }
    const float lmid = __fmul_rn(0.5f, __fadd_rn(l1, l2));
    B->lmid = lmid;
    GenXnew gx{x, dc, passive, lmid, move};
    float s = 0.0f;
    int64_t i = 0;
    for (; i < ne && (i < OC_PREFIX || !(s >= 1e-30f)); ++i) s = __fadd_rn(s, gx(i));
    B->s = s;
    B->i0 = i;
    B->W = OC_W0;
    B->ssDone = i >= ne ? 1 : 0;
}

// aggT[c] = transducer of chunk c = [i0 + c OC_CH, ...) under the binade of s
__global__ void __launch_bounds__(OC_CT) k_ss_agg(const OcBis* __restrict__ B, int mode, int ne,
                                                  const float* __restrict__ x, const float* __restrict__ dc,
                                                  const uint8_t* __restrict__ passive, float move,
                                                  seqsum::Tr* __restrict__ aggT)
{
    using namespace seqsum;
    if (B->done || B->ssDone) return;
    const int64_t i0 = B->i0;
    const int64_t nch = (ne - i0 + OC_CH - 1) / OC_CH;
    const int64_t nw = nch < B->W ? nch : B->W;
    int64_t m;
    int ue;
    decompose(B->s, m, ue);
    const float lmid = B->lmid;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ Tr wt[OC_CT / 64];
    for (int64_t c = blockIdx.x; c < nw; c += gridDim.x) {   // uniform over the workgroup
        const int64_t c0 = i0 + c * OC_CH;
        Tr F = tr_identity();
#pragma unroll
        for (int k = 0; k < OC_CPER; ++k) {
            const int64_t idx = c0 + int64_t(tid) * OC_CPER + k;
            if (idx < ne) F = tr_compose(F, tr_elem(oc_gen(mode, idx, x, dc, passive, lmid, move), ue));
        }
        // ordered wave reduction: lane i takes lane i + d's aggregate after its own
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Tr o = shfl_down_tr(F, d);
            if ((lane & (2 * d - 1)) == 0 && lane + d < 64) F = tr_compose(F, o);
        }
        if (lane == 0) wt[wave] = F;
        __syncthreads();
        if (tid == 0) {
            Tr acc = tr_identity();
            for (int w = 0; w < OC_CT / 64; ++w) acc = tr_compose(acc, wt[w]);
            aggT[c] = acc;
        }
        __syncthreads();
    }
}

// one wavefront: compose the chunk aggregates in order from i0; at the first
// binade crossing, find its element, add it with a real float add and restart
// after it; otherwise the sum is complete
__global__ void __launch_bounds__(64) k_ss_carry(OcBis* __restrict__ B, int mode, int ne, const float* __restrict__ x,
                                                 const float* __restrict__ dc, const uint8_t* __restrict__ passive,
                                                 float move, const seqsum::Tr* __restrict__ aggT)
{
    using namespace seqsum;
    if (B->done || B->ssDone) return;
    const int lane = threadIdx.x;
    const int64_t i0 = B->i0;
    const float s = B->s;
    int64_t m;
    int ue;
    decompose(s, m, ue);
    const int64_t nch = (ne - i0 + OC_CH - 1) / OC_CH;
    const int64_t nw = nch < B->W ? nch : B->W;
    int64_t cross = -1;   // chunk holding the crossing
    for (int64_t b = 0; b < nw && cross < 0; b += 64) {
        const int64_t c = b + lane;
        Tr inc = (c < nw) ? aggT[c] : tr_identity();
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Tr o = shfl_up_tr(inc, d);
            if (lane >= d) inc = tr_compose(o, inc);
        }
        const bool crosses = c < nw && m + tr_apply(inc, m) >= M_TOP;
        const unsigned long long ball = __ballot(crosses);
        if (ball) {
            const int first = __ffsll((long long)ball) - 1;
            Tr ex = shfl_up_tr(inc, 1);
            if (lane == 0) ex = tr_identity();
            const int64_t r = tr_apply(ex, m);
            m += __shfl(r, first, 64);          // m before the crossing chunk
            cross = b + first;
        } else {
            const Tr tot{__shfl(inc.r0, 63, 64), __shfl(inc.r1, 63, 64)};
            m += tr_apply(tot, m);
        }
    }
    if (cross < 0) {   // no crossing in the window: move past it (same binade), double it
        if (lane == 0) {
            B->s = ldexpf(float(m), ue);
            const int64_t next = i0 + nw * OC_CH;
            B->i0 = next < ne ? next : ne;
            B->ssDone = next >= ne ? 1 : 0;
            B->W = 2 * B->W;
        }
        return;
    }
    // the crossing chunk: 64 lanes x 32 elements, a wave scan, then one lane walks its run
    const float lmid = B->lmid;
    const int64_t cs = i0 + cross * OC_CH;
    constexpr int RUN = OC_CH / 64;
    Tr T = tr_identity();
    for (int k = 0; k < RUN; ++k) {
        const int64_t idx = cs + int64_t(lane) * RUN + k;
        if (idx < ne) T = tr_compose(T, tr_elem(oc_gen(mode, idx, x, dc, passive, lmid, move), ue));
    }
    Tr inc = T;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Tr o = shfl_up_tr(inc, d);
        if (lane >= d) inc = tr_compose(o, inc);
    }
    const bool crosses = m + tr_apply(inc, m) >= M_TOP;
    const unsigned long long ball = __ballot(crosses);
    const int first = __ffsll((long long)ball) - 1;   // the chunk crosses, so some lane does
    Tr ex = shfl_up_tr(inc, 1);
    if (lane == 0) ex = tr_identity();
    if (lane == first) {
        int64_t mm = m + tr_apply(ex, m);
        for (int k = 0; k < RUN; ++k) {
            const int64_t idx = cs + int64_t(lane) * RUN + k;
            const float v = oc_gen(mode, idx, x, dc, passive, lmid, move);
            const int64_t r = tr_apply(tr_elem(v, ue), mm);
            if (mm + r >= M_TOP) {
                B->s = __fadd_rn(ldexpf(float(mm), ue), v);
                B->i0 = idx + 1;
                B->ssDone = (idx + 1 >= ne) ? 1 : 0;
                break;
            }
            mm += r;
        }
    }
}

// start of an exact sum of x (mode 1) or x (1 - x) (mode 2): the elements are
// added one by one until the sum is a normal float, then the rounds take over
__global__ void k_es_start(OcBis* __restrict__ B, int mode, int ne, const float* __restrict__ x)
{
    float s = 0.0f;
    int64_t i = 0;
    for (; i < ne && (i < OC_PREFIX || !(s >= 1e-30f)); ++i) s = __fadd_rn(s, oc_gen(mode, i, x, nullptr, nullptr, 0.0f, 0.0f));
    B->done = 0;
    B->err = 0;
    B->s = s;
    B->i0 = i;
    B->W = OC_W0;
    B->ssDone = i >= ne ? 1 : 0;
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
    if (!B->ssDone) {   // more binade crossings than the rounds allow: cannot happen for OC densities
        B->err = 1;
        B->done = 1;
        return;
    }
    const float vol = ((B->s) / (float(ne)));
    if (__fsub_rn(vol, V0) > 0) B->l1 = B->lmid;
    else B->l2 = B->lmid;
    ++B->steps;
}

// out[0] = last lmid, out[1] = bisection steps
__global__ void k_bis_out(const OcBis* __restrict__ B, float* __restrict__ out)
{
    out[0] = B->lmid;
    out[1] = float(B->steps);
}

// xnew for the final multiplier and change = max|xnew - x| (driver.cpp:184)
__global__ void k_oc_apply(int ne, const float* __restrict__ x, const float* __restrict__ dc,
                           const uint8_t* __restrict__ passive, const float* __restrict__ lm, float move,
                           float* __restrict__ xnew, unsigned* __restrict__ changeBits)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    float d = 0.0f;
    if (e < ne) {
        const float v = (passive && passive[e]) ? 1.0f : oc_xnew(x[e], dc[e], lm[0], move);
        xnew[e] = v;
        d = fabsf(__fsub_rn(v, x[e]));
    }
    // max of non-negative floats = max of their bit patterns
    unsigned b = __float_as_uint(d);
    for (int w = 32; w >= 1; w >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, w, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(changeBits, b);
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
            double c = 0.0;
            for (int i = 0; i < 24; ++i) {
                double rd = 0.0;
#pragma unroll
                for (int j = 0; j < 24; ++j) rd = rd + sK[i * 24 + j] * ue[j];
                c = c + ue[i] * rd;
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
    if (p->ocAggCap < nchMax) {
        if (p->ocAgg) T3D_HIP(hipFree(p->ocAgg));
        p->ocAgg = nullptr;
        T3D_HIP(hipMalloc(&p->ocAgg, sizeof(seqsum::Tr) * size_t(nchMax)));
        p->ocAggCap = nchMax;
    }
    OcBis* B = static_cast<OcBis*>(p->ocBis);
    seqsum::Tr* agg = static_cast<seqsum::Tr*>(p->ocAgg);
    const unsigned aggGrid = unsigned(std::min<int64_t>(nchMax, OC_AGG_GRID));
    hipLaunchKernelGGL(k_bis_init, dim3(1), dim3(1), 0, st, B);
    // bisection steps in blocks of 8, checked on the host between blocks; the
    // float bisection from [0, 1e9] ends within ~180 steps even when l1 stays 0
    // one step at a time: rounds in groups of OC_GROUP, the sum's completion
    // (ssDone) read back between groups; a finished bisection ends the loop
    int hostDone[2] = {0, 0};
    int hostSS[4] = {0, 0, 0, 0};   // done, err, s, ssDone
    for (int step = 0; step < 256 && !hostDone[0]; ++step) {
        hipLaunchKernelGGL(k_bis_start, dim3(1), dim3(1), 0, st, B, ne, x, dc, passive, move);
        for (int r = 0; r < OC_ROUNDS; r += OC_GROUP) {
            for (int k = 0; k < OC_GROUP; ++k) {
                hipLaunchKernelGGL(k_ss_agg, dim3(aggGrid), dim3(OC_CT), 0, st, (const OcBis*)B, 0, ne, x, dc,
                                   passive, move, agg);
                hipLaunchKernelGGL(k_ss_carry, dim3(1), dim3(64), 0, st, B, 0, ne, x, dc, passive, move,
                                   (const seqsum::Tr*)agg);
            }
            T3D_HIP(hipGetLastError());
            T3D_HIP(hipMemcpyAsync(hostSS, &B->done, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
            T3D_HIP(hipStreamSynchronize(st));
            if (hostSS[0] || hostSS[3]) break;   // bisection finished, or this step's sum complete
        }
        hipLaunchKernelGGL(k_bis_end, dim3(1), dim3(1), 0, st, B, ne, V0);
        T3D_HIP(hipGetLastError());
        T3D_HIP(hipMemcpyAsync(hostDone, &B->done, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
    }
    T3D_CHECK(hostDone[0] && !hostDone[1], T3D_ERR_STATE, "OC bisection: exact volume sum did not complete");
    hipLaunchKernelGGL(k_bis_out, dim3(1), dim3(1), 0, st, (const OcBis*)B, scal);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipMemsetAsync(scal + 2, 0, sizeof(float), st));
    hipLaunchKernelGGL(k_oc_apply, dim3(unsigned((p->ne + 255) / 256)), dim3(256), 0, st, p->ne, x, dc, passive,
                       scal, move, xnew, reinterpret_cast<unsigned*>(scal + 2));
    T3D_HIP(hipGetLastError());
    // sharpness and volume fraction of xnew, exact, spread over the GPU like the bisection's sums
    for (int mode = 2; mode >= 1; --mode) {
        hipLaunchKernelGGL(k_es_start, dim3(1), dim3(1), 0, st, B, mode, ne, (const float*)xnew);
        int hs[4] = {0, 0, 0, 0};
        for (int r = 0; r < 32 * OC_ROUNDS && !hs[3]; r += OC_GROUP) {
            for (int k = 0; k < OC_GROUP; ++k) {
                hipLaunchKernelGGL(k_ss_agg, dim3(aggGrid), dim3(OC_CT), 0, st, (const OcBis*)B, mode, ne,
                                   (const float*)xnew, (const float*)nullptr, (const uint8_t*)nullptr, 0.0f, agg);
                hipLaunchKernelGGL(k_ss_carry, dim3(1), dim3(64), 0, st, B, mode, ne, (const float*)xnew,
                                   (const float*)nullptr, (const uint8_t*)nullptr, 0.0f, (const seqsum::Tr*)agg);
            }
            T3D_HIP(hipGetLastError());
            T3D_HIP(hipMemcpyAsync(hs, &B->done, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
            T3D_HIP(hipStreamSynchronize(st));
        }
        T3D_CHECK(hs[3], T3D_ERR_STATE, "OC statistics: exact sum did not complete");
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
