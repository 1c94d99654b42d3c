// chol.hpp -- the coarsest-level Cholesky solve L L^T x = b of the V-cycle in
// one workgroup of CHOL_MAX threads, thread i owning row i
// (coarsest.cpp:92-122).  Shared by the FP64 multigrid (mg.hip), the mixed
// solver (mixed.hip) and the batched mixed solver (batch.hip).
//
// Every row subtracts its terms in the reference's order, so the result is
// bit-identical to it:
//   forward  x_i = (b_i - L_i0 x_0 - L_i1 x_1 - ... - L_i,i-1 x_i-1) / L_ii
//            (the row-form sum of coarsest.cpp:101-108, increasing k);
//   backward the saxpy form of coarsest.cpp:113-121: x_k = y_k / L_kk, then
//            y_i -= L_ki x_k for i < k, k running from N-1 down to 0.
//
// Blocked by wavefronts: the rows are cut into blocks of 64, one per wave.
// For block t the owning wave solves its 64x64 diagonal block on its own,
// passing each new x_k to the other lanes with v_readlane (no workgroup
// barrier inside the block); one barrier then publishes the 64 values.  The
// next block's rows apply the block's columns at once, from a 64x64 tile in
// LDS; all other rows apply them one block later, from global memory, while
// the next diagonal block is being solved.  The waves that are not solving
// also stage the next diagonal tile and the next off-diagonal tile into LDS.
// Every row still applies the columns in the reference's order.  One barrier
// per 64 rows and direction instead of two per row.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace t3d {

constexpr int CHOL_MAX = 1024;   // rows; the workgroup has CHOL_MAX threads
constexpr int CHOL_T = 64;       // rows per diagonal block (one wavefront)

struct CholShared {
    double diag[2][CHOL_T * CHOL_T];   // diagonal tile of the current / next block
    double sub[2][CHOL_T * CHOL_T];    // next block's rows x current block's columns
    double x[CHOL_MAX];                // solved values, published per block
};

__device__ __forceinline__ double chol_bcast(double v, int lane)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(u)), lane));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(u >> 32)), lane));
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

// tiles A[r][c] = M[ra + r][ca + c] and B[r][c] = M[rb + r][cb + c] (zero
// outside the N x N matrix), copied by helper threads h = 0 .. H-1 with all
// loads of a round in flight before the stores; consecutive threads read one
// row segment
__device__ __forceinline__ void chol_stage2(double* __restrict__ A, int ra, int ca, double* __restrict__ B, int rb,
                                            int cb, const double* __restrict__ M, int N, int h, int H)
{
    constexpr int E = 2 * CHOL_T * CHOL_T;
    constexpr int R = (E + (CHOL_MAX - CHOL_T) - 1) / (CHOL_MAX - CHOL_T);   // rounds at H = CHOL_MAX - 64
    for (int e0 = h; e0 < E; e0 += R * H) {
        double v[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int e = e0 + u * H;
            const int f = e & (CHOL_T * CHOL_T - 1);
            const bool second = e >= CHOL_T * CHOL_T;
            const int gr = (second ? rb : ra) + (f >> 6), gc = (second ? cb : ca) + (f & 63);
            v[u] = (e < E && gr >= 0 && gr < N && gc >= 0 && gc < N) ? M[size_t(gr) * N + gc] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int e = e0 + u * H;
            if (e < E) (e >= CHOL_T * CHOL_T ? B : A)[e & (CHOL_T * CHOL_T - 1)] = v[u];
        }
    }
}

// s -= M[row][k0 + q] * xs[k0 + q] for q = 0 .. kn-1 (ascending) or kn-1 .. 0
// (descending), 16 loads in flight at a time
template <bool DESC>
__device__ __forceinline__ double chol_apply_global(double s, const double* __restrict__ Mrow, const double* xs,
                                                    int k0, int kn)
{
#pragma unroll 1
    for (int c = 0; c < kn; c += 16) {
        double l[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = DESC ? kn - 1 - (c + u) : c + u;
            l[u] = (c + u < kn) ? Mrow[k0 + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = DESC ? kn - 1 - (c + u) : c + u;
            if (c + u < kn) s = s - l[u] * xs[k0 + q];
        }
    }
    return s;
}

// L row-major N x N lower factor, LT its transpose (LT[i][k] = L[k][i]);
// mask bit (i % 3) of mask[i / 3] zeroes x_i.  Launch: one workgroup of
// CHOL_MAX threads, N <= CHOL_MAX.
__device__ __forceinline__ void chol_solve_body(int N, const double* __restrict__ L, const double* __restrict__ LT,
                                                const double* __restrict__ b, const uint8_t* __restrict__ mask,
                                                double* __restrict__ x)
{
    __shared__ CholShared sh;
    const int i = int(threadIdx.x);
    const int nt = int(blockDim.x);
    const int lane = i & (CHOL_T - 1);
    const int w = i / CHOL_T;
    const bool act = i < N;
    const int nb = (N + CHOL_T - 1) / CHOL_T;
    // rows at or beyond N carry s = +0 and divide by 1: they stay +0, and
    // their zero tile entries subtract +0 from the active rows (s - (+0) == s)
    double s = act ? b[i] : 0.0;

    // ---- forward: L y = b ----
    chol_stage2(sh.diag[0], 0, 0, sh.sub[0], 0, 0, L, N, i, nt);
    __syncthreads();
    for (int t = 0; t < nb; ++t) {
        const int k0 = t * CHOL_T;
        const int kn = min(CHOL_T, N - k0);
        const int cur = t & 1, nxt = cur ^ 1;
        if (w == t) {
            const double* T = sh.diag[cur] + lane * CHOL_T;
            const double d = act ? T[lane] : 1.0;
            double tr[16], tn[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) tr[u] = T[u];
#pragma unroll 1
            for (int c = 0; c < CHOL_T; c += 16) {
#pragma unroll
                for (int u = 0; u < 16; ++u) tn[u] = T[(c + 16 + u) & (CHOL_T - 1)];   // next chunk
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int q = c + u;
                    const double xq = s / d;
                    s = lane == q ? xq : s;
                    const double v = chol_bcast(s, q);
                    const double sn = s - tr[u] * v;
                    s = lane > q ? sn : s;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) tr[u] = tn[u];
            }
            sh.x[i] = s;
        } else {
            const int h = w < t ? i : i - CHOL_T;
            if (t + 1 < nb)
                chol_stage2(sh.diag[nxt], k0 + CHOL_T, k0 + CHOL_T, sh.sub[nxt], k0 + CHOL_T, k0, L, N, h,
                            nt - CHOL_T);
            // rows beyond the next block: the previous block's columns
            if (w > t && act && t >= 1) s = chol_apply_global<false>(s, L + size_t(i) * N, sh.x, k0 - CHOL_T, CHOL_T);
        }
        __syncthreads();
        if (w == t + 1 && t + 1 < nb) {
            const double* S = sh.sub[nxt] + lane * CHOL_T;
#pragma unroll 16
            for (int q = 0; q < kn; ++q) s = s - S[q] * sh.x[k0 + q];
        }
    }

    // ---- backward: L^T x = y, saxpy order ----
    __syncthreads();
    {
        const int k0 = (nb - 1) * CHOL_T;
        chol_stage2(sh.diag[0], k0, k0, sh.sub[0], k0, k0, LT, N, i, nt);
    }
    __syncthreads();
    for (int t = 0; t < nb; ++t) {
        const int blk = nb - 1 - t;
        const int k0 = blk * CHOL_T;
        const int kn = min(CHOL_T, N - k0);
        const int cur = t & 1, nxt = cur ^ 1;
        if (w == blk) {
            const double* T = sh.diag[cur] + lane * CHOL_T;   // T[q] = L[k0+q][k0+lane]
            const double d = act ? T[lane] : 1.0;
            double tr[16], tn[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) tr[u] = T[CHOL_T - 1 - u];
#pragma unroll 1
            for (int c = CHOL_T - 16; c >= 0; c -= 16) {
#pragma unroll
                for (int u = 0; u < 16; ++u) tn[u] = T[(c - 1 - u) & (CHOL_T - 1)];   // next chunk
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int q = c + 15 - u;
                    const double xq = s / d;
                    s = lane == q ? xq : s;
                    const double v = chol_bcast(s, q);
                    const double sn = s - tr[u] * v;
                    s = lane < q ? sn : s;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) tr[u] = tn[u];
            }
            sh.x[i] = s;
        } else {
            const int h = w < blk ? i : i - CHOL_T;
            if (blk > 0)
                chol_stage2(sh.diag[nxt], k0 - CHOL_T, k0 - CHOL_T, sh.sub[nxt], k0 - CHOL_T, k0, LT, N, h,
                            nt - CHOL_T);
            // rows before the previous block: the columns of the block after this one
            if (w < blk && act && blk + 1 < nb) {
                const int k1 = k0 + CHOL_T;
                s = chol_apply_global<true>(s, LT + size_t(i) * N, sh.x, k1, min(CHOL_T, N - k1));
            }
        }
        __syncthreads();
        if (w == blk - 1) {
            const double* S = sh.sub[nxt] + lane * CHOL_T;
#pragma unroll 16
            for (int q = kn - 1; q >= 0; --q) s = s - S[q] * sh.x[k0 + q];
        }
    }
    if (act) {
        const unsigned m = mask[i / 3];
        x[i] = ((m >> (i % 3)) & 1u) ? 0.0 : s;
    }
}

}  // namespace t3d
