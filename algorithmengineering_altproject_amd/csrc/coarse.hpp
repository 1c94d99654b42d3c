// coarse.hpp -- the coarsest-level solve of the mixed-precision V-cycle
// through the inverse Cholesky factor, in FP64.
//
// The reference solves L L^T x = b by forward and backward substitution
// (coarsest.cpp:92-122), a chain of N dependent divisions that one workgroup
// must walk row by row.  The mixed solver instead keeps Ainv = Li^T Li with
// Li = L^-1 (both computed on the device once per multigrid set-up) and applies
//     x = Ainv b                    (x masked on output, as the substitution)
// as one dense matrix-vector product spread over many workgroups: one
// wavefront per row, lanes striding over the row, a fixed-order butterfly
// sum.  Same factor, same FP64 arithmetic class; the rounding differs from the
// substitution at the level of the matrix's condition number times 2^-53.
// The FP64 plain path keeps the reference's substitution (chol.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "chol.hpp"

namespace t3d {

constexpr int CO_WAVES = 4;   // wavefronts per workgroup of the products
constexpr int CO_ROWS = 1;    // rows per wavefront (grid size = ceil(N / (CO_WAVES * CO_ROWS)))

inline unsigned coarse_blocks(int N)
{
    return unsigned((N + CO_WAVES * CO_ROWS - 1) / (CO_WAVES * CO_ROWS));
}

// column j of Li = L^-1 (row-major N x N, zero above the diagonal on entry):
//   Li[j][j] = 1 / L[j][j],  Li[i][j] = -(sum_{k=j}^{i-1} L[i][k] Li[k][j]) / L[i][i]
// one thread per column; also writes LiT = Li^T when given.  Each thread reads back only
// the column it wrote itself (and zeros above the diagonal).
__device__ __forceinline__ void tri_inv_body(int N, const double* __restrict__ L, double* Li, double* LiT, int j0,
                                             int lane)
{
    const int j = j0 + lane;
    const bool act = j < N;
    const int jc = act ? j : j0;   // inactive lanes read a valid column and write nothing
    for (int i = j0; i < N; ++i) {
        const double* Lr = L + size_t(i) * N;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = j0;
        for (; k + 4 <= i; k += 4) {
            a0 = fma(Lr[k], Li[size_t(k) * N + jc], a0);
            a1 = fma(Lr[k + 1], Li[size_t(k + 1) * N + jc], a1);
            a2 = fma(Lr[k + 2], Li[size_t(k + 2) * N + jc], a2);
            a3 = fma(Lr[k + 3], Li[size_t(k + 3) * N + jc], a3);
        }
        for (; k < i; ++k) a0 = fma(Lr[k], Li[size_t(k) * N + jc], a0);
        if (act && i >= j) {
            const double v = (i == j) ? 1.0 / Lr[i] : -((a0 + a1) + (a2 + a3)) / Lr[i];
            Li[size_t(i) * N + j] = v;
            if (LiT) LiT[size_t(j) * N + i] = v;
        }
    }
}

// broadcast of lane l's double to the wavefront (l uniform)
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b & 0xffffffffll), l));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Columns [j0, j0 + NC) of Li = L^-1 (L row-major N x N, lower; its upper part
// is never read) by column-oriented forward substitution in 32-row blocks, one
// 256-thread workgroup per column group.  Thread t owns rows t, t + 256, ...,
// t + 256 (R - 1) and keeps their entries of the NC columns in registers
// (x = e_j at the start).  For each block from the one holding j0:
//   * the wavefront owning the block's 32 rows (lanes l0 .. l0 + 31) solves the
//     diagonal block: y_k = x_k / L[k][k] broadcast by readlane, then
//     x_i -= L[i][k] y_k for the block's later rows, in k order;
//   * the block's y go to LDS (double-buffered: one barrier per block);
//   * every thread subtracts L[i][k] y_k, k over the block in order, from its
//     rows below the block, reading its own 32-entry segment of row i.
// Rows above j0 stay zero, so all N entries of each column are written (the
// zero upper part of Li included).  Requires N <= 256 R.  The sums run in a
// different order from tri_inv_body; both are FP64 forward substitutions.
template <int R, int NC>
__device__ __forceinline__ void tri_inv_blk_body(int N, const double* __restrict__ L, double* __restrict__ Li,
                                                 int j0)
{
    __shared__ double ys[2][32][NC];
    const int t = int(threadIdx.x), lane = t & 63, wave = t >> 6;
    double x[R][NC];
#pragma unroll
    for (int m = 0; m < R; ++m)
#pragma unroll
        for (int c = 0; c < NC; ++c) x[m][c] = (t + 256 * m == j0 + c) ? 1.0 : 0.0;
    int buf = 0;
    for (int k0 = j0 & ~31; k0 < N; k0 += 32, buf ^= 1) {
        const int mk = k0 >> 8, wk = (k0 & 255) >> 6, l0 = k0 & 63;
        if (wave == wk) {
            const int r = lane - l0, i = k0 + r;
            const bool blk = r >= 0 && r < 32;
            const bool own = blk && i < N;
            double xv[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) xv[c] = 0.0;
#pragma unroll
            for (int m = 0; m < R; ++m)
                if (m == mk)
#pragma unroll
                    for (int c = 0; c < NC; ++c) xv[c] = x[m][c];
            const double* Lr = L + size_t(own ? i : 0) * N + k0;
            double lr[32];
#pragma unroll
            for (int s = 0; s < 32; ++s) lr[s] = (own && s < r) ? Lr[s] : 0.0;   // k0 + s < i < N
            const double dinv = own ? 1.0 / Lr[r] : 0.0;
#pragma unroll
            for (int s = 0; s < 32; ++s) {
                if (k0 + s < N) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const double y = readlane_f64(xv[c] * dinv, l0 + s);
                        if (r == s) xv[c] = y;
                        else if (r > s) xv[c] = fma(-lr[s], y, xv[c]);
                    }
                }
            }
            if (blk) {
#pragma unroll
                for (int c = 0; c < NC; ++c) ys[buf][r][c] = own ? xv[c] : 0.0;
#pragma unroll
                for (int m = 0; m < R; ++m)
                    if (m == mk)
#pragma unroll
                        for (int c = 0; c < NC; ++c) x[m][c] = xv[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < R; ++m) {
            if (256 * m + 255 < k0 + 32) continue;   // every row of this slot is above the block's end
            const int i = t + 256 * m;
            if (i >= k0 + 32 && i < N) {
                const double* Lr = L + size_t(i) * N + k0;
                double a[32];
#pragma unroll
                for (int s = 0; s < 32; ++s) a[s] = Lr[s];
#pragma unroll
                for (int s = 0; s < 32; ++s)
#pragma unroll
                    for (int c = 0; c < NC; ++c) x[m][c] = fma(-a[s], ys[buf][s][c], x[m][c]);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int i = t + 256 * m;
        if (i < N)
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (j0 + c < N) Li[size_t(i) * N + j0 + c] = x[m][c];
    }
}

template <int R, int NC>
static __global__ void __launch_bounds__(256) k_tri_inv_blk(int N, const double* __restrict__ L,
                                                            double* __restrict__ Li)
{
    tri_inv_blk_body<R, NC>(N, L, Li, int(blockIdx.x) * NC);
}

static __global__ void __launch_bounds__(64) k_tri_inv_cols(int N, const double* __restrict__ L, double* Li)
{
    tri_inv_body(N, L, Li, nullptr, int(blockIdx.x) * 64, int(threadIdx.x));
}

// Li = L^-1 on stream st: the blocked form up to N = 8192, else one thread per
// column (Li cleared first, as that form writes only the lower part)
static inline hipError_t launch_tri_inv(hipStream_t st, int N, const double* L, double* Li)
{
    if (N <= 2048)
        hipLaunchKernelGGL((k_tri_inv_blk<8, 4>), dim3(unsigned((N + 3) / 4)), dim3(256), 0, st, N, L, Li);
    else if (N <= 4096)
        hipLaunchKernelGGL((k_tri_inv_blk<16, 4>), dim3(unsigned((N + 3) / 4)), dim3(256), 0, st, N, L, Li);
    else if (N <= 8192)
        hipLaunchKernelGGL((k_tri_inv_blk<32, 2>), dim3(unsigned((N + 1) / 2)), dim3(256), 0, st, N, L, Li);
    else {
        const hipError_t e = hipMemsetAsync(Li, 0, sizeof(double) * size_t(N) * size_t(N), st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_tri_inv_cols, dim3(unsigned((N + 63) / 64)), dim3(64), 0, st, N, L, Li);
    }
    return hipGetLastError();
}

// Ainv = Li^T Li for the lower-triangular Li (row-major N x N, zero above the
// diagonal): the coarsest solve becomes one dense product.  Output tile
// (ti, tj) of 64 x 64 per 256-thread workgroup, 4 x 4 entries per thread,
// k from the tile's first nonzero row in chunks of 32 staged in LDS.
__device__ __forceinline__ void gram_body(int N, const double* __restrict__ Li, double* __restrict__ Ainv, int ti,
                                          int tj)
{
    __shared__ double Sa[32][65], Sb[32][65];
    const int i0 = 64 * ti, j0 = 64 * tj;
    const int t = int(threadIdx.x);
    const int ri = (t / 16) * 4, rj = (t % 16) * 4;
    double acc[4][4] = {};
    for (int k0 = (i0 > j0 ? i0 : j0) & ~31; k0 < N; k0 += 32) {
        for (int e = t; e < 32 * 64; e += 256) {
            const int k = e / 64, c = e % 64;
            const bool kin = k0 + k < N;
            Sa[k][c] = (kin && i0 + c < N) ? Li[size_t(k0 + k) * N + i0 + c] : 0.0;
            Sb[k][c] = (kin && j0 + c < N) ? Li[size_t(k0 + k) * N + j0 + c] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < 32; ++k)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(Sa[k][ri + a], Sb[k][rj + b], acc[a][b]);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (i0 + ri + a < N && j0 + rj + b < N) Ainv[size_t(i0 + ri + a) * N + j0 + rj + b] = acc[a][b];
}

// rows gw, gw + GW, ... of out = M v over the lower (TRI = 0: k <= i) or upper
// (TRI = 1: k >= i) triangle, or all (TRI = 2), of the row-major N x N matrix
// M, one wavefront per row.  With mask, masked rows are zeroed; with out32, an
// FP32 copy is written too.
template <int TRI>
__device__ __forceinline__ void coarse_mv_body(int N, const double* __restrict__ M, const double* __restrict__ v,
                                               const uint8_t* __restrict__ mask, double* __restrict__ out,
                                               float* __restrict__ out32, int gw, int GW)
{
    const int lane = int(threadIdx.x) & 63;
    for (int i = gw; i < N; i += GW) {
        const double* Mi = M + size_t(i) * N;
        const int k0 = TRI == 1 ? i : 0;
        const int k1 = TRI == 0 ? i + 1 : N;
        // eight independent partial sums, so a wavefront has eight row loads in flight
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int u = 0; k0 + u * 64 < k1; u += 8) {
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                const int k = k0 + (u + w) * 64 + lane;
                if (k < k1) acc[w] = fma(Mi[k], v[k], acc[w]);
            }
        }
        double a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
        if (lane == 0) {
            if (mask && ((mask[i / 3] >> (i % 3)) & 1u)) a = 0.0;
            out[i] = a;
            if (out32) out32[i] = float(a);
        }
    }
}

}  // namespace t3d
