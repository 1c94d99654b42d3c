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
