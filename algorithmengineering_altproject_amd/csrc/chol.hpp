// chol.hpp -- the coarsest-level Cholesky solve L L^T x = b of the V-cycle,
// one workgroup, one row per thread (coarsest.cpp:92-122).  Shared by the
// FP64 multigrid (mg.hip) and the batched mixed solver (batch.hip).
// Forward: the row-form sum x_i = (b_i - sum_{k<i} L_ik x_k) / L_ii is evaluated
// as a column sweep that subtracts the k-th term from every row i > k at step k,
// so each row still subtracts its terms in increasing k (coarsest.cpp:101-108).
// Backward: the reference's saxpy form (coarsest.cpp:113-121), row k receives
// the updates of i = N-1 .. k+1 in that order.  Mask applied at the end.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace t3d {

constexpr int CHOL_MAX = 1024;

__device__ __forceinline__ void chol_solve_body(int N, const double* __restrict__ L, const double* __restrict__ LT,
                                                const double* __restrict__ b, const uint8_t* __restrict__ mask,
                                                double* __restrict__ x)
{
    __shared__ double xk;
    const int i = threadIdx.x;
    const bool act = i < N;
    double s = act ? b[i] : 0.0;
    // forward
    for (int k0 = 0; k0 < N; k0 += 8) {
        double l8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) l8[q] = (act && k0 + q < i) ? L[size_t(i) * N + k0 + q] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + q;
            if (k >= N) break;
            if (i == k) {
                s = s / L[size_t(k) * N + k];
                xk = s;
            }
            __syncthreads();
            const double v = xk;
            if (act && i > k) s = s - l8[q] * v;
            __syncthreads();
        }
    }
    // backward: s holds y_i
    for (int k0 = N - 1; k0 >= 0; k0 -= 8) {
        double l8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 - q;
            l8[q] = (act && k >= 0 && k > i) ? LT[size_t(i) * N + k] : 0.0;   // L[k][i]
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 - q;
            if (k < 0) break;
            if (i == k) {
                s = s / L[size_t(k) * N + k];
                xk = s;
            }
            __syncthreads();
            const double v = xk;
            if (act && i < k) s = s - l8[q] * v;
            __syncthreads();
        }
    }
    if (act) {
        const unsigned m = mask[i / 3];
        x[i] = ((m >> (i % 3)) & 1u) ? 0.0 : s;
    }
}

}  // namespace t3d
