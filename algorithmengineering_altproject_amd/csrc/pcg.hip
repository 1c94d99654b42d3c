// pcg.hip -- device-resident preconditioned conjugate gradients.
//
// Reference: PCG_free (cpp/src/solver.cpp:87-217).  The whole iteration stays
// on the device: alpha/beta are formed by the kernels that consume them from
// per-iteration scalar slots, the convergence test sets a device flag that
// turns every later kernel of the solve into a no-op, and the host only polls
// the flag between chunks of enqueued iterations.  Dot products use the
// canonical reduction tree; all vectors are in the full node-major layout.
#include <hip/hip_runtime.h>

#include <cmath>

#include "t3d_internal.hpp"

namespace t3d {

// per-iteration scalar slots in p->pcgSc.  The direction update
// p = z + beta p (solver.cpp:203-210) and the solution update x += alpha p
// (solver.cpp:184-188) are done by the next iteration's operator
// (launch_kxu_f64_fusep; after the last iteration by k_pcg_fold_x) and r.z by
// the V-cycle's last pass (mg_apply_dev), in the same arithmetic and
// summation order.
struct PcgSlots {
    double* normb2;   // [1]
    double* stop;     // [1]
    double* rz;       // [maxIt + 1]
    double* denom;    // [maxIt]
    double* rn2;      // [maxIt]
};

__global__ void k_sub(int64_t n, const double* __restrict__ b, const double* __restrict__ Ap, double* __restrict__ r)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - Ap[i];
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_pcg_stop(const double* __restrict__ normb2, double tol, double* __restrict__ stop,
                           int* __restrict__ flags)
{
    const double nb = sqrt(*normb2);
    *stop = (1.0e-30 < nb ? nb : 1.0e-30) * tol;
    flags[0] = 0;
    flags[1] = -1;
}

// x += alpha p ; r = r - alpha Ap ; tile partials of r.r (canonical layout)
// r -= alpha Ap with the canonical tile partials of r.r (solver.cpp:189-199);
// x += alpha p of this iteration is applied by the next operator
__global__ void __launch_bounds__(256) k_pcg_update_xr(int64_t n, const double* __restrict__ rz,
                                                       const double* __restrict__ denom,
                                                       const double* __restrict__ Ap,
                                                       double* __restrict__ r, double* __restrict__ partial,
                                                       const int* __restrict__ done)
{
    if (*done) return;
    __shared__ double sh[CANON_LANES];
    const double d = *denom;
    const double alpha = *rz / (1.0e-30 < d ? d : 1.0e-30);
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) {
            const double v = r[i] - alpha * Ap[i];
            r[i] = v;
            s = s + v * v;
        }
    }
    // canonical tile tree (same as k_canon_dot_tiles)
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

// x += alpha p of the last iteration run (the one no later operator applies)
__global__ void k_pcg_fold_x(int64_t n, const double* __restrict__ rz, const double* __restrict__ denom,
                             const double* __restrict__ p, double* __restrict__ x)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const double d = *denom;
    const double alpha = *rz / (1.0e-30 < d ? d : 1.0e-30);
    if (i < n) x[i] = x[i] + alpha * p[i];
}

__global__ void k_pcg_check(const double* __restrict__ rn2, const double* __restrict__ stop, int it,
                            int* __restrict__ flags)
{
    if (flags[0]) return;
    if (sqrt(*rn2) < *stop) {
        flags[0] = 1;
        flags[1] = it + 1;
    }
}

static int pcg_alloc(t3d_problem* p, int maxIt)
{
    const size_t nD = size_t(3 * p->g.nN);
    // each work vector on first use (the mixed-precision solver shares wAp)
    if (!p->wr) T3D_HIP(hipMalloc(&p->wr, sizeof(double) * nD));
    if (!p->wz) T3D_HIP(hipMalloc(&p->wz, sizeof(double) * nD));
    if (!p->wp) T3D_HIP(hipMalloc(&p->wp, sizeof(double) * nD));
    if (!p->wp2) T3D_HIP(hipMalloc(&p->wp2, sizeof(double) * nD));
    if (!p->wAp) T3D_HIP(hipMalloc(&p->wAp, sizeof(double) * nD));
    if (!p->pcgFlags) T3D_HIP(hipMalloc(&p->pcgFlags, sizeof(int) * 4));
    if (!p->hostFlags) T3D_HIP(hipHostMalloc(&p->hostFlags, sizeof(int) * 4, hipHostMallocDefault));
    const int need = 3 * (maxIt + 1) + 8;
    if (need > p->pcgScCap) {
        if (p->pcgSc) T3D_HIP(hipFree(p->pcgSc));
        T3D_HIP(hipMalloc(&p->pcgSc, sizeof(double) * need));
        p->pcgScCap = need;
    }
    return ensure_red(p, int64_t(nD));
}

}  // namespace t3d

using namespace t3d;

extern "C" {

int t3d_pcg(t3d_problem* p, t3d_mg* M, const double* b, double* x, double tol, int maxIt, int* iters)
{
    T3D_CHECK(p && b && x && maxIt >= 0, T3D_ERR_ARG, "t3d_pcg: bad argument");
    int rc = pcg_alloc(p, maxIt);
    if (rc) return rc;
    hipStream_t st = p->stream;
    const int64_t n = 3 * p->g.nN;
    const int B = 256;
    const unsigned nb = unsigned((n + B - 1) / B);
    const unsigned ntile = unsigned((n + CANON_TILE - 1) / CANON_TILE);
    PcgSlots S;
    S.normb2 = p->pcgSc;
    S.stop = p->pcgSc + 1;
    S.rz = p->pcgSc + 8;
    S.denom = S.rz + (maxIt + 1);
    S.rn2 = S.denom + maxIt;
    int* flags = p->pcgFlags;
    double *r = p->wr, *z = p->wz, *pp = p->wp, *Ap = p->wAp;
    // p of iteration it lives in pb[it & 1]: the operator of iteration it >= 1
    // forms it from z (r without a preconditioner) and p of iteration it - 1
    double* pb[2] = {p->wp, p->wp2};
    // r = b - K x
    if ((rc = launch_kxu_f64(p, x, Ap))) return rc;
    hipLaunchKernelGGL(k_sub, dim3(nb), dim3(B), 0, st, n, b, Ap, r);
    T3D_HIP(hipGetLastError());
    const double* zr = r;
    if (M) {
        if ((rc = mg_apply_dev(M, r, z, nullptr))) return rc;
        zr = z;
    }
    if ((rc = canon_dot_dev(p, r, zr, n, S.rz))) return rc;
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(B), 0, st, n, zr, pb[0]);
    T3D_HIP(hipGetLastError());
    if ((rc = canon_dot_dev(p, b, b, n, S.normb2))) return rc;
    hipLaunchKernelGGL(k_pcg_stop, dim3(1), dim3(1), 0, st, S.normb2, tol, S.stop, flags);
    T3D_HIP(hipGetLastError());
    const int CHUNK = 32;
    int result = maxIt;
    for (int it0 = 0; it0 < maxIt; it0 += CHUNK) {
        const int it1 = std::min(maxIt, it0 + CHUNK);
        for (int it = it0; it < it1; ++it) {
            pp = pb[it & 1];
            if (it == 0) rc = launch_kxu_f64(p, pp, Ap, flags);
            else rc = launch_kxu_f64_fusep(p, M ? z : r, pb[(it - 1) & 1], S.rz + it, S.rz + it - 1, pp, Ap, flags,
                                           x, S.denom + it - 1);
            if (rc) return rc;
            if ((rc = canon_dot_dev(p, pp, Ap, n, S.denom + it, flags))) return rc;
            hipLaunchKernelGGL(k_pcg_update_xr, dim3(ntile), dim3(CANON_LANES), 0, st, n, S.rz + it, S.denom + it,
                               Ap, r, p->red, flags);
            T3D_HIP(hipGetLastError());
            if ((rc = canon_finish(p, ntile, S.rn2 + it, flags))) return rc;
            hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(1), 0, st, S.rn2 + it, S.stop, it, flags);
            T3D_HIP(hipGetLastError());
            // z = M r with r.z in the V-cycle's last pass; p = z + beta p is
            // formed by the next iteration's operator
            if (M) {
                if ((rc = mg_apply_dev(M, r, z, flags, S.rz + it + 1))) return rc;
            } else {
                hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, st, S.rn2 + it, S.rz + it + 1, flags);
                T3D_HIP(hipGetLastError());
            }
        }
        T3D_HIP(hipMemcpyAsync(p->hostFlags, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        if (p->hostFlags[0]) {
            result = p->hostFlags[1];
            break;
        }
    }
    if (result > 0) {
        const int k = result - 1;
        hipLaunchKernelGGL(k_pcg_fold_x, dim3(nb), dim3(B), 0, st, n, S.rz + k, S.denom + k, pb[k & 1], x);
        T3D_HIP(hipGetLastError());
    }
    T3D_HIP(hipStreamSynchronize(st));
    if (iters) *iters = result;
    return T3D_OK;
}

// per-iteration scalars of the last solve: rn2[k] = ||r_k+1||^2 (k < n)
int t3d_pcg_history(t3d_problem* p, int maxIt, int n, double* rn2_host)
{
    T3D_CHECK(p && p->pcgSc && rn2_host && n <= maxIt, T3D_ERR_ARG, "t3d_pcg_history: bad argument");
    const double* rn2 = p->pcgSc + 8 + (maxIt + 1) + maxIt;
    T3D_HIP(hipDeviceSynchronize());
    T3D_HIP(hipMemcpy(rn2_host, rn2, sizeof(double) * n, hipMemcpyDeviceToHost));
    return T3D_OK;
}

}  // extern "C"
