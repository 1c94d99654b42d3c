// filter.hip -- PDE (Helmholtz) filter on the device.
//
// Reference: SetupPDEFilter / MatTimesVec_PDE / ApplyPDEFilter
// (cpp/src/filter.cpp:9-275).  Node-centric gathers reproduce the reference's
// accumulation orders exactly:
//   * element -> node right-hand side and the Jacobi diagonal: contributions in
//     increasing compact (Morton) element order (filter.cpp:63-69, 177-182);
//   * the 8x8 matvec: incident elements in colour order 0..7, each row a
//     sequential 8-term sum (filter.cpp:98-127);
//   * node -> element: 8 node values summed in local order, times 1/8, to float
//     (filter.cpp:267-271).
// CG reductions use the canonical tree; the CG loop runs device-resident with
// the same early-exit rules (absolute 1e-2 before, 1e-3 inside, at most 80).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "t3d_internal.hpp"

namespace t3d {

__device__ __forceinline__ int fcorner(int dx, int dy, int dz) { return (dy ? (dx ? 1 : 0) : (dx ? 2 : 3)) + 4 * dz; }
__device__ __forceinline__ void fcorner_off(int a, int& ox, int& oy, int& oz)
{
    const int b = a & 3;
    ox = (b == 1 || b == 2) ? 1 : 0;
    oy = (b <= 1) ? 1 : 0;
    oz = a >> 2;
}

// incident solid cells of node (ix,iy,iz): compact ids and local corner, sorted by compact id
__device__ __forceinline__ int incident_sorted(const Grid& g, const int32_t* __restrict__ cell2c, int ix, int iy, int iz,
                                               int* ids, int* corner)
{
    int k = 0;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = ix - dx, ty = iy - dy, cz = iz - dz;
                if (cx < 0 || ty < 0 || cz < 0 || cx >= g.resX || ty >= g.resY || cz >= g.resZ) continue;
                const int e = cell2c[(int64_t(cz) * g.resY + ty) * g.resX + cx];
                if (e < 0) continue;
                // insertion by compact id
                int j = k++;
                const int a = fcorner(dx, dy, dz);
                while (j > 0 && ids[j - 1] > e) {
                    ids[j] = ids[j - 1];
                    corner[j] = corner[j - 1];
                    --j;
                }
                ids[j] = e;
                corner[j] = a;
            }
    return k;
}

// rhs[n] = sum over incident elements (compact order) of (double)src[e] * (1/8)
__global__ void __launch_bounds__(256) k_filter_rhs(Grid g, const int32_t* __restrict__ cell2c,
                                                    const float* __restrict__ src, double* __restrict__ rhs)
{
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= g.nN) return;
    const int ix = int(n % g.nnx), iy = int((n / g.nnx) % g.nny), iz = int(n / (int64_t(g.nnx) * g.nny));
    int ids[8], cor[8];
    const int k = incident_sorted(g, cell2c, ix, iy, iz, ids, cor);
    double s = 0.0;
    for (int j = 0; j < k; ++j) s = s + double(src[ids[j]]) * (1.0 / 8.0);
    rhs[n] = s;
}

// y = A x, node-centric, colour order (MatTimesVec_PDE)
__global__ void __launch_bounds__(256) k_pde_matvec(Grid g, const int32_t* __restrict__ cell2c,
                                                    const double* __restrict__ Kf, const double* __restrict__ x,
                                                    double* __restrict__ y, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= g.nN) return;
    const int ix = int(n % g.nnx), iy = int((n / g.nnx) % g.nny), iz = int(n / (int64_t(g.nnx) * g.nny));
    double acc = 0.0;
    for (int col = 0; col < 8; ++col) {
        const int dx = (ix ^ col) & 1, dy = (iy ^ (col >> 1) ^ 1) & 1, dz = (iz ^ (col >> 2)) & 1;
        const int cx = ix - dx, ty = iy - dy, cz = iz - dz;
        if (cx < 0 || ty < 0 || cz < 0 || cx >= g.resX || ty >= g.resY || cz >= g.resZ) continue;
        if (cell2c[(int64_t(cz) * g.resY + ty) * g.resX + cx] < 0) continue;
        const int a = fcorner(dx, dy, dz);
        double f = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            int ox, oy, oz;
            fcorner_off(b, ox, oy, oz);
            const int64_t nb = (int64_t(cz + oz) * g.nny + (ty + oy)) * g.nnx + (cx + ox);
            f = f + Kf[a * 8 + b] * x[nb];
        }
        acc = acc + f;
    }
    y[n] = acc;
}

// x0 = warm ? lastX : 0 (warm decided from ||rhs|| and ||rhs - lastRhs||, filter.cpp:185-199)
__global__ void k_filter_warm(const double* __restrict__ nrm2, const double* __restrict__ dif2,
                              int* __restrict__ flags)
{
    const double nr = sqrt(*nrm2), df = sqrt(*dif2);
    const int warm = (nr > 0.0) && (df / (1.0e-30 < nr ? nr : 1.0e-30) < 1000.0);
    flags[2] = warm;   // 0 done, 1 iterations, 2 warm
    flags[0] = 0;
    flags[1] = -1;
}
__global__ void k_filter_x0(int64_t n, const double* __restrict__ lastX, const int* __restrict__ flags,
                            double* __restrict__ x)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = flags[2] ? lastX[i] : 0.0;
}

// canonical tile partials of (a[i]-b[i])^2
__global__ void __launch_bounds__(256) k_tiles_diff2(int64_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ out)
{
    __shared__ double sh[CANON_LANES];
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) {
            const double d = a[i] - b[i];
            s = s + d * d;
        }
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
        if (l == 0) out[blockIdx.x] = v;
    }
}

// r = rhs - Ap ; z = diag * r ; tile partials of r.z
__global__ void __launch_bounds__(256) k_filter_r0(int64_t n, const double* __restrict__ rhs,
                                                   const double* __restrict__ Ap, const double* __restrict__ diag,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   double* __restrict__ out)
{
    __shared__ double sh[CANON_LANES];
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) {
            const double rv = rhs[i] - Ap[i];
            r[i] = rv;
            const double zv = diag[i] * rv;
            z[i] = zv;
            s = s + rv * zv;
        }
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
        if (l == 0) out[blockIdx.x] = v;
    }
}

// rz: if 0 -> 1 ; early exit when ||r|| < 1e-2 (filter.cpp:219-239)
__global__ void k_filter_start(double* __restrict__ rz, const double* __restrict__ rn2, int* __restrict__ flags)
{
    if (*rz == 0) *rz = 1.0;
    if (sqrt(*rn2) < 1.0e-2) flags[0] = 1;
}

// x += alpha p ; r -= alpha Ap ; partials of r.r
__global__ void __launch_bounds__(256) k_filter_xr(int64_t n, const double* __restrict__ rz,
                                                   const double* __restrict__ denom, const double* __restrict__ p,
                                                   const double* __restrict__ Ap, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ out,
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
            x[i] += alpha * p[i];
            const double rv = r[i] - alpha * Ap[i];
            r[i] = rv;
            s = s + rv * rv;
        }
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
        if (l == 0) out[blockIdx.x] = v;
    }
}

__global__ void k_filter_check(const double* __restrict__ rn2, int it, int* __restrict__ flags)
{
    if (flags[0]) return;
    flags[1] = it + 1;
    if (sqrt(*rn2) < 1.0e-3) flags[0] = 1;
}

// z = diag*r ; partials of r.z  (guarded)
__global__ void __launch_bounds__(256) k_filter_z(int64_t n, const double* __restrict__ r,
                                                  const double* __restrict__ diag, double* __restrict__ z,
                                                  double* __restrict__ out, const int* __restrict__ done)
{
    if (*done) return;
    __shared__ double sh[CANON_LANES];
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) {
            const double zv = diag[i] * r[i];
            z[i] = zv;
            s = s + r[i] * zv;
        }
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
        if (l == 0) out[blockIdx.x] = v;
    }
}

__global__ void k_filter_p(int64_t n, const double* __restrict__ rz_new, const double* __restrict__ rz_old,
                           const double* __restrict__ z, double* __restrict__ p, const int* __restrict__ done)
{
    if (*done) return;
    const double o = *rz_old;
    const double beta = *rz_new / (1.0e-30 < o ? o : 1.0e-30);
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// dst[e] = (float)((sum_a x[n_a]) * (1/8))
__global__ void k_filter_out(Grid g, int ne, const int32_t* __restrict__ c2cell, const double* __restrict__ x,
                             float* __restrict__ dst)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int cell = c2cell[e];
    const int cx = cell % g.resX, ty = (cell / g.resX) % g.resY, cz = cell / (g.resX * g.resY);
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        int ox, oy, oz;
        fcorner_off(a, ox, oy, oz);
        s = s + x[(int64_t(cz + oz) * g.nny + (ty + oy)) * g.nnx + (cx + ox)];
    }
    dst[e] = float(s * (1.0 / 8.0));
}

}  // namespace t3d

using namespace t3d;

struct t3d_filter {
    t3d_problem* p = nullptr;
    double kernel[64];
    double* dK = nullptr;      // [64]
    double* diag = nullptr;    // [nN]
    double* lastX = nullptr;   // [nN]
    double* lastRhs = nullptr; // [nN]
    double *rhs = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *pp = nullptr, *Ap = nullptr;
    double* sc = nullptr;      // scalars: [0] nrm2 [1] dif2 [2] rz0 [3] rn0 ; per it: rz[81], denom[80], rn2[80]
    int* flags = nullptr;      // [0] done [1] iterations [2] warm
    int hostFlags[4] = {0, 0, 0, 0};
};

extern "C" {

int t3d_filter_destroy(t3d_filter* f)
{
    if (!f) return T3D_OK;
    void* b[] = {f->dK, f->diag, f->lastX, f->lastRhs, f->rhs, f->x, f->r, f->z, f->pp, f->Ap, f->sc, f->flags};
    for (void* q : b)
        if (q) (void)hipFree(q);
    delete f;
    return T3D_OK;
}

int t3d_filter_create(t3d_problem* p, float rMin, t3d_filter** out)
{
    T3D_CHECK(p && out, T3D_ERR_ARG, "t3d_filter_create: null argument");
    *out = nullptr;
    t3d_filter* f = new t3d_filter();
    f->p = p;
    // 8x8 Helmholtz kernel iRmin^2 * KEF0 + KEF1 with 2x2x2 Gauss points (filter.cpp:9-54)
    const int S[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int T[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int P[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double N[8][8], G[24][8];
    for (int q = 0; q < 8; ++q) {
        const double s = S[q] / std::sqrt(3.0), t = T[q] / std::sqrt(3.0), u = P[q] / std::sqrt(3.0);
        for (int a = 0; a < 8; ++a) {
            const double fs = S[a] < 0 ? (1 - s) : (1 + s);
            const double ft = T[a] < 0 ? (1 - t) : (1 + t);
            const double fu = P[a] < 0 ? (1 - u) : (1 + u);
            N[q][a] = 0.125 * fs * ft * fu;
            G[3 * q][a] = (S[a] < 0 ? -0.125 : 0.125) * ft * fu;
            G[3 * q + 1][a] = (T[a] < 0 ? -0.125 : 0.125) * fs * fu;
            G[3 * q + 2][a] = (P[a] < 0 ? -0.125 : 0.125) * fs * ft;
        }
    }
    const double iR = (double(rMin) * double(1.0f)) / (2.0 * std::sqrt(3.0));
    double rowSum[8];
    for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < 8; ++j) {
            double k0 = 0.0, k1 = 0.0;
            for (int k = 0; k < 24; ++k) k0 += G[k][i] * G[k][j];
            for (int q = 0; q < 8; ++q) k1 += N[q][i] * (1.0 / 8.0) * N[q][j];
            f->kernel[i * 8 + j] = iR * iR * k0 + k1;
        }
    }
    for (int a = 0; a < 8; ++a) {
        double s = 0.0;
        for (int b = 0; b < 8; ++b) s += f->kernel[a * 8 + b];
        rowSum[a] = s;
    }
    // Jacobi diagonal: element order contributions of the row sums (filter.cpp:62-70)
    const HostModel& m = p->host;
    std::vector<double> dg(p->g.nN, 0.0);
    for (int e = 0; e < m.numElements; ++e)
        for (int a = 0; a < 8; ++a) dg[m.eNodMat[size_t(8) * e + a]] += rowSum[a];
    for (double& v : dg) v = v > 0 ? 1.0 / v : 1.0;
    const size_t nN = size_t(p->g.nN);
    const size_t tiles = (nN + CANON_TILE - 1) / CANON_TILE;
    (void)tiles;
    auto fail = [&]() {
        set_error("t3d_filter_create: device allocation failed");
        t3d_filter_destroy(f);
        return T3D_ERR_HIP;
    };
    if (hipMalloc(&f->dK, sizeof(double) * 64) != hipSuccess ||
        hipMemcpy(f->dK, f->kernel, sizeof(double) * 64, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&f->diag, sizeof(double) * nN) != hipSuccess ||
        hipMemcpy(f->diag, dg.data(), sizeof(double) * nN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&f->lastX, sizeof(double) * nN) != hipSuccess ||
        hipMalloc(&f->lastRhs, sizeof(double) * nN) != hipSuccess ||
        hipMemset(f->lastX, 0, sizeof(double) * nN) != hipSuccess ||
        hipMemset(f->lastRhs, 0, sizeof(double) * nN) != hipSuccess ||
        hipMalloc(&f->rhs, sizeof(double) * nN) != hipSuccess || hipMalloc(&f->x, sizeof(double) * nN) != hipSuccess ||
        hipMalloc(&f->r, sizeof(double) * nN) != hipSuccess || hipMalloc(&f->z, sizeof(double) * nN) != hipSuccess ||
        hipMalloc(&f->pp, sizeof(double) * nN) != hipSuccess || hipMalloc(&f->Ap, sizeof(double) * nN) != hipSuccess ||
        hipMalloc(&f->sc, sizeof(double) * 512) != hipSuccess || hipMalloc(&f->flags, sizeof(int) * 4) != hipSuccess)
        return fail();
    if (ensure_red(p, int64_t(nN))) return fail();
    // the set-up copies and fills above ran on the null stream
    T3D_HIP(hipDeviceSynchronize());
    *out = f;
    return T3D_OK;
}

int t3d_filter_info(const t3d_filter* f, double* kernel64, double* diag_host)
{
    T3D_CHECK(f, T3D_ERR_ARG, "t3d_filter_info: null");
    if (kernel64) std::memcpy(kernel64, f->kernel, sizeof(double) * 64);
    T3D_HIP(hipDeviceSynchronize());
    if (diag_host) T3D_HIP(hipMemcpy(diag_host, f->diag, sizeof(double) * f->p->g.nN, hipMemcpyDeviceToHost));
    return T3D_OK;
}

// ApplyPDEFilter (filter.cpp:175-275): src/dst are float element arrays (compact order, device)
int t3d_filter_apply_async(t3d_filter* f, const float* src, float* dst)
{
    t3d_problem* p = f->p;
    hipStream_t st = p->stream;
    const Grid& g = p->g;
    const int64_t n = g.nN;
    const int B = 256;
    const unsigned nb = unsigned((n + B - 1) / B);
    const unsigned nt = unsigned((n + CANON_TILE - 1) / CANON_TILE);
    double* S = f->sc;
    double* rz = S + 8;       // [81]
    double* denom = rz + 81;  // [80]
    double* rn2 = denom + 80; // [80]
    int rc;
    hipLaunchKernelGGL(k_filter_rhs, dim3(nb), dim3(B), 0, st, g, p->cellToCompact, src, f->rhs);
    if ((rc = canon_dot_dev(p, f->rhs, f->rhs, n, S + 0))) return rc;
    hipLaunchKernelGGL(k_tiles_diff2, dim3(nt), dim3(CANON_LANES), 0, st, n, f->rhs, f->lastRhs, p->red);
    if ((rc = canon_finish(p, nt, S + 1))) return rc;
    hipLaunchKernelGGL(k_filter_warm, dim3(1), dim3(1), 0, st, S + 0, S + 1, f->flags);
    hipLaunchKernelGGL(k_filter_x0, dim3(nb), dim3(B), 0, st, n, f->lastX, f->flags, f->x);
    hipLaunchKernelGGL(k_pde_matvec, dim3(nb), dim3(B), 0, st, g, p->cellToCompact, f->dK, f->x, f->Ap, nullptr);
    hipLaunchKernelGGL(k_filter_r0, dim3(nt), dim3(CANON_LANES), 0, st, n, f->rhs, f->Ap, f->diag, f->r, f->z,
                       p->red);
    if ((rc = canon_finish(p, nt, rz))) return rc;
    T3D_HIP(hipMemcpyAsync(f->pp, f->z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if ((rc = canon_dot_dev(p, f->r, f->r, n, S + 3))) return rc;
    hipLaunchKernelGGL(k_filter_start, dim3(1), dim3(1), 0, st, rz, S + 3, f->flags);
    for (int it = 0; it < 80; ++it) {
        hipLaunchKernelGGL(k_pde_matvec, dim3(nb), dim3(B), 0, st, g, p->cellToCompact, f->dK, f->pp, f->Ap,
                           f->flags);
        if ((rc = canon_dot_dev(p, f->pp, f->Ap, n, denom + it, f->flags))) return rc;
        hipLaunchKernelGGL(k_filter_xr, dim3(nt), dim3(CANON_LANES), 0, st, n, rz + it, denom + it, f->pp, f->Ap,
                           f->x, f->r, p->red, f->flags);
        if ((rc = canon_finish(p, nt, rn2 + it, f->flags))) return rc;
        hipLaunchKernelGGL(k_filter_check, dim3(1), dim3(1), 0, st, rn2 + it, it, f->flags);
        hipLaunchKernelGGL(k_filter_z, dim3(nt), dim3(CANON_LANES), 0, st, n, f->r, f->diag, f->z, p->red, f->flags);
        if ((rc = canon_finish(p, nt, rz + it + 1, f->flags))) return rc;
        hipLaunchKernelGGL(k_filter_p, dim3(nb), dim3(B), 0, st, n, rz + it + 1, rz + it, f->z, f->pp, f->flags);
        T3D_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_filter_out, dim3(unsigned((p->ne + B - 1) / B)), dim3(B), 0, st, g, p->ne,
                       p->compactToCell, f->x, dst);
    T3D_HIP(hipGetLastError());
    // warm-start state for the next call
    std::swap(f->lastX, f->x);
    std::swap(f->lastRhs, f->rhs);
    T3D_HIP(hipMemcpyAsync(f->hostFlags, f->flags, sizeof(int) * 3, hipMemcpyDeviceToHost, st));
    return T3D_OK;
}

int t3d_filter_last_iters(const t3d_filter* f) { return f->hostFlags[1]; }

int t3d_filter_apply(t3d_filter* f, const float* src, float* dst, int* iters)
{
    T3D_CHECK(f && src && dst, T3D_ERR_ARG, "t3d_filter_apply: null argument");
    int rc = t3d_filter_apply_async(f, src, dst);
    if (rc) return rc;
    T3D_HIP(hipStreamSynchronize(f->p->stream));
    // -1 = early exit before the CG loop (filter.cpp:228-238)
    if (iters) *iters = f->hostFlags[1];
    return T3D_OK;
}

}  // extern "C"
