// kxu.hip -- matrix-free SIMP elasticity operator y = K_ff u on the node grid.
//
// Reference: K_times_u_finest + ProcessBlock_8_Free (cpp/src/fea.cpp:692-837).
// The reference loops over 8 parity colours and, per element, forms
// fe[i] = sum_{j=0..23} Ke[i][j]*ue[j] (sequential j) and scatters fe*E.  Each
// free DOF therefore receives its <= 8 element contributions in colour order
// 0..7, starting from +0.0.  The node-centric kernels below gather instead of
// scatter: one thread owns a node, visits its incident cells in colour order
// and evaluates the three rows of Ke belonging to the node's local index, in
// the same operation order -- the result is bit-identical, with no atomics.
#include <hip/hip_runtime.h>

#include "t3d_internal.hpp"

namespace t3d {

// local index a of a node sitting at offset (dx,dy,dz) inside its cell;
// inverse of the table (fea.cpp:257-264):
// a: 0(0,1,0) 1(1,1,0) 2(1,0,0) 3(0,0,0) 4(0,1,1) 5(1,1,1) 6(1,0,1) 7(0,0,1)
__device__ __forceinline__ int local_index(int dx, int dy, int dz)
{
    const int a2 = dy ? (dx ? 1 : 0) : (dx ? 2 : 3);
    return a2 + 4 * dz;
}
__device__ __forceinline__ void local_offset(int a, int& dx, int& dy, int& dz)
{
    const int b = a & 3;
    dx = (b == 1 || b == 2) ? 1 : 0;
    dy = (b <= 1) ? 1 : 0;
    dz = a >> 2;
}

// ---------------------------------------------------------------------------
// v1: straightforward node-centric FP64 operator (gathers through the caches)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_kxu_f64_node(Grid g, const double* __restrict__ Ke,
                                                      const double* __restrict__ Ecell,
                                                      const uint8_t* __restrict__ nodeFixed,
                                                      const double* __restrict__ u, double* __restrict__ y)
{
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= g.nN) return;
    const int ix = int(n % g.nnx);
    const int iy = int((n / g.nnx) % g.nny);
    const int iz = int(n / (int64_t(g.nnx) * g.nny));
    const unsigned fixed = nodeFixed[n];
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (fixed != 7u) {
        for (int col = 0; col < 8; ++col) {
            // the incident cell of colour col: colour bits are (x&1, ey&1, z&1)
            // with ey = resY-1-ty and resY even, so ey&1 = 1 - (ty&1)
            const int dx = (ix ^ col) & 1;
            const int dy = (iy ^ (col >> 1) ^ 1) & 1;
            const int dz = (iz ^ (col >> 2)) & 1;
            const int cx = ix - dx, ty = iy - dy, cz = iz - dz;
            if (cx < 0 || ty < 0 || cz < 0 || cx >= g.resX || ty >= g.resY || cz >= g.resZ) continue;
            const int64_t cell = (int64_t(cz) * g.resY + ty) * g.resX + cx;
            const double E = Ecell[cell];
            if (E <= 1.0e-16) continue;   // void cell or skipped element (fea.cpp:743-744)
            double ue[24];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                int ox, oy, oz;
                local_offset(b, ox, oy, oz);
                const int64_t nb = (int64_t(cz + oz) * g.nny + (ty + oy)) * g.nnx + (cx + ox);
                ue[3 * b + 0] = u[3 * nb + 0];
                ue[3 * b + 1] = u[3 * nb + 1];
                ue[3 * b + 2] = u[3 * nb + 2];
            }
            const int a = local_index(dx, dy, dz);
            const double* K0 = Ke + (3 * a) * 24;
            double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
            for (int j = 0; j < 24; ++j) {
                f0 = f0 + K0[j] * ue[j];
                f1 = f1 + K0[24 + j] * ue[j];
                f2 = f2 + K0[48 + j] * ue[j];
            }
            acc0 = acc0 + f0 * E;
            acc1 = acc1 + f1 * E;
            acc2 = acc2 + f2 * E;
        }
    }
    y[3 * n + 0] = (fixed & 1u) ? 0.0 : acc0;
    y[3 * n + 1] = (fixed & 2u) ? 0.0 : acc1;
    y[3 * n + 2] = (fixed & 4u) ? 0.0 : acc2;
}

int launch_kxu_f64(t3d_problem* p, const double* u, double* y)
{
    const int B = 256;
    hipLaunchKernelGGL(k_kxu_f64_node, dim3(unsigned((p->g.nN + B - 1) / B)), dim3(B), 0, p->stream, p->g, p->Ke,
                       p->Ecell, p->nodeFixed, u, y);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

}  // namespace t3d

using namespace t3d;

extern "C" {

int t3d_kxu_f64(t3d_problem* p, const double* u, double* y)
{
    T3D_CHECK(p && u && y, T3D_ERR_ARG, "t3d_kxu_f64: null argument");
    int rc = launch_kxu_f64(p, u, y);
    if (rc) return rc;
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_kxu_free_host(t3d_problem* p, const double* E_h, const double* uF_h, double* yF_h)
{
    T3D_CHECK(p && E_h && uF_h && yF_h, T3D_ERR_ARG, "t3d_kxu_free_host: null argument");
    double *E = nullptr, *uf = nullptr, *u = nullptr, *y = nullptr;
    const size_t nD = size_t(3 * p->g.nN);
    int rc = T3D_OK;
    auto cleanup = [&]() {
        if (E) (void)hipFree(E);
        if (uf) (void)hipFree(uf);
        if (u) (void)hipFree(u);
        if (y) (void)hipFree(y);
    };
    if (hipMalloc(&E, sizeof(double) * p->ne) != hipSuccess || hipMalloc(&uf, sizeof(double) * p->nF) != hipSuccess ||
        hipMalloc(&u, sizeof(double) * nD) != hipSuccess || hipMalloc(&y, sizeof(double) * nD) != hipSuccess) {
        cleanup();
        set_error("t3d_kxu_free_host: allocation failed");
        return T3D_ERR_HIP;
    }
    if (hipMemcpy(E, E_h, sizeof(double) * p->ne, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(uf, uF_h, sizeof(double) * p->nF, hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        set_error("t3d_kxu_free_host: upload failed");
        return T3D_ERR_HIP;
    }
    if ((rc = t3d_set_modulus(p, E)) == T3D_OK && (rc = t3d_free_to_full(p, uf, u)) == T3D_OK &&
        (rc = t3d_kxu_f64(p, u, y)) == T3D_OK && (rc = t3d_full_to_free(p, y, uf)) == T3D_OK) {
        if (hipMemcpy(yF_h, uf, sizeof(double) * p->nF, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("t3d_kxu_free_host: download failed");
            rc = T3D_ERR_HIP;
        }
    }
    cleanup();
    return rc;
}

int t3d_time_op(t3d_problem* p, int op, const void* in, void* out, int reps, double* avg_ms)
{
    T3D_CHECK(p && in && out && reps > 0 && avg_ms, T3D_ERR_ARG, "t3d_time_op: bad argument");
    hipEvent_t e0, e1;
    T3D_HIP(hipEventCreate(&e0));
    T3D_HIP(hipEventCreate(&e1));
    int rc = T3D_OK;
    T3D_HIP(hipEventRecord(e0, p->stream));
    for (int r = 0; r < reps && rc == T3D_OK; ++r) {
        switch (op) {
            case 0: rc = launch_kxu_f64(p, (const double*)in, (double*)out); break;
            default: set_error("t3d_time_op: unknown op"); rc = T3D_ERR_ARG;
        }
    }
    T3D_HIP(hipEventRecord(e1, p->stream));
    T3D_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    T3D_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = double(ms) / reps;
    return rc;
}

}  // extern "C"
