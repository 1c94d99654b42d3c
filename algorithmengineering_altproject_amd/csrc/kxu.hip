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

// ---------------------------------------------------------------------------
// v2: LDS-tiled node-centric FP64 operator with wave-uniform sub-lattices.
//
// A workgroup owns a 16x8x4 node tile.  Wave w (of 8) owns the 64 nodes whose
// coordinate parities are (w&1, (w>>1)&1, (w>>2)&1).  For a fixed colour the
// incident cell offset -- and hence the node's local index a and the three Ke
// rows it uses -- depend only on the parities, so they are uniform across the
// wave: Ke is read through the scalar cache into SGPR operands while u and E
// come from LDS.  Summation order per DOF is colour 0..7 as in the reference.
// ---------------------------------------------------------------------------
// Each thread owns NZ nodes of its sub-lattice stacked 4 apart in z, so one
// scalar Ke load feeds NZ multiply-adds.
template <int NZ>
struct KxuTile {
    static constexpr int TX = 16, TY = 8, TZ = 4 * NZ;
    static constexpr int UX = TX + 2, UY = TY + 2, UZ = TZ + 2;   // node tile + halo
    static constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1;   // cell tile + halo
    static constexpr int UN = UX * UY * UZ;
    static constexpr int EN = EX * EY * EZ;
};

// The PCG's direction update p = z + beta p_prev (solver.cpp:203-210) done by
// the operator that consumes p: every tile forms p for its nodes and halo
// from z and p_prev while staging them, and stores it for the nodes it owns
// (p_prev and p are different buffers, so a neighbour's halo reads the old p).
// With x set, the same tile also applies the previous iteration's solution
// update x += alpha p_prev to the nodes it owns (solver.cpp:184-188), alpha =
// rzOld / max(1e-30, denomOld) as the iteration computed it: the update the
// x/r kernel no longer makes, with the same operations on the same values.
struct FuseP {
    const double* z;
    const double* pPrev;
    const double* rzNew;   // beta = rzNew / max(1e-30, rzOld)
    const double* rzOld;
    double* p;
    double* x = nullptr;
    const double* denomOld = nullptr;
};

template <int NZ, bool FUSEP = false>
__global__ void __launch_bounds__(512) k_kxu_f64_lds(Grid g, const double* __restrict__ Ke,
                                                     const double* __restrict__ Ecell,
                                                     const uint8_t* __restrict__ nodeFixed,
                                                     const double* __restrict__ u, double* __restrict__ y,
                                                     const int* __restrict__ done, FuseP fp = FuseP{})
{
    if (done && *done) return;
    using T = KxuTile<NZ>;
    __shared__ double sU[3 * T::UN];   // SoA: component c at sU[c*UN + idx]
    __shared__ double sE[T::EN];
    const TileId tile = xcd_tile();
    const int x0 = tile.x * T::TX, y0 = tile.y * T::TY, z0 = tile.z * T::TZ;
    const int tid = threadIdx.x;
    double beta = 0.0, alphaPrev = 0.0;
    if constexpr (FUSEP) {
        const double o = *fp.rzOld;
        beta = *fp.rzNew / (1.0e-30 < o ? o : 1.0e-30);
        if (fp.x) {
            const double d = *fp.denomOld;
            alphaPrev = o / (1.0e-30 < d ? d : 1.0e-30);
        }
    }
    // cooperative loads (zero outside the grid)
    for (int i = tid; i < T::UN; i += 512) {
        const int lx = i % T::UX, ly = (i / T::UX) % T::UY, lz = i / (T::UX * T::UY);
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < g.nnx && Y < g.nny && Z < g.nnz) {
            const int64_t n = (int64_t(Z) * g.nny + Y) * g.nnx + X;
            if constexpr (FUSEP) {
                const double q0 = fp.pPrev[3 * n], q1 = fp.pPrev[3 * n + 1], q2 = fp.pPrev[3 * n + 2];
                v0 = fp.z[3 * n] + beta * q0;
                v1 = fp.z[3 * n + 1] + beta * q1;
                v2 = fp.z[3 * n + 2] + beta * q2;
                if (lx >= 1 && lx <= T::TX && ly >= 1 && ly <= T::TY && lz >= 1 && lz <= T::TZ) {
                    fp.p[3 * n] = v0; fp.p[3 * n + 1] = v1; fp.p[3 * n + 2] = v2;
                    if (fp.x) {
                        double* xn = fp.x + 3 * n;
                        const double x0v = xn[0], x1v = xn[1], x2v = xn[2];
                        xn[0] = x0v + alphaPrev * q0;
                        xn[1] = x1v + alphaPrev * q1;
                        xn[2] = x2v + alphaPrev * q2;
                    }
                }
            } else {
                v0 = u[3 * n]; v1 = u[3 * n + 1]; v2 = u[3 * n + 2];
            }
        }
        sU[i] = v0; sU[T::UN + i] = v1; sU[2 * T::UN + i] = v2;
    }
    for (int i = tid; i < T::EN; i += 512) {
        const int lx = i % T::EX, ly = (i / T::EX) % T::EY, lz = i / (T::EX * T::EY);
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        double e = 0.0;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < g.resX && Y < g.resY && Z < g.resZ)
            e = Ecell[(int64_t(Z) * g.resY + Y) * g.resX + X];
        sE[i] = e;
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int px = w & 1, py = (w >> 1) & 1, pz = (w >> 2) & 1;
    const int lx = 2 * (lane & 7) + px, ly = 2 * ((lane >> 3) & 3) + py, lz = 2 * (lane >> 5) + pz;
    double acc[NZ][3];
#pragma unroll
    for (int k = 0; k < NZ; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
#pragma unroll 1
    for (int col = 0; col < 8; ++col) {
        const int dx = (px ^ col) & 1;
        const int dy = (py ^ (col >> 1) ^ 1) & 1;
        const int dz = (pz ^ (col >> 2)) & 1;
        const int cxl = lx - dx, cyl = ly - dy, czl = lz - dz;   // incident cell, tile coords
        const int a = __builtin_amdgcn_readfirstlane(local_index(dx, dy, dz));
        const double* __restrict__ K0 = Ke + (3 * a) * 24;
        double E[NZ], f[NZ][3];
        int ub[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            E[k] = sE[((czl + 4 * k + 1) * T::EY + (cyl + 1)) * T::EX + (cxl + 1)];
            ub[k] = ((czl + 4 * k + 1) * T::UY + (cyl + 1)) * T::UX + (cxl + 1);
            f[k][0] = f[k][1] = f[k][2] = 0.0;
        }
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int ox = (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0;
            const int oy = ((b & 3) <= 1) ? 1 : 0;
            const int oz = b >> 2;
            const int off = (oz * T::UY + oy) * T::UX + ox;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int j = 3 * b + c;
                const double k0 = K0[j], k1 = K0[24 + j], k2 = K0[48 + j];
#pragma unroll
                for (int k = 0; k < NZ; ++k) {
                    const double ue = sU[c * T::UN + ub[k] + off];
                    f[k][0] = f[k][0] + k0 * ue;
                    f[k][1] = f[k][1] + k1 * ue;
                    f[k][2] = f[k][2] + k2 * ue;
                }
            }
        }
        // elements with E <= 1e-16 (and void cells, E = 0) contribute nothing (fea.cpp:743-744)
#pragma unroll
        for (int k = 0; k < NZ; ++k)
            if (!(E[k] <= 1.0e-16)) {
                acc[k][0] = acc[k][0] + f[k][0] * E[k];
                acc[k][1] = acc[k][1] + f[k][1] * E[k];
                acc[k][2] = acc[k][2] + f[k][2] * E[k];
            }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int X = x0 + lx, Y = y0 + ly, Z = z0 + lz + 4 * k;
        if (X < g.nnx && Y < g.nny && Z < g.nnz) {
            const int64_t n = (int64_t(Z) * g.nny + Y) * g.nnx + X;
            const unsigned fixed = nodeFixed[n];
            y[3 * n + 0] = (fixed & 1u) ? 0.0 : acc[k][0];
            y[3 * n + 1] = (fixed & 2u) ? 0.0 : acc[k][1];
            y[3 * n + 2] = (fixed & 4u) ? 0.0 : acc[k][2];
        }
    }
}

template <int NZ>
static int launch_kxu_lds(t3d_problem* p, const double* u, double* y, const int* done = nullptr)
{
    using T = KxuTile<NZ>;
    const Grid& g = p->g;
    dim3 grid(unsigned((g.nnx + T::TX - 1) / T::TX), unsigned((g.nny + T::TY - 1) / T::TY),
              unsigned((g.nnz + T::TZ - 1) / T::TZ));
    hipLaunchKernelGGL(k_kxu_f64_lds<NZ>, grid, dim3(512), 0, p->stream, g, p->Ke, p->Ecell, p->nodeFixed, u, y, done);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

__global__ void __launch_bounds__(256) k_stream_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b)
{
    const int64_t i = 2 * (int64_t(blockIdx.x) * blockDim.x + threadIdx.x);
    if (i + 1 < n) {
        const double2 v = *reinterpret_cast<const double2*>(a + i);
        *reinterpret_cast<double2*>(b + i) = v;
    } else if (i < n) {
        b[i] = a[i];
    }
}

int launch_kxu_f64(t3d_problem* p, const double* u, double* y, const int* done)
{
    return launch_kxu_lds<1>(p, u, y, done);
}

// p = z + beta p_prev, then y = K p (see FuseP)
int launch_kxu_f64_fusep(t3d_problem* p, const double* z, const double* pPrev, const double* rzNew,
                         const double* rzOld, double* pOut, double* y, const int* done, double* x,
                         const double* denomOld)
{
    using T = KxuTile<1>;
    const Grid& g = p->g;
    dim3 grid(unsigned((g.nnx + T::TX - 1) / T::TX), unsigned((g.nny + T::TY - 1) / T::TY),
              unsigned((g.nnz + T::TZ - 1) / T::TZ));
    const FuseP fp{z, pPrev, rzNew, rzOld, pOut, x, denomOld};
    hipLaunchKernelGGL((k_kxu_f64_lds<1, true>), grid, dim3(512), 0, p->stream, g, p->Ke, p->Ecell, p->nodeFixed,
                       (const double*)pOut, y, done, fp);
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
    (void)hipDeviceSynchronize();   // the uploads ran on the null stream
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
            case 101: rc = launch_kxu_lds<2>(p, (const double*)in, (double*)out); break;
            case 1: {   // streaming copy of 3*nN doubles: bandwidth / counter calibration
                const int64_t n = 3 * p->g.nN;
                hipLaunchKernelGGL(k_stream_copy, dim3(unsigned((n / 2 + 255) / 256)), dim3(256), 0, p->stream, n,
                                   (const double*)in, (double*)out);
                break;
            }
            case 100: {
                const int B = 256;
                hipLaunchKernelGGL(k_kxu_f64_node, dim3(unsigned((p->g.nN + B - 1) / B)), dim3(B), 0, p->stream,
                                   p->g, p->Ke, p->Ecell, p->nodeFixed, (const double*)in, (double*)out);
                break;
            }
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
