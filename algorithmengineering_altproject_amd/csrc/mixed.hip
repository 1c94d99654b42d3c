// mixed.hip -- mixed-precision MG-PCG: FP32 operator and V-cycle inside a
// flexible FP64 outer conjugate-gradient loop.
//
// Same preconditioner and operator as the FP64 plain path (PCG_free,
// cpp/src/solver.cpp:87-217; V-cycle, cpp/src/multigrid/preconditioner.cpp:88-265;
// K_times_u_finest, cpp/src/fea.cpp:692-837), with:
//   * K*u in FP32 (explicit FMA), p kept only in FP32; the update
//     p = z + beta*p and the dot p.Ap are fused into the operator kernel;
//   * the V-cycle (restriction, Jacobi smoothing, prolongation) in FP32 with
//     FP32 diagonals; the coarsest Cholesky solve stays FP64;
//   * the recursive residual r in FP32 and the x updates accumulated in an
//     FP32 increment dx that is added into the FP64 x at every residual
//     replacement and at the end;
//   * FP64 dot products, alpha and beta; beta either the reference's
//     (r.z)_new/(r.z)_old or the flexible (Polak-Ribiere) form
//     -alpha (z_new.Ap)/(r.z)_old, which tolerates the FP32 preconditioner
//     being only approximately linear;
//   * every `trueEvery` iterations the residual is recomputed as b - K x with
//     the FP64 operator (residual replacement; the direction is kept), and a
//     solve only ends when that true residual passes the tolerance.
//
// Reductions: every block of the producing kernel writes one FP64 partial
// (a fixed tree over its threads); a one-block finishing kernel sums the
// partials in block order and derives the scalars of the iteration (alpha, the
// convergence flag, beta), so every result is deterministic.  The kernel
// boundary orders the partials; no device-wide fence per block is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "t3d_internal.hpp"

#include "coarse.hpp"
#include "mx_kernels.hpp"

namespace t3d {

template <int NZ, bool FUSED>
__global__ void __launch_bounds__(256) k_kx32(Grid g, const float* __restrict__ Ke, const float* __restrict__ Ecell,
                                              const uint8_t* __restrict__ nodeFixed, const float* __restrict__ u,
                                              const float* __restrict__ z, float* __restrict__ pcur,
                                              float* __restrict__ y, double* __restrict__ partial,
                                              const double* __restrict__ S, const int* __restrict__ done)
{
    using T = kx::Tile<NZ>;
    using kx::TX;
    using kx::TY;
    if (FUSED && *done) return;
    __shared__ float sU[3 * T::UN];
    __shared__ float sE[T::EN];
    __shared__ double sh[4];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * NZ;
    const int tid = threadIdx.x;
    const float beta = FUSED ? float(S[S_BETA]) : 0.0f;
    // stage the node halo: rows of UX nodes are contiguous 3*UX floats in memory
    constexpr int ROW = 3 * T::UX;
    for (int f = tid; f < ROW * T::UY * T::UZ; f += 256) {
        const int row = f / ROW;
        const int k = f - row * ROW;
        const int lx = k / 3, c = k - 3 * lx;
        const int ly = row % T::UY, lz = row / T::UY;
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        float v = 0.f;
        if (X >= 0 && X < g.nnx && Y >= 0 && Y < g.nny && Z >= 0 && Z < g.nnz) {
            const int64_t gi = 3 * ((int64_t(Z) * g.nny + Y) * g.nnx + X) + c;
            v = u[gi];
            if (FUSED) {
                v = fmaf(beta, v, z[gi]);
                if (lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && lz >= 1 && lz <= NZ) pcur[gi] = v;
            }
        }
        sU[c * T::UN + row * T::UX + lx] = v;
    }
    // cell halo; cells with E <= 1e-16 (and void cells) contribute nothing (fea.cpp:743-744)
    for (int i = tid; i < T::EN; i += 256) {
        const int lx = i % T::EX, ly = (i / T::EX) % T::EY, lz = i / (T::EX * T::EY);
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        float e = 0.f;
        if (X >= 0 && X < g.resX && Y >= 0 && Y < g.resY && Z >= 0 && Z < g.resZ) {
            e = Ecell[(int64_t(Z) * g.resY + Y) * g.resX + X];
            if (e <= 1.0e-16f) e = 0.f;
        }
        sE[i] = e;
    }
    __syncthreads();
    const int tx = tid % TX, ty = tid / TX;
    const int X = x0 + tx, Y = y0 + ty;
    const int nz = min(NZ, g.nnz - z0);
    double dot = 0.0;
    if (X < g.nnx && Y < g.nny) {
        float w[3][3][3][3];   // [z][y][x][component] around the current node
        float e[2][2][2];      // [dz][dy][dx]: cell (X-dx, Y-dy, Z-dz)
        auto plane = [&](int L, float (&P)[3][3][3]) {
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix)
#pragma unroll
                    for (int c = 0; c < 3; ++c) P[iy][ix][c] = sU[c * T::UN + (L * T::UY + ty + iy) * T::UX + tx + ix];
        };
        auto eplane = [&](int L, float (&P)[2][2]) {
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) P[dy][dx] = sE[(L * T::EY + ty + 1 - dy) * T::EX + tx + 1 - dx];
        };
        plane(0, w[1]);
        plane(1, w[2]);
        eplane(0, e[0]);
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
            if (j >= nz) break;
#pragma unroll
            for (int a = 0; a < 27; ++a) {
                (&w[0][0][0][0])[a] = (&w[1][0][0][0])[a];
                (&w[1][0][0][0])[a] = (&w[2][0][0][0])[a];
            }
            plane(j + 2, w[2]);
#pragma unroll
            for (int a = 0; a < 4; ++a) (&e[1][0][0])[a] = (&e[0][0][0])[a];
            eplane(j + 1, e[0]);
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int a = corner_of(dx, dy, dz);
                        float f0 = 0.f, f1 = 0.f, f2 = 0.f;
#pragma unroll
                        for (int b = 0; b < 8; ++b) {
                            const int ox = (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0;
                            const int oy = ((b & 3) <= 1) ? 1 : 0;
                            const int oz = b >> 2;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const float uv = w[oz - dz + 1][oy - dy + 1][ox - dx + 1][c];
                                const int jj = 3 * b + c;
                                f0 = fmaf(Ke[(3 * a) * 24 + jj], uv, f0);
                                f1 = fmaf(Ke[(3 * a + 1) * 24 + jj], uv, f1);
                                f2 = fmaf(Ke[(3 * a + 2) * 24 + jj], uv, f2);
                            }
                        }
                        const float E = e[dz][dy][dx];
                        a0 = fmaf(E, f0, a0);
                        a1 = fmaf(E, f1, a1);
                        a2 = fmaf(E, f2, a2);
                    }
            const int64_t n = (int64_t(z0 + j) * g.nny + Y) * g.nnx + X;
            const unsigned fx = nodeFixed[n];
            const float y0v = (fx & 1u) ? 0.f : a0, y1v = (fx & 2u) ? 0.f : a1, y2v = (fx & 4u) ? 0.f : a2;
            y[3 * n] = y0v;
            y[3 * n + 1] = y1v;
            y[3 * n + 2] = y2v;
            if (FUSED) {
                dot = fma(double(w[1][1][1][0]), double(y0v), dot);
                dot = fma(double(w[1][1][1][1]), double(y1v), dot);
                dot = fma(double(w[1][1][1][2]), double(y2v), dot);
            }
        }
    }
    if (FUSED) {
        const double t = block_sum_256(dot, sh);
        if (tid == 0) partial[linear_block()] = t;
    }
}
// x += dx + alpha p (the increment of the last iteration, not yet applied by
// an operator launch) ; dx = 0 ; followed by k_mx_applied
__global__ void __launch_bounds__(256) k_mx_fold(int64_t n, double* __restrict__ x, float* __restrict__ dx,
                                                 const float* __restrict__ pA, const float* __restrict__ pB,
                                                 const double* __restrict__ S, const int* __restrict__ done)
{
    if (done && *done) return;
    const float a = float(S[S_ALPHA_DX]);
    const float* p = S[S_PLAST] != 0.0 ? pA : pB;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        x[i] += double(fmaf(a, p[i], dx[i]));
        dx[i] = 0.f;
    }
}
// the pending increment is in x: the next operator launch adds nothing
__global__ void k_mx_applied(double* __restrict__ S, const int* __restrict__ done)
{
    if (done && *done) return;
    S[S_ALPHA_DX] = 0.0;
}
__global__ void k_mx_init(double tol, double* __restrict__ S, int* __restrict__ flags)
{
    const double nb = sqrt(S[S_NORMB2]);
    S[S_STOP] = (1.0e-30 < nb ? nb : 1.0e-30) * tol;
    S[S_BETA] = 0.0;
    S[S_ALPHA_DX] = 0.0;
    S[S_PLAST] = 0.0;
    flags[0] = 0;
    flags[1] = -1;
}
__global__ void k_mx_clear_done(int* __restrict__ flags)
{
    flags[0] = 0;
    flags[1] = -1;
}
// FP32 inverse Jacobi diagonal 1 / max(1e-30, D) (the smoother's and the
// coarsest diagonal solve's divisor, preconditioner.cpp:230-253)
__global__ void k_dinv(int64_t n, const double* __restrict__ D, float* __restrict__ Dinv)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) Dinv[i] = float(1.0 / (1.0e-30 < D[i] ? D[i] : 1.0e-30));
}

// xcd != 0: XCD-grouped tile order (each XCD works through a contiguous range
// of tiles, x fastest, so x- and y-neighbours that share halo nodes run on
// one L2); the partial slot is the tile index in both orders
// PIECES = 0: the Walsh form (KW); 2, 3: the MFMA form with bf16 x3 / x6 products (KA)
template <int NZ, bool FUSED, int CYT, int PIECES = 0>
__global__ void __launch_bounds__(32 * CYT, CYT >= 16 ? 4 : 5) k_kxw32(Grid g, const float* __restrict__ KW, const float* __restrict__ Ecell,
                                                  const uint8_t* __restrict__ nodeFixed, const float* __restrict__ u,
                                                  const float* __restrict__ z, float* __restrict__ pcur,
                                                  float* __restrict__ y, double* __restrict__ partial,
                                                  const double* __restrict__ S, const int* __restrict__ done, int xcd,
                                                  float* __restrict__ dxv, int pid,
                                                  const unsigned short* __restrict__ KA = nullptr)
{
    unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (xcd) {
        const TileId t = xcd_tile();
        bx = t.x;
        by = t.y;
        bz = t.z;
    }
    const unsigned slot = (bz * gridDim.y + by) * gridDim.x + bx;
    kxw32_body<NZ, FUSED, CYT, PIECES>(g, KW, Ecell, nodeFixed, u, z, pcur, y, partial, S, done, dxv, pid, bx, by,
                                       bz, slot, gridDim.x, KA);
}

// ---------------------------------------------------------------------------
// The true residual of the residual replacement, r = b - K x, in FP64 with
// the Walsh form: the scheme of kxw32_body (a thread per voxel column
// sweeping z, node planes staged in LDS one step ahead, face transforms
// shared by the voxels below and above, the four-column node gather in a
// fixed order) on FP64 planes, FP64 moduli (p->Ecell, zero at or below 1e-16
// as fea.cpp:743-744) and the 45 FP64 coefficients of K^/64.  The node gather
// ends in r = b - (K x, 0 at fixed DOFs), stored in FP32, and the block's
// partial of the FP64 r.r.  256-thread tiles of 32 x 8 voxel columns (31 x 7
// owned node columns) and R64_NZ node planes.
// ---------------------------------------------------------------------------
constexpr int R64_NZ = 16;

__device__ __forceinline__ double buf_ld64(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void walsh2d(double a00, double a10, double a01, double a11, double F[2][2])
{
    const double s0 = a00 + a10, d0 = a00 - a10, s1 = a01 + a11, d1 = a01 - a11;
    F[0][0] = s0 + s1;
    F[0][1] = d0 + d1;
    F[1][0] = s0 - s1;
    F[1][1] = d0 - d1;
}

static dim3 resid64_grid(const Grid& g)
{
    return dim3(unsigned((g.nnx + wk::NX - 1) / wk::NX), unsigned((g.nny + wk::NY - 1) / wk::NY),
                unsigned((g.nnz + R64_NZ - 1) / R64_NZ));
}

__global__ void __launch_bounds__(256) k_resid64(Grid g, const double* __restrict__ KW, const double* __restrict__ Ecell,
                                                 const uint8_t* __restrict__ nodeFixed, const double* __restrict__ x,
                                                 const double* __restrict__ b, float* __restrict__ r,
                                                 double* __restrict__ partial, const int* __restrict__ done)
{
    using wk::IN;
    using wk::NNZ;
    using wk::OUT;
    constexpr int CX = wk::CX, CY = wk::CY, NX = wk::NX, NY = wk::NY, PX = wk::PX, PN = wk::PN;
    constexpr int NT = CX * CY;
    constexpr int SPT = (3 * PN + NT - 1) / NT;   // staging slots per thread
    constexpr int SPN = 3 * PN + 64;              // staged plane + 64 write-only dummy slots
    constexpr int NZ = R64_NZ;
    if (done && *done) return;
    __shared__ double sP[3][SPN];
    __shared__ double sC[12][NT];
    __shared__ double sh[NT / 64];
    const int tid = threadIdx.x;
    const int tx = tid % CX, ty = tid / CX;
    const int x0 = blockIdx.x * NX, y0 = blockIdx.y * NY, z0 = blockIdx.z * NZ;
    const int cx = x0 - 1 + tx, cy = y0 - 1 + ty;   // this thread's voxel column
    const bool colIn = cx >= 0 && cx < g.resX && cy >= 0 && cy < g.resY;
    const int X = x0 + tx, Y = y0 + ty;             // the node this thread gathers
    const bool ownNode = tx < NX && ty < NY && X < g.nnx && Y < g.nny;
    uint32_t voff[SPT];   // byte offset of staging slot j within a node plane (BUF_OOB: loads 0)
    int slds[SPT];        // its LDS index c * PN + row * PX + lx (a dummy slot if none)
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int f = tid + j * NT;
        const int row = f / (3 * PX);
        const int k = f - row * (3 * PX);
        const int lx = k / 3, c = k - 3 * lx;
        const int Xs = x0 - 1 + lx, Ys = y0 - 1 + row;
        const bool slot = f < 3 * PN;
        const bool in = slot && Xs >= 0 && Xs < g.nnx && Ys >= 0 && Ys < g.nny;
        slds[j] = slot ? c * PN + row * PX + lx : 3 * PN + (tid & 63);
        voff[j] = in ? uint32_t(8 * (3 * (Ys * g.nnx + Xs) + c)) : BUF_OOB;
    }
    const int64_t plane = int64_t(3) * g.nnx * g.nny;
    const int64_t layer = int64_t(g.resX) * g.resY;
    const uint32_t eoff = colIn ? uint32_t(8 * (cy * g.resX + cx)) : BUF_OOB;
    const uint32_t noff = ownNode ? uint32_t(Y * g.nnx + X) : BUF_OOB;
    double fu[SPT];
    auto fetch = [&](int Z) {
        const bool zin = Z >= 0 && Z < g.nnz;
        const __amdgpu_buffer_rsrc_t ru = buf_rsrc(x + (zin ? Z * plane : 0), zin ? uint32_t(8 * plane) : 0u);
#pragma unroll
        for (int j = 0; j < SPT; ++j) fu[j] = buf_ld64(ru, voff[j]);
    };
    auto commit = [&](int q) {
#pragma unroll
        for (int j = 0; j < SPT; ++j) sP[q][slds[j]] = fu[j];
    };
    auto modulus = [&](int cz) {
        const bool in = cz >= 0 && cz < g.resZ;
        return buf_ld64(buf_rsrc(Ecell + (in ? cz * layer : 0), in ? uint32_t(8 * layer) : 0u), eoff);
    };
    auto fixed_bits = [&](int Z) {
        const bool zin = Z >= 0 && Z < g.nnz;
        const int64_t pn = int64_t(g.nnx) * g.nny;
        return unsigned(__builtin_amdgcn_raw_buffer_load_b8(
            buf_rsrc(nodeFixed + (zin ? Z * pn : 0), zin ? uint32_t(pn) : 0u), noff, 0, 0));
    };
    auto face = [&](int q, double F[3][2][2]) {
        const int bb = ty * PX + tx;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            walsh2d(sP[q][c * PN + bb], sP[q][c * PN + bb + 1], sP[q][c * PN + bb + PX], sP[q][c * PN + bb + PX + 1],
                    F[c]);
    };
    double Fp[3][2][2], Fn[3][2][2], carry[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) (&carry[c][0][0])[k] = 0.0;
    fetch(z0 - 1);
    commit(0);
    fetch(z0);
    commit(1);
    double Enext = modulus(z0 - 1);
    unsigned fxNext = 0u;
    __syncthreads();
    face(0, Fp);
    double dot = 0.0;
    const int nsteps = min(NZ, g.nnz - z0) + 1;
    for (int s = 0; s < nsteps; ++s) {
        const int qn = (s + 1) % 3;
        const bool more = s + 1 < nsteps;
        const int cz = z0 - 1 + s;
        const double E = Enext <= 1.0e-16 ? 0.0 : Enext;
        const unsigned fx = fxNext;
        if (more) {   // loads of the next step, in flight during this one
            fetch(z0 + s + 1);
            Enext = modulus(cz + 1);
            fxNext = fixed_bits(cz + 1);
        }
        face(qn, Fn);
        double uh[24], vh[24];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int my = 0; my < 2; ++my)
#pragma unroll
                for (int mx = 0; mx < 2; ++mx) {
                    uh[(2 * my + mx) * 3 + c] = E * (Fp[c][my][mx] + Fn[c][my][mx]);
                    uh[(4 + 2 * my + mx) * 3 + c] = E * (Fp[c][my][mx] - Fn[c][my][mx]);
                }
#pragma unroll
        for (int i = 0; i < 24; ++i) vh[i] = 0.0;
#pragma unroll
        for (int k = 0; k < NNZ; ++k) vh[OUT[k]] = fma(KW[k], uh[IN[k]], vh[OUT[k]]);
        double A4[3][2][2];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int my = 0; my < 2; ++my)
#pragma unroll
                for (int mx = 0; mx < 2; ++mx) {
                    const double v0 = vh[(2 * my + mx) * 3 + c], v1 = vh[(4 + 2 * my + mx) * 3 + c];
                    A4[c][my][mx] = carry[c][my][mx] + (v0 + v1);
                    carry[c][my][mx] = v0 - v1;
                }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) (&Fp[c][0][0])[k] = (&Fn[c][0][0])[k];
        if (s > 0) {   // corner forces of plane cz (plane z0 - 1 belongs to the block below)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double C4[2][2];
                walsh2d(A4[c][0][0], A4[c][0][1], A4[c][1][0], A4[c][1][1], C4);
                sC[0 * 3 + c][tid] = C4[0][0];
                sC[1 * 3 + c][tid] = C4[0][1];
                sC[2 * 3 + c][tid] = C4[1][0];
                sC[3 * 3 + c][tid] = C4[1][1];
            }
            __syncthreads();
            if (ownNode) {   // node (X, Y) collects corner (ox, oy) of voxel column (X - ox, Y - oy)
                const bool zin = cz < g.nnz;
                const __amdgpu_buffer_rsrc_t rb =
                    buf_rsrc(b + (zin ? cz * plane : 0), zin ? uint32_t(8 * plane) : 0u);
                const __amdgpu_buffer_rsrc_t rr =
                    buf_rsrc(r + (zin ? cz * plane : 0), zin ? uint32_t(4 * plane) : 0u);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double c00 = sC[3 * 3 + c][ty * CX + tx];
                    const double c10 = sC[2 * 3 + c][ty * CX + tx + 1];
                    const double c01 = sC[1 * 3 + c][(ty + 1) * CX + tx];
                    const double c11 = sC[0 * 3 + c][(ty + 1) * CX + tx + 1];
                    const double yv = ((fx >> c) & 1u) ? 0.0 : (c00 + c10) + (c01 + c11);
                    const double rv = buf_ld64(rb, 8 * (3 * noff + c)) - yv;
                    buf_st(float(rv), rr, 4 * (3 * noff + c));
                    dot = fma(rv, rv, dot);
                }
            }
        }
        if (more) commit((s + 2) % 3);
        __syncthreads();
    }
    const double t = block_sum_256(dot, sh);
    if (tid == 0) partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

__global__ void __launch_bounds__(1024) k_fin_alpha(const double* __restrict__ partial, unsigned m,
                                                    double* __restrict__ S, const int* __restrict__ done)
{
    fin_alpha_body(partial, m, S, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

__global__ void __launch_bounds__(1024) k_fin_rn2(const double* __restrict__ partial, unsigned m,
                                                  double* __restrict__ S, int* __restrict__ flags, int it, int check,
                                                  int guard)
{
    fin_rn2_body(partial, m, S, flags, it, check, guard, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

__global__ void __launch_bounds__(1024) k_fin_beta(const double* __restrict__ partial, unsigned m,
                                                   double* __restrict__ S, int flexible, int first,
                                                   const int* __restrict__ done)
{
    fin_beta_body(partial, m, S, flexible, first, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

// restriction of an FP32 level in one pass (span 2 or 4; restrict_c32_t)
__global__ void __launch_bounds__(64) k_restrict_c32(LevelF Lf, LevelF Lc, int span, const float* __restrict__ rf,
                                                     const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                                     const float* __restrict__ Dc, float* __restrict__ xc,
                                                     double* __restrict__ rcd, const int* __restrict__ done, int norm)
{
    if (done && *done) return;
    if (span == 4)
        restrict_c32_t<4>(Lf, Lc, rf, cmask, rc, Dc, xc, rcd, norm, blockIdx.x, blockIdx.y, blockIdx.z);
    else
        restrict_c32_t<2>(Lf, Lc, rf, cmask, rc, Dc, xc, rcd, norm, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class TIn>
__global__ void __launch_bounds__(64) k_restrict_x(LevelF Lf, int cnx, int span, const TIn* __restrict__ rf,
                                                   float* __restrict__ t, const int* __restrict__ done)
{
    restrict_x_body<TIn>(Lf, cnx, span, rf, t, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

__global__ void __launch_bounds__(64) k_restrict_yz(LevelF Lf, LevelF Lc, int span, const float* __restrict__ t,
                                                    const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                                    const float* __restrict__ Dc, float* __restrict__ xc,
                                                    double* __restrict__ rcd, const int* __restrict__ done,
                                                    int norm)
{
    restrict_yz_body(Lf, Lc, span, t, cmask, rc, Dc, xc, rcd, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(),
                     gridDim.x, norm);
}

__global__ void __launch_bounds__(64) k_prolong_smooth32(LevelF Lf, LevelF Lc, int span, const float* __restrict__ xc,
                                                         const float* __restrict__ rf, const float* __restrict__ Df,
                                                         const uint8_t* __restrict__ fmask, float weight,
                                                         float* __restrict__ xf, const int* __restrict__ done)
{
    prolong_smooth32_body(Lf, Lc, span, xc, rf, Df, fmask, weight, xf, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

__global__ void __launch_bounds__(256) k_xr_rows(int nnx, int NR, const float* __restrict__ Ap,
                                                 float* __restrict__ r, double* __restrict__ partial,
                                                 const double* __restrict__ S, int cnx, int span,
                                                 float* __restrict__ t, const int* __restrict__ done)
{
    xr_rows_body<RPB>(nnx, NR, Ap, r, partial, S, cnx, span, t, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

__global__ void __launch_bounds__(256) k_rx_rows(int nnx, int NR, const float* __restrict__ r, int cnx, int span,
                                                 float* __restrict__ t, const int* __restrict__ done)
{
    rx_rows_body<RPB>(nnx, NR, r, cnx, span, t, done, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

template <int U, int TW = 2, bool PARTIALS = true>
__global__ void __launch_bounds__(256) k_pro_rows(LevelF Lf, LevelF Lc, int span, const float* __restrict__ xc,
                                                  const float* __restrict__ rf, const float* __restrict__ Df,
                                                  const uint8_t* __restrict__ fmask, float weight,
                                                  float* __restrict__ xf, const float* __restrict__ Ap,
                                                  double* __restrict__ partial, const int* __restrict__ done)
{
    pro_rows_body<RPB, U, TW, PARTIALS>(Lf, Lc, span, xc, rf, Df, fmask, weight, xf, Ap, partial, done, blockIdx.x,
                                        blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

// ---------------------------------------------------------------------------
// MATLAB's 'Standard' V-cycle of the Galerkin form (TOP3D_XL.m:783-803) in
// FP32, as the FP64 path's (mg.hip) with the mixed solver's arithmetic:
//   down:  x_l = w r_l Dinv_l;  r_{l+1} = P^T (r_l - A_l x_l)
//   up:    x_l += P x_{l+1};    x_l += w (r_l - A_l x_l) Dinv_l
// every result masked; A_0 is the FP32 operator of the solver (fixed DOFs 0),
// A_l (l >= 1) the level's Galerkin element matrices in FP32.
// ---------------------------------------------------------------------------
// op 0: x = w r Dinv;  op 1: x = r - y;  op 2: x += w (r - y) Dinv  (masked -> 0); one node per thread
__global__ void __launch_bounds__(256) k_std32_point(int op, int64_t nN, const float* __restrict__ r,
                                                     const float* y, const float* __restrict__ Dinv,
                                                     const uint8_t* __restrict__ mask, float w, float* x,
                                                     const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= nN) return;
    const unsigned m = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t i = 3 * n + c;
        float out = 0.f;
        if (!((m >> c) & 1u)) {
            if (op == 0)
                out = w * r[i] * Dinv[i];
            else if (op == 1)
                out = r[i] - y[i];
            else
                out = fmaf(w * (r[i] - y[i]), Dinv[i], x[i]);
        }
        x[i] = out;
    }
}

// FP32 copy of a level's element matrices, entry-major: out[e * ncell + cell] =
// in[cell * 576 + e] (row-major 24 x 24 per cell, local DOF 3a + c), so that
// neighbouring nodes read neighbouring cells' entries
__global__ void __launch_bounds__(256) k_kel32(int64_t ncell, const double* __restrict__ in, float* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 576 * ncell) return;
    const int64_t e = i / ncell, c = i - e * ncell;
    out[i] = float(in[c * 576 + e]);
}

// y = A_l x from the level's FP32 element matrices (entry-major, k_kel32): one
// thread per node gathers the rows of its corner in each of the (up to) eight
// cells around it (k_gal_apply's order)
__global__ void __launch_bounds__(256) k_gal_apply32(LevelF L, const float* __restrict__ Kel,
                                                     const float* __restrict__ x, float* __restrict__ y,
                                                     const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= L.nN) return;
    const int ix = int(n % L.nnx), iy = int((n / L.nnx) % L.nny), iz = int(n / (int64_t(L.nnx) * L.nny));
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = ix - dx, cy = iy - dy, cz = iz - dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= L.resX || cy >= L.resY || cz >= L.resZ) continue;
                const int a = 4 * dz + (dy ? (dx ? 1 : 0) : (dx ? 2 : 3));   // the node's corner in the cell
                const int64_t nc = int64_t(L.resX) * L.resY * L.resZ;
                const float* K = Kel + int64_t(3 * a * 24) * nc + (int64_t(cz) * L.resY + cy) * L.resX + cx;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const int ox = node_off_x(b), oy = node_off_y(b), oz = b >> 2;
                    const int64_t nb = (int64_t(cz + oz) * L.nny + (cy + oy)) * L.nnx + (cx + ox);
                    const float u0 = x[3 * nb], u1 = x[3 * nb + 1], u2 = x[3 * nb + 2];
                    const float* Kb = K + int64_t(3 * b) * nc;
                    acc0 = fmaf(Kb[0], u0, fmaf(Kb[nc], u1, fmaf(Kb[2 * nc], u2, acc0)));
                    acc1 = fmaf(Kb[24 * nc], u0, fmaf(Kb[25 * nc], u1, fmaf(Kb[26 * nc], u2, acc1)));
                    acc2 = fmaf(Kb[48 * nc], u0, fmaf(Kb[49 * nc], u1, fmaf(Kb[50 * nc], u2, acc2)));
                }
            }
    y[3 * n] = acc0;
    y[3 * n + 1] = acc1;
    y[3 * n + 2] = acc2;
}

// xf += P xc (trilinear, masked -> 0), one fine node per thread
__global__ void __launch_bounds__(256) k_prolong_add32(LevelF Lf, LevelF Lc, int span, const float* __restrict__ xc,
                                                       const uint8_t* __restrict__ fmask, float* __restrict__ xf,
                                                       const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lf.nN) return;
    const int fx = int(n % Lf.nnx), fy = int((n / Lf.nnx) % Lf.nny), fz = int(n / (int64_t(Lf.nnx) * Lf.nny));
    const Prolong P(Lc, span, fx, fy, fz);
    float v[3];
    P.eval(Lc, xc, v);
    const unsigned m = fmask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) xf[3 * n + c] = ((m >> c) & 1u) ? 0.f : xf[3 * n + c] + v[c];
}

// the finest post-smoothing z += w (r - y) Dinv (masked -> 0) with the
// partials of r.z and z.Ap that give beta (fixed grid, block b writes
// partial[b] and partial[gridDim.x + b])
__global__ void __launch_bounds__(256) k_std32_final(int64_t nN, const float* __restrict__ r,
                                                     const float* __restrict__ y, const float* __restrict__ Dinv,
                                                     const uint8_t* __restrict__ mask, float w, float* __restrict__ z,
                                                     const float* __restrict__ Ap, double* __restrict__ partial,
                                                     const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[4];
    double srz = 0.0, szap = 0.0;
    for (int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; n < nN; n += int64_t(gridDim.x) * blockDim.x) {
        const unsigned m = mask[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t i = 3 * n + c;
            const float ri = r[i];
            const float out = ((m >> c) & 1u) ? 0.f : fmaf(w * (ri - y[i]), Dinv[i], z[i]);
            z[i] = out;
            srz = fma(double(ri), double(out), srz);
            szap = fma(double(out), double(Ap[i]), szap);
        }
    }
    const double a = block_sum_256(srz, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = a;
    const double b = block_sum_256(szap, sh);
    if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = b;
}

__global__ void k_d2f(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    d2f_body(n, a, b, blockIdx.x, blockIdx.y, blockIdx.z, linear_block(), gridDim.x);
}

// Ainv = Li^T Li into the packed lower tiles (coarse.hpp), one tile per workgroup
__global__ void __launch_bounds__(256) k_gram(int N, const double* __restrict__ Li, double* __restrict__ Ap)
{
    int I, J;
    coarse_tile_ij(blockIdx.x, I, J);
    gram_body(N, Li, Ap, I, J);
}

// the coarsest solve's product over the packed tiles (coarse.hpp): partials, then their sums
__global__ void __launch_bounds__(256) k_coarse_tiles(int N, const double* __restrict__ Ap,
                                                      const double* __restrict__ v, double* __restrict__ P,
                                                      const int* __restrict__ done)
{
    if (done && *done) return;
    coarse_tiles_loop(coarse_tiles(N), [=](unsigned e) { return CoarseTile{N, Ap, v, P, e, true}; });
}

__global__ void __launch_bounds__(CT) k_coarse_sum(int N, const double* __restrict__ P,
                                                   const uint8_t* __restrict__ mask, double* __restrict__ out,
                                                   float* __restrict__ out32, const int* __restrict__ done)
{
    if (done && *done) return;
    coarse_sum_body(N, P, mask, out, out32, int(blockIdx.x));
}

}  // namespace t3d

using namespace t3d;

#include "mixed_internal.hpp"

namespace t3d {
// xc64 (and x32, if given) = Ainv rc64 = Li^T Li rc64, masked
static void launch_coarse(t3d_mixed* s, const uint8_t* mask, float* x32, const int* done)
{
    hipStream_t st = s->p->stream;
    const int N = s->NcInv;
    hipLaunchKernelGGL(k_coarse_tiles, dim3(coarse_tile_blocks(coarse_tiles(N))), dim3(256), 0, st, N, (const double*)s->Ainv,
                       (const double*)s->rc64, s->cP, done);
    hipLaunchKernelGGL(k_coarse_sum, dim3(unsigned(coarse_nt(N))), dim3(CT), 0, st, N, (const double*)s->cP, mask,
                       s->xc64, x32, done);
}

}  // namespace t3d

namespace t3d {

constexpr int KX_NZ = 8;

static dim3 kx_grid(const Grid& g, int nz)
{
    return dim3(unsigned((g.nnx + kx::TX - 1) / kx::TX), unsigned((g.nny + kx::TY - 1) / kx::TY),
                unsigned((g.nnz + nz - 1) / nz));
}

dim3 kww_grid(const Grid& g)
{
    return dim3(unsigned((g.nnx + wk::NX - 1) / wk::NX), unsigned((g.nny + KW_CY - 2) / (KW_CY - 1)),
                unsigned((g.nnz + KW_NZ - 1) / KW_NZ));
}


// y = K u in FP32; form 0: Walsh form, 1: node-gather form, 2: MFMA bf16 x3,
// 3: MFMA bf16 x6
static void launch_kx32(t3d_mixed* s, int form, const float* u, float* y, const int* done = nullptr)
{
    t3d_problem* p = s->p;
    if (form == 2)
        hipLaunchKernelGGL((k_kxw32<KW_NZ, false, KW_CY, 2>), kww_grid(p->g), dim3(32 * KW_CY), 0, p->stream, p->g,
                           (const float*)nullptr, (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, u, nullptr,
                           nullptr, y, nullptr, nullptr, done, s->kwXcd, nullptr, 0, (const unsigned short*)s->KA2);
    else if (form == 3)
        hipLaunchKernelGGL((k_kxw32<KW_NZ, false, KW_CY, 3>), kww_grid(p->g), dim3(32 * KW_CY), 0, p->stream, p->g,
                           (const float*)nullptr, (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, u, nullptr,
                           nullptr, y, nullptr, nullptr, done, s->kwXcd, nullptr, 0, (const unsigned short*)s->KA3);
    else if (form == 0)
        hipLaunchKernelGGL((k_kxw32<KW_NZ, false, KW_CY>), kww_grid(p->g), dim3(32 * KW_CY), 0, p->stream, p->g,
                           (const float*)s->KW, (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, u, nullptr,
                           nullptr, y, nullptr, nullptr, done, s->kwXcd, nullptr, 0);
    else
        hipLaunchKernelGGL((k_kx32<KX_NZ, false>), kx_grid(p->g, KX_NZ), dim3(256), 0, p->stream, p->g,
                           (const float*)s->Ke32, (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, u, nullptr,
                           nullptr, y, nullptr, nullptr, nullptr);
}



unsigned row_blocks(const LevelF& L) { return unsigned((int64_t(L.nny) * L.nnz + RPB - 1) / RPB); }

// the solver's fused operator launch: p = z + beta p_prev, dx += alpha_prev
// p_prev, Ap = K p, p.Ap partials; pid: 1 if pcur is s->pA
static void launch_fused(t3d_mixed* s, const float* pprev, float* pcur, float* Ap, int pid)
{
    t3d_problem* p = s->p;
    if (s->form == 2 || s->form == 3)
        hipLaunchKernelGGL((s->form == 2 ? k_kxw32<KW_NZ, true, KW_CY, 2> : k_kxw32<KW_NZ, true, KW_CY, 3>),
                           kww_grid(p->g), dim3(32 * KW_CY), 0, p->stream, p->g, (const float*)nullptr,
                           (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, pprev, (const float*)s->z, pcur, Ap,
                           s->part, (const double*)s->sc, (const int*)s->iflags, s->kwXcd, s->dx, pid,
                           (const unsigned short*)(s->form == 2 ? s->KA2 : s->KA3));
    else
        hipLaunchKernelGGL((k_kxw32<KW_NZ, true, KW_CY>), kww_grid(p->g), dim3(32 * KW_CY), 0, p->stream, p->g,
                           (const float*)s->KW, (const float*)p->EcellF, (const uint8_t*)p->nodeFixed, pprev,
                           (const float*)s->z, pcur, Ap, s->part, (const double*)s->sc, (const int*)s->iflags,
                           s->kwXcd, s->dx, pid);
}

// x += dx + pending alpha p ; dx = 0 (skipped for a finished problem when done != nullptr)
static void launch_fold(t3d_mixed* s, double* x, const int* done)
{
    t3d_problem* p = s->p;
    const int64_t n = 3 * p->g.nN;
    const unsigned fb = unsigned(std::min<int64_t>(VB, (n + 255) / 256));
    hipLaunchKernelGGL(k_mx_fold, dim3(fb), dim3(256), 0, p->stream, n, x, s->dx, (const float*)s->pA,
                       (const float*)s->pB, (const double*)s->sc, done);
    hipLaunchKernelGGL(k_mx_applied, dim3(1), dim3(1), 0, p->stream, s->sc, done);
}

// r = b - K64 x (FP64 Walsh form, r in FP32) with the FP64 r.r partials in
// s->part, then rn2 = their sum and, with check, the convergence test of
// iteration it (k_fin_rn2); skipped for a finished problem when done != nullptr
static int launch_true_residual(t3d_mixed* s, const double* b, const double* x, const int* done, int it, int check,
                                int guard)
{
    t3d_problem* p = s->p;
    const dim3 gr = resid64_grid(p->g);
    hipLaunchKernelGGL(k_resid64, gr, dim3(256), 0, p->stream, p->g, (const double*)s->KW64,
                       (const double*)p->Ecell, (const uint8_t*)p->nodeFixed, x, b, s->r, s->part, done);
    hipLaunchKernelGGL(k_fin_rn2, dim3(1), dim3(1024), 0, p->stream, (const double*)s->part, gr.x * gr.y * gr.z,
                       s->sc, s->iflags, it, check, guard);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

// z32 = M32(r); rz and beta from the finest prolongation.  tmpReady: the
// x-pass of the finest restriction is already in s->tmp (fused into the x, r
// update)
// z32 = M32(r) with MATLAB's 'Standard' V-cycle (the kernels above); rz and beta
static int mixed_vcycle_std(t3d_mixed* s, const MGView& v, const float* r, int flexible, int first, const int* done)
{
    t3d_problem* p = s->p;
    hipStream_t st = p->stream;
    const int L = v.nlev - 1;
    const bool chol = v.Nc > 0;
    const float w = float(v.weight);
    T3D_CHECK(L >= 1 && s->ys32[0], T3D_ERR_STATE, "mixed 'Standard' V-cycle: t3d_mixed_setup not called");
    auto nodes = [](int64_t nN) { return dim3(unsigned((nN + 255) / 256)); };
    // y = A_l x
    auto apply = [&](int l, const float* x, float* y) {
        if (l == 0)
            launch_kx32(s, s->form, x, y, done);
        else
            hipLaunchKernelGGL(k_gal_apply32, nodes(v.nN[l]), dim3(256), 0, st, levf(v, l), (const float*)s->Kel32[l],
                               x, y, done);
    };
    for (int l = 0; l < L; ++l) {
        const bool last = (l + 1 == L);
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        const int span = v.span[l + 1];
        const float* rl = l == 0 ? r : s->r32[l];
        float* xl = l == 0 ? s->z : s->x32[l];
        float* yl = s->ys32[l];
        hipLaunchKernelGGL(k_std32_point, nodes(Lf.nN), dim3(256), 0, st, 0, Lf.nN, rl, (const float*)nullptr,
                           (const float*)s->D32[l], v.mask[l], w, xl, done);
        apply(l, xl, yl);
        hipLaunchKernelGGL(k_std32_point, nodes(Lf.nN), dim3(256), 0, st, 1, Lf.nN, rl, (const float*)yl,
                           (const float*)nullptr, v.mask[l], w, yl, done);
        // r_{l+1} = P^T (r_l - A_l x_l)
        const dim3 g2(unsigned((Lc.nnx + 63) / 64), unsigned(Lc.nny), unsigned(Lc.nnz));
        const float* Dc = (last && !chol) ? (const float*)s->D32[L] : nullptr;
        double* rcd = (last && chol) ? s->rc64 : nullptr;
        if (l > 0 && (span == 2 || span == 4)) {   // coarse levels: one pass
            hipLaunchKernelGGL(k_restrict_c32, g2, dim3(64), 0, st, Lf, Lc, span, (const float*)yl, v.mask[l + 1],
                               s->r32[l + 1], Dc, s->x32[L], rcd, done, v.norm);
            T3D_HIP(hipGetLastError());
            continue;
        }
        if (l == 0) {
            hipLaunchKernelGGL(k_rx_rows, dim3(row_blocks(Lf)), dim3(256), sizeof(float) * RPB * 3 * Lf.nnx, st,
                               Lf.nnx, Lf.nny * Lf.nnz, (const float*)yl, Lc.nnx, span, s->tmp, done);
        } else {
            const dim3 g1(unsigned((Lc.nnx + 63) / 64), unsigned(Lf.nny), unsigned(Lf.nnz));
            hipLaunchKernelGGL(k_restrict_x<float>, g1, dim3(64), 0, st, Lf, Lc.nnx, span, (const float*)yl, s->tmp,
                               done);
        }
        hipLaunchKernelGGL(k_restrict_yz, g2, dim3(64), 0, st, Lf, Lc, span, (const float*)s->tmp, v.mask[l + 1],
                           s->r32[l + 1], Dc, s->x32[L], rcd, done, v.norm);
        T3D_HIP(hipGetLastError());
    }
    if (chol) {
        T3D_CHECK(s->NcInv == v.Nc, T3D_ERR_STATE, "mixed V-cycle: t3d_mixed_setup not called after t3d_mg_setup");
        launch_coarse(s, v.mask[L], s->x32[L], done);
    }
    for (int l = L - 1; l >= 1; --l) {
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        hipLaunchKernelGGL(k_prolong_add32, nodes(Lf.nN), dim3(256), 0, st, Lf, Lc, v.span[l + 1],
                           (const float*)s->x32[l + 1], v.mask[l], s->x32[l], done);
        apply(l, s->x32[l], s->ys32[l]);
        hipLaunchKernelGGL(k_std32_point, nodes(Lf.nN), dim3(256), 0, st, 2, Lf.nN, (const float*)s->r32[l],
                           (const float*)s->ys32[l], (const float*)s->D32[l], v.mask[l], w, s->x32[l], done);
    }
    // finest: z = w r Dinv + P x_1 (row kernel, the pre-smoothed value recomputed), then the post-smoothing
    const LevelF L0 = levf(v, 0), L1 = levf(v, 1);
    const unsigned nb = row_blocks(L0);
    if (3 * L0.nnx <= 256 * 5)
        hipLaunchKernelGGL((k_pro_rows<5, 1, false>), dim3(nb), dim3(256), sizeof(float) * RPB * 3 * L1.nnx, st, L0,
                           L1, v.span[1], (const float*)s->x32[1], r, (const float*)s->D32[0], v.mask[0], w, s->z,
                           (const float*)nullptr, (double*)nullptr, done);
    else
        hipLaunchKernelGGL((k_pro_rows<8, 1, false>), dim3(nb), dim3(256), sizeof(float) * RPB * 3 * L1.nnx, st, L0,
                           L1, v.span[1], (const float*)s->x32[1], r, (const float*)s->D32[0], v.mask[0], w, s->z,
                           (const float*)nullptr, (double*)nullptr, done);
    apply(0, s->z, s->ys32[0]);
    const unsigned fb = unsigned(std::min<int64_t>(VB, (L0.nN + 255) / 256));
    hipLaunchKernelGGL(k_std32_final, dim3(fb), dim3(256), 0, st, L0.nN, r, (const float*)s->ys32[0],
                       (const float*)s->D32[0], v.mask[0], w, s->z, (const float*)s->Ap, s->part, done);
    hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(1024), 0, st, (const double*)s->part, fb, s->sc, flexible, first,
                       done);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

int mixed_vcycle(t3d_mixed* s, const MGView& v, const float* r, int flexible, int first, bool tmpReady,
                        const int* done)
{
    if (v.vtype == 1) return mixed_vcycle_std(s, v, r, flexible, first, done);
    t3d_problem* p = s->p;
    hipStream_t st = p->stream;
    const int L = v.nlev - 1;
    const bool chol = v.Nc > 0;
    T3D_CHECK(L >= 1, T3D_ERR_STATE, "mixed V-cycle needs at least two levels");
    for (int l = 0; l < L; ++l) {
        const bool last = (l + 1 == L);
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        const int span = v.span[l + 1];
        const float* Dc = (last && !chol) ? s->D32[L] : nullptr;
        double* rcd = (last && chol) ? s->rc64 : nullptr;
        const dim3 g2(unsigned((Lc.nnx + 63) / 64), unsigned(Lc.nny), unsigned(Lc.nnz));
        if (l > 0 && (span == 2 || span == 4)) {   // coarse levels: one pass
            hipLaunchKernelGGL(k_restrict_c32, g2, dim3(64), 0, st, Lf, Lc, span, (const float*)s->r32[l],
                               v.mask[l + 1], s->r32[l + 1], Dc, s->x32[L], rcd, done, v.norm);
            T3D_HIP(hipGetLastError());
            continue;
        }
        if (l == 0) {
            if (!tmpReady)
                hipLaunchKernelGGL(k_rx_rows, dim3(row_blocks(Lf)), dim3(256), sizeof(float) * RPB * 3 * Lf.nnx, st,
                                   Lf.nnx, Lf.nny * Lf.nnz, r, Lc.nnx, span, s->tmp, done);
        } else {
            const dim3 g1(unsigned((Lc.nnx + 63) / 64), unsigned(Lf.nny), unsigned(Lf.nnz));
            hipLaunchKernelGGL(k_restrict_x<float>, g1, dim3(64), 0, st, Lf, Lc.nnx, span, (const float*)s->r32[l],
                               s->tmp, done);
        }
        hipLaunchKernelGGL(k_restrict_yz, g2, dim3(64), 0, st, Lf, Lc, span, (const float*)s->tmp, v.mask[l + 1],
                           s->r32[l + 1], Dc, s->x32[L], rcd, done, v.norm);
        T3D_HIP(hipGetLastError());
    }
    if (chol) {
        T3D_CHECK(s->NcInv == v.Nc, T3D_ERR_STATE, "mixed V-cycle: t3d_mixed_setup not called after t3d_mg_setup");
        launch_coarse(s, v.mask[L], s->x32[L], done);
        T3D_HIP(hipGetLastError());
    }
    const float w = float(v.weight);
    for (int l = L - 1; l >= 1; --l) {
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        const dim3 gp(unsigned((Lf.nnx + 63) / 64), unsigned(Lf.nny), unsigned(Lf.nnz));
        hipLaunchKernelGGL(k_prolong_smooth32, gp, dim3(64), 0, st, Lf, Lc, v.span[l + 1], (const float*)s->x32[l + 1],
                           (const float*)s->r32[l], (const float*)s->D32[l], v.mask[l], w, s->x32[l], done);
        T3D_HIP(hipGetLastError());
    }
    const LevelF L0 = levf(v, 0), L1 = levf(v, 1);
    const unsigned nb = row_blocks(L0);
    // positions per pass: 5 chunks cover rows of up to 426 nodes in one pass
    if (3 * L0.nnx <= 256 * 5)
        hipLaunchKernelGGL(k_pro_rows<5>, dim3(nb), dim3(256), sizeof(float) * RPB * 3 * L1.nnx, st, L0, L1, v.span[1],
                       (const float*)s->x32[1], r, (const float*)s->D32[0], v.mask[0], w, s->z, (const float*)s->Ap,
                       s->part, done);
    else
        hipLaunchKernelGGL(k_pro_rows<8>, dim3(nb), dim3(256), sizeof(float) * RPB * 3 * L1.nnx, st, L0, L1, v.span[1],
                       (const float*)s->x32[1], r, (const float*)s->D32[0], v.mask[0], w, s->z, (const float*)s->Ap,
                       s->part, done);
    hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(1024), 0, st, (const double*)s->part, nb, s->sc, flexible, first,
                       done);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

// The fused operator kernel's memory streams as a plain copy (the bandwidth
// this GPU gives that access mix without a stencil): reads u, z, dx and the
// cell moduli, writes pcur = z + w u, y = z and dx += w u, one float4 of each
// node vector per thread (3 nN / 4 of them; the moduli by the first ncell / 4
// threads, added as w E).  w is a kernel argument (0 when timed), so no
// access can be elided at compile time.
__global__ void __launch_bounds__(256) k_copy_fused_streams(int64_t n4, int64_t e4, const float4* __restrict__ u,
                                                            const float4* __restrict__ z, float4* __restrict__ dx,
                                                            const float4* __restrict__ E, float4* __restrict__ pcur,
                                                            float4* __restrict__ y, float w)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    const float4 a = u[i], b = z[i];
    float4 c = dx[i];
    float4 pv = b;
    if (i < e4) {
        const float4 e = E[i];
        pv.x = fmaf(w, e.x, pv.x);
        pv.y = fmaf(w, e.y, pv.y);
        pv.z = fmaf(w, e.z, pv.z);
        pv.w = fmaf(w, e.w, pv.w);
    }
    pv.x = fmaf(w, a.x, pv.x);
    pv.y = fmaf(w, a.y, pv.y);
    pv.z = fmaf(w, a.z, pv.z);
    pv.w = fmaf(w, a.w, pv.w);
    c.x = fmaf(w, a.x, c.x);
    c.y = fmaf(w, a.y, c.y);
    c.z = fmaf(w, a.z, c.z);
    c.w = fmaf(w, a.w, c.w);
    pcur[i] = pv;
    y[i] = b;
    dx[i] = c;
}

}  // namespace t3d

namespace t3d {

// helpers of the batched solver (batch.hip), one problem each, on its stream

// r = b - K64 x in FP32 ; rn2 (no convergence test)
int mixed_true_residual(t3d_mixed* s, const double* b, const double* x)
{
    return launch_true_residual(s, b, x, nullptr, -1, 0, 0);
}

// the residual replacement of iteration it, skipped when the problem is done:
// x += dx, r = b - K64 x, rn2 and the convergence test
int mixed_replace_residual(t3d_mixed* s, const double* b, double* x, int it)
{
    launch_fold(s, x, (const int*)s->iflags);
    return launch_true_residual(s, b, x, (const int*)s->iflags, it, 1, 1);
}

// x += dx ; r = b - K64 x ; returns the FP64 r.r and the stopping threshold
int mixed_confirm(t3d_mixed* s, const double* b, double* x, double* rn2, double* stop)
{
    t3d_problem* p = s->p;
    int rc = mixed_fold(s, x);
    if (rc) return rc;
    if ((rc = mixed_true_residual(s, b, x))) return rc;
    T3D_HIP(hipMemcpyAsync(stop, s->sc + S_STOP, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipMemcpyAsync(rn2, s->sc + S_RN2, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

// x += dx ; dx = 0 (unconditional)
int mixed_fold(t3d_mixed* s, double* x)
{
    launch_fold(s, x, nullptr);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

// continue a problem whose convergence was not confirmed: clear its flag, run
// the V-cycle of iteration itDone with the replaced residual, and place the
// direction of iteration itDone where the iteration itNext reads p_prev
int mixed_resume(t3d_mixed* s, int itDone, int itNext, int flexible)
{
    t3d_problem* p = s->p;
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mx_clear_done, dim3(1), dim3(1), 0, p->stream, s->iflags);
    if ((rc = mixed_vcycle(s, v, s->r, flexible, 0, false, s->iflags))) return rc;
    float* last = (itDone & 1) ? s->pA : s->pB;   // p_cur of iteration itDone
    float* next = (itNext & 1) ? s->pB : s->pA;   // p_prev of iteration itNext
    if (last != next)
        T3D_HIP(hipMemcpyAsync(next, last, sizeof(float) * 3 * p->g.nN, hipMemcpyDeviceToDevice, p->stream));
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

}  // namespace t3d

extern "C" {

int t3d_mixed_destroy(t3d_mixed* s)
{
    if (!s) return T3D_OK;
    void* b[] = {s->Ke32, s->KW, s->KW64, s->KA2, s->KA3, s->r, s->dx, s->tmp, s->rc64, s->xc64,
                 s->Li, s->ownAinv ? s->Ainv : nullptr, s->yc64, s->cP,
                 s->pA, s->pB, s->Ap, s->z, s->part, s->sc, s->iflags};
    for (void* q : b)
        if (q) (void)hipFree(q);
    for (hipEvent_t e : s->tev)
        if (e) (void)hipEventDestroy(e);
    for (int l = 0; l < 8; ++l) {
        if (s->D32[l]) (void)hipFree(s->D32[l]);
        if (s->r32[l]) (void)hipFree(s->r32[l]);
        if (s->x32[l]) (void)hipFree(s->x32[l]);
        if (s->Kel32[l]) (void)hipFree(s->Kel32[l]);
        if (s->ys32[l]) (void)hipFree(s->ys32[l]);
    }
    delete s;
    return T3D_OK;
}

int t3d_mixed_create(t3d_problem* p, t3d_mg* M, t3d_mixed** out)
{
    T3D_CHECK(p && M && out, T3D_ERR_ARG, "t3d_mixed_create: null argument");
    *out = nullptr;
    MGView v;
    int rc = mg_view(M, v);
    if (rc) return rc;
    T3D_CHECK(v.nlev >= 2, T3D_ERR_STATE, "t3d_mixed_create: the hierarchy needs at least two levels");
    t3d_mixed* s = new t3d_mixed();
    s->p = p;
    s->M = M;
    s->nlev = v.nlev;
    auto fail = [&]() {
        set_error("t3d_mixed_create: device allocation failed");
        t3d_mixed_destroy(s);
        return T3D_ERR_HIP;
    };
    float K32[576];
    for (int i = 0; i < 576; ++i) K32[i] = float(p->host.Ke[i]);
    // Walsh form: K^ = (H x I3) Ke (H x I3)^T; check its sparsity against the kernel's pattern
    float KW[wk::NNZ];
    double KW64[wk::NNZ];
    {
        static const int off[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0},
                                      {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};   // o(b), fea.cpp node order
        double Hm[8][8];
        for (int m = 0; m < 8; ++m)
            for (int b = 0; b < 8; ++b) {
                const int dotp = (m & 1) * off[b][0] + ((m >> 1) & 1) * off[b][1] + ((m >> 2) & 1) * off[b][2];
                Hm[m][b] = (dotp & 1) ? -1.0 : 1.0;
            }
        static double Kh[24][24];
        double mx = 0.0;
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double acc = 0.0;
                for (int b = 0; b < 8; ++b)
                    for (int bb = 0; bb < 8; ++bb)
                        acc += Hm[i / 3][b] * p->host.Ke[(3 * b + i % 3) * 24 + 3 * bb + j % 3] * Hm[j / 3][bb];
                Kh[i][j] = acc;
                mx = std::max(mx, std::fabs(acc));
            }
        const unsigned char outIdx[wk::NNZ] = {3,  3,  3,  4,  4,  5,  5,  6,  6,  7,  7,  7,  8,  8,  9,
                                               9,  10, 10, 11, 11, 11, 12, 12, 13, 13, 14, 14, 14, 15, 15,
                                               16, 16, 16, 17, 17, 18, 18, 18, 19, 19, 20, 20, 21, 22, 23};
        const unsigned char inIdx[wk::NNZ] = {3,  7,  14, 4,  6,  5,  12, 4,  6,  3,  7,  14, 8,  13, 9,
                                              20, 10, 17, 11, 16, 18, 5,  12, 8,  13, 3,  7,  14, 15, 19,
                                              11, 16, 18, 10, 17, 11, 16, 18, 15, 19, 9,  20, 21, 22, 23};
        bool inPat[24][24] = {};
        for (int k = 0; k < wk::NNZ; ++k) {
            inPat[outIdx[k]][inIdx[k]] = true;
            KW64[k] = Kh[outIdx[k]][inIdx[k]] / 64.0;
            KW[k] = float(KW64[k]);
        }
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j)
                T3D_CHECK(inPat[i][j] || std::fabs(Kh[i][j]) <= 1e-12 * mx, T3D_ERR_STATE,
                          "t3d_mixed_create: element matrix does not have the Walsh-form structure");
    }
    const int64_t nD = 3 * p->g.nN;
    const dim3 kg = kww_grid(p->g);
    const LevelF L0 = levf(v, 0);
    // partials: the largest layout any caller uses (Walsh tiles down to 8 planes
    // deep, row kernels down to 2 rows per block, the vector kernels)
    const int64_t kwMax = int64_t(kg.x) * kg.y * ((p->g.nnz + 7) / 8);
    const int64_t rowsMax = (int64_t(L0.nny) * L0.nnz + 1) / 2;
    const dim3 rg = resid64_grid(p->g);
    const int64_t npart = std::max<int64_t>({kwMax, 2 * rowsMax, 2 * int64_t(VB), int64_t(rg.x) * rg.y * rg.z}) + 16;
    int64_t ntmp = 0;
    for (int l = 0; l + 1 < v.nlev; ++l) {
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        ntmp = std::max<int64_t>(ntmp, 3 * int64_t(Lc.nnx) * Lf.nny * Lf.nnz);
    }
    // MFMA A operands: A[i][k] = Ke[3 b_i + c_i][3 b_k + c_k] for the output
    // row order of the MFMA form (mfma_row_c / mfma_row_b: each lane receives
    // all eight nodes of its components) and the component-major inputs
    // k = 8 c_k + b_k (24x24 padded to 32x32), split into 2 or 3 bf16 pieces
    // (round to nearest even), in the instruction's lane layout
    // [K step][piece][lane][8]: lane l holds row l%32, columns 16 ks + 8 (l/32) + e
    std::vector<unsigned short> KAh2(2 * 2 * 64 * 8), KAh3(2 * 3 * 64 * 8);
    {
        auto rne = [](float x) {
            unsigned u;
            std::memcpy(&u, &x, 4);
            return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
        };
        auto val = [](unsigned short h) {
            const unsigned u = unsigned(h) << 16;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        for (int ks = 0; ks < 2; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int i = l % 32, k = 16 * ks + 8 * (l / 32) + e;
                    float a = 0.0f;
                    const int ci = mfma_row_c(i), bi = mfma_row_b(i);
                    if (ci >= 0 && k < 24) {
                        const int ck = k / 8, bk = k % 8;
                        a = float(p->host.Ke[(3 * bi + ci) * 24 + 3 * bk + ck]);
                    }
                    const unsigned short h = rne(a);
                    const float r1 = a - val(h);
                    const unsigned short m = rne(r1);
                    KAh2[((ks * 2 + 0) * 64 + l) * 8 + e] = h;
                    KAh2[((ks * 2 + 1) * 64 + l) * 8 + e] = m;
                    KAh3[((ks * 3 + 0) * 64 + l) * 8 + e] = h;
                    KAh3[((ks * 3 + 1) * 64 + l) * 8 + e] = m;
                    KAh3[((ks * 3 + 2) * 64 + l) * 8 + e] = rne(r1 - val(m));
                }
    }
    if (hipMalloc(&s->Ke32, sizeof(float) * 576) != hipSuccess ||
        hipMalloc(&s->KA2, sizeof(unsigned short) * KAh2.size()) != hipSuccess ||
        hipMemcpy(s->KA2, KAh2.data(), sizeof(unsigned short) * KAh2.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->KA3, sizeof(unsigned short) * KAh3.size()) != hipSuccess ||
        hipMemcpy(s->KA3, KAh3.data(), sizeof(unsigned short) * KAh3.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->Ke32, K32, sizeof(float) * 576, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->KW, sizeof(float) * wk::NNZ) != hipSuccess ||
        hipMemcpy(s->KW, KW, sizeof(float) * wk::NNZ, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->KW64, sizeof(double) * wk::NNZ) != hipSuccess ||
        hipMemcpy(s->KW64, KW64, sizeof(double) * wk::NNZ, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->pA, sizeof(float) * nD) != hipSuccess || hipMalloc(&s->pB, sizeof(float) * nD) != hipSuccess ||
        hipMalloc(&s->Ap, sizeof(float) * nD) != hipSuccess || hipMalloc(&s->z, sizeof(float) * nD) != hipSuccess ||
        hipMalloc(&s->r, sizeof(float) * nD) != hipSuccess || hipMalloc(&s->dx, sizeof(float) * nD) != hipSuccess ||
        hipMemset(s->z, 0, sizeof(float) * nD) != hipSuccess ||
        hipMalloc(&s->tmp, sizeof(float) * ntmp) != hipSuccess ||
        hipMalloc(&s->part, sizeof(double) * npart) != hipSuccess ||
        hipMalloc(&s->sc, sizeof(double) * 64) != hipSuccess || hipMemset(s->sc, 0, sizeof(double) * 64) != hipSuccess ||
        hipMalloc(&s->iflags, sizeof(int) * 4) != hipSuccess ||
        hipMemset(s->iflags, 0, sizeof(int) * 4) != hipSuccess)
        return fail();
    for (int l = 0; l < v.nlev; ++l) {
        if (hipMalloc(&s->D32[l], sizeof(float) * 3 * v.nN[l]) != hipSuccess) return fail();
        if (l > 0 && (hipMalloc(&s->r32[l], sizeof(float) * 3 * v.nN[l]) != hipSuccess ||
                      hipMalloc(&s->x32[l], sizeof(float) * 3 * v.nN[l]) != hipSuccess))
            return fail();
    }
    const int64_t nc = 3 * v.nN[v.nlev - 1];
    if (hipMalloc(&s->rc64, sizeof(double) * nc) != hipSuccess || hipMalloc(&s->xc64, sizeof(double) * nc) != hipSuccess)
        return fail();
    // the set-up copies and fills above ran on the null stream
    T3D_HIP(hipDeviceSynchronize());
    *out = s;
    return T3D_OK;
}

// refresh the FP32 inverse diagonals after t3d_mg_setup
int t3d_mixed_setup(t3d_mixed* s)
{
    T3D_CHECK(s, T3D_ERR_ARG, "t3d_mixed_setup: null");
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    hipStream_t st = s->p->stream;
    if (v.vtype == 1) {   // 'Standard' V-cycle: FP32 element matrices and per-level work vectors
        const int L = v.nlev - 1;
        for (int l = 0; l < L; ++l) {
            const size_t ny = size_t(3) * size_t(v.nN[l]);
            if (s->ys32n[l] < ny) {
                if (s->ys32[l]) (void)hipFree(s->ys32[l]);
                s->ys32[l] = nullptr;
                s->ys32n[l] = 0;
                T3D_HIP(hipMalloc(&s->ys32[l], sizeof(float) * ny));
                s->ys32n[l] = ny;
            }
            if (l == 0) continue;
            T3D_CHECK(v.Kel[l], T3D_ERR_STATE, "t3d_mixed_setup: the 'Standard' V-cycle needs the Galerkin form");
            const size_t nk = size_t(576) * size_t(v.resX[l]) * size_t(v.resY[l]) * size_t(v.resZ[l]);
            if (s->Kel32n[l] < nk) {
                if (s->Kel32[l]) (void)hipFree(s->Kel32[l]);
                s->Kel32[l] = nullptr;
                s->Kel32n[l] = 0;
                T3D_HIP(hipMalloc(&s->Kel32[l], sizeof(float) * nk));
                s->Kel32n[l] = nk;
            }
            hipLaunchKernelGGL(k_kel32, dim3(unsigned((nk + 255) / 256)), dim3(256), 0, st, int64_t(nk / 576),
                               v.Kel[l], s->Kel32[l]);
            T3D_HIP(hipGetLastError());
        }
    }
    for (int l = 0; l < v.nlev; ++l) {
        const int64_t n = 3 * v.nN[l];
        hipLaunchKernelGGL(k_dinv, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, v.D[l], s->D32[l]);
        T3D_HIP(hipGetLastError());
    }
    // inverse of the coarsest Cholesky factor (the coarsest size is fixed by
    // the hierarchy; the factor exists only when the Galerkin matrix is SPD)
    s->NcInv = 0;
    if (v.Nc > 0 && v.Ainv) {   // the multigrid object keeps the coarsest inverse
        if (s->ownAinv && s->Ainv) (void)hipFree(s->Ainv);
        if (s->Li) (void)hipFree(s->Li);
        s->Li = nullptr;
        s->Ainv = const_cast<double*>(v.Ainv);
        s->ownAinv = false;
        s->NcInv = v.Nc;
    } else if (v.Nc > 0) {
        const size_t n2 = size_t(v.Nc) * size_t(v.Nc);
        if (!s->ownAinv) {
            s->Ainv = nullptr;
            s->ownAinv = true;
        }
        if (!s->Li) T3D_HIP(hipMalloc(&s->Li, sizeof(double) * n2));
        if (!s->Ainv) T3D_HIP(hipMalloc(&s->Ainv, sizeof(double) * coarse_packed_doubles(v.Nc)));
        T3D_HIP(launch_tri_inv(st, v.Nc, v.L, s->Li));
        hipLaunchKernelGGL(k_gram, dim3(coarse_tiles(v.Nc)), dim3(256), 0, st, v.Nc, (const double*)s->Li, s->Ainv);
        T3D_HIP(hipGetLastError());
        s->NcInv = v.Nc;
    }
    if (s->NcInv > 0 && s->cPn < coarse_partial_doubles(s->NcInv)) {
        if (s->cP) (void)hipFree(s->cP);
        s->cP = nullptr;
        s->cPn = 0;
        T3D_HIP(hipMalloc(&s->cP, sizeof(double) * coarse_partial_doubles(s->NcInv)));
        s->cPn = coarse_partial_doubles(s->NcInv);
    }
    T3D_HIP(hipStreamSynchronize(st));
    return T3D_OK;
}

// z = M32(r): one application of the mixed solver's FP32 V-cycle (the
// multigrid object's cycle type) to an FP32 full-layout residual
int t3d_mixed_vcycle(t3d_mixed* s, const float* r, float* z)
{
    T3D_CHECK(s && r && z, T3D_ERR_ARG, "t3d_mixed_vcycle: null argument");
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    hipStream_t st = s->p->stream;
    hipLaunchKernelGGL(k_mx_clear_done, dim3(1), dim3(1), 0, st, s->iflags);
    T3D_HIP(hipMemsetAsync(s->Ap, 0, sizeof(float) * 3 * size_t(s->p->g.nN), st));   // the beta partials read it
    if ((rc = mixed_vcycle(s, v, r, 0, 1, false, s->iflags))) return rc;
    T3D_HIP(hipMemcpyAsync(z, s->z, sizeof(float) * 3 * size_t(s->p->g.nN), hipMemcpyDeviceToDevice, st));
    T3D_HIP(hipStreamSynchronize(st));
    return T3D_OK;
}

// r = b - K64 x (the replacement's FP64 Walsh-form residual) into r (FP32) and its FP64 r.r
int t3d_mixed_true_residual(t3d_mixed* s, const double* b, const double* x, float* r, double* rn2)
{
    T3D_CHECK(s && b && x && r && rn2, T3D_ERR_ARG, "t3d_mixed_true_residual: null argument");
    hipStream_t st = s->p->stream;
    int rc = launch_true_residual(s, b, x, nullptr, -1, 0, 0);
    if (rc) return rc;
    T3D_HIP(hipMemcpyAsync(r, s->r, sizeof(float) * 3 * size_t(s->p->g.nN), hipMemcpyDeviceToDevice, st));
    T3D_HIP(hipMemcpyAsync(rn2, s->sc + S_RN2, sizeof(double), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));
    return T3D_OK;
}

// x = the mixed solver's coarsest solve of b (device arrays of the coarsest
// size); a test hook for the inverse-factor form
int t3d_mixed_coarse_solve(t3d_mixed* s, const double* b, double* x)
{
    T3D_CHECK(s && b && x, T3D_ERR_ARG, "t3d_mixed_coarse_solve: null argument");
    T3D_CHECK(s->NcInv > 0, T3D_ERR_STATE, "t3d_mixed_coarse_solve: no Cholesky factor at the coarsest level");
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    hipStream_t st = s->p->stream;
    T3D_HIP(hipMemcpyAsync(s->rc64, b, sizeof(double) * s->NcInv, hipMemcpyDeviceToDevice, st));
    launch_coarse(s, v.mask[v.nlev - 1], nullptr, nullptr);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipMemcpyAsync(x, s->xc64, sizeof(double) * s->NcInv, hipMemcpyDeviceToDevice, st));
    T3D_HIP(hipStreamSynchronize(st));
    return T3D_OK;
}

// y = K u in FP32 (full layout, fixed DOFs masked), the modulus last set with
// t3d_set_modulus; the single-precision counterpart of t3d_kxu_f64.
// form 0: Walsh form (used by the solver), 1: node-gather form
int t3d_kxu_f32(t3d_mixed* s, int form, const float* u, float* y)
{
    T3D_CHECK(s && u && y && form >= 0 && form <= 3, T3D_ERR_ARG, "t3d_kxu_f32: bad argument");
    launch_kx32(s, form, u, y);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(s->p->stream));
    return T3D_OK;
}

// selects the operator of the solver: 0 Walsh form (default), 2 MFMA bf16 x3,
// 3 MFMA bf16 x6
int t3d_mixed_set_operator(t3d_mixed* s, int form)
{
    T3D_CHECK(s && (form == 0 || form == 2 || form == 3), T3D_ERR_ARG,
              "t3d_mixed_set_operator: form must be 0, 2 or 3");
    s->form = form;
    return T3D_OK;
}

// HIP-event timing on the problem's stream, average milliseconds per launch:
// op 0..3: the FP32 operator y = K u of form 0 (Walsh) / 1 (node gather) /
// 2 (MFMA bf16 x3) / 3 (MFMA bf16 x6); op 4: the solver's fused kernel of the
// selected operator (p = z + beta u staged and written to the solver's p
// buffer, y = K p, block partials of p.y), as launched once per PCG iteration
int t3d_mixed_time_op(t3d_mixed* s, int op, const float* u, float* y, int reps, float* ms)
{
    T3D_CHECK(s && u && y && reps > 0 && ms && op >= 0 && op <= 7, T3D_ERR_ARG, "t3d_mixed_time_op: bad argument");
    const int order = s->kwXcd;
    if (op == 5 || op == 6) {   // 5: fused, 6: Walsh-form operator, in the other tile order
        s->kwXcd = !order;
        op = op == 5 ? 4 : 0;
    }
    t3d_problem* p = s->p;
    hipEvent_t e0, e1;
    T3D_HIP(hipEventCreate(&e0));
    T3D_HIP(hipEventCreate(&e1));
    T3D_HIP(hipEventRecord(e0, p->stream));
    const int64_t n4 = 3 * p->g.nN / 4;
    for (int i = 0; i < reps; ++i) {
        if (op < 4)
            launch_kx32(s, op, u, y);
        else if (op == 4)
            launch_fused(s, u, s->pB, y, 0);
        else
            hipLaunchKernelGGL(t3d::k_copy_fused_streams, dim3(unsigned((n4 + 255) / 256)), dim3(256), 0, p->stream, n4,
                               p->g.ncell / 4, (const float4*)u, (const float4*)s->z, (float4*)s->dx,
                               (const float4*)p->EcellF, (float4*)s->pB, (float4*)y, 0.0f);
    }
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipEventRecord(e1, p->stream));
    T3D_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    T3D_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / float(reps);
    s->kwXcd = order;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return T3D_OK;
}

// per-launch HIP-event timing of the solver's fused operator kernel inside
// t3d_pcg_mixed: on != 0 creates the events and zeroes the totals
int t3d_mixed_set_timing(t3d_mixed* s, int on)
{
    T3D_CHECK(s, T3D_ERR_ARG, "t3d_mixed_set_timing: null");
    if (on && !s->tev[0])
        for (hipEvent_t& e : s->tev) T3D_HIP(hipEventCreate(&e));
    s->timing = on != 0;
    s->timedMs = 0.0;
    s->timedLaunches = 0;
    return T3D_OK;
}

int t3d_mixed_get_timing(const t3d_mixed* s, double* ms_total, long long* launches)
{
    T3D_CHECK(s && ms_total && launches, T3D_ERR_ARG, "t3d_mixed_get_timing: null");
    *ms_total = s->timedMs;
    *launches = s->timedLaunches;
    return T3D_OK;
}

// Mixed-precision MG-PCG on the full layout (b, x FP64, device).  The problem's
// modulus and the multigrid object (after t3d_mg_setup + t3d_mixed_setup) are
// used; trueEvery > 0 recomputes the residual in FP64 every trueEvery iterations.
int t3d_pcg_mixed(t3d_mixed* s, const double* b, double* x, double tol, int maxIt, int trueEvery, int flexible,
                  int* iters)
{
    T3D_CHECK(s && b && x && maxIt >= 0, T3D_ERR_ARG, "t3d_pcg_mixed: bad argument");
    t3d_problem* p = s->p;
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    hipStream_t st = p->stream;
    const Grid& g = p->g;
    const int64_t n = 3 * g.nN;
    float* r = s->r;
    double* S = s->sc;
    int* flags = s->iflags;
    const dim3 kg = kww_grid(g);   // tiles of the fused operator kernel
    const unsigned nfused = kg.x * kg.y * kg.z;
    const LevelF L0 = levf(v, 0), L1 = levf(v, 1);
    const unsigned xrb = row_blocks(L0);
    // the 'Standard' V-cycle restricts r - A_0 x_0, not r: the x/r update then skips the restriction's x-pass
    const bool stdCycle = v.vtype == 1;
    // r = b - K64 x ; stop = tol max(1e-30, |b|) ; z = M r ; rz ; beta = 0
    if ((rc = canon_dot_dev(p, b, b, n, S + S_NORMB2))) return rc;
    hipLaunchKernelGGL(k_mx_init, dim3(1), dim3(1), 0, st, tol, S, flags);
    if ((rc = launch_true_residual(s, b, x, nullptr, -1, 0, 0))) return rc;
    T3D_HIP(hipMemsetAsync(s->Ap, 0, sizeof(float) * n, st));
    T3D_HIP(hipMemsetAsync(s->pA, 0, sizeof(float) * n, st));
    T3D_HIP(hipMemsetAsync(s->dx, 0, sizeof(float) * n, st));
    if ((rc = mixed_vcycle(s, v, r, flexible, 1, false, flags))) return rc;
    // the initial residual norm, against which the solve checks for divergence
    double rn2First = 0.0, stopHost = 0.0;
    T3D_HIP(hipMemcpyAsync(&rn2First, S + S_RN2, sizeof(double), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipMemcpyAsync(&stopHost, S + S_STOP, sizeof(double), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));
    // iteration it reads p_prev and writes p_cur; the two buffers alternate
    auto pPrevOf = [&](int it) { return (it & 1) ? s->pB : s->pA; };
    auto pCurOf = [&](int it) { return (it & 1) ? s->pA : s->pB; };
    const int CHUNK = 32;
    int chunk = CHUNK;
    int result = maxIt;
    bool confirmed = false;
    int it0 = 0;
    while (it0 < maxIt) {
        const int it1 = std::min(maxIt, it0 + chunk);
        for (int it = it0; it < it1; ++it) {
            // p = z + beta p_prev ; Ap = K32 p ; alpha = rz / p.Ap
            if (s->timing) T3D_HIP(hipEventRecord(s->tev[2 * (it - it0)], st));
            launch_fused(s, pPrevOf(it), pCurOf(it), s->Ap, it & 1);
            if (s->timing) T3D_HIP(hipEventRecord(s->tev[2 * (it - it0) + 1], st));
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(1024), 0, st, (const double*)s->part, nfused, S,
                               (const int*)flags);
            const bool replace = trueEvery > 0 && ((it + 1) % trueEvery) == 0;
            // dx += alpha p ; r -= alpha Ap ; rn2 ; convergence test
            hipLaunchKernelGGL(k_xr_rows, dim3(xrb), dim3(256), sizeof(float) * RPB * 3 * L0.nnx, st, L0.nnx,
                               L0.nny * L0.nnz, (const float*)s->Ap, r, s->part,
                               (const double*)S, L1.nnx, v.span[1], (replace || stdCycle) ? nullptr : s->tmp,
                               (const int*)flags);
            hipLaunchKernelGGL(k_fin_rn2, dim3(1), dim3(1024), 0, st, (const double*)s->part, xrb, S, flags, it,
                               replace ? 0 : 1, 1);
            if (replace) {
                // x += dx, then residual replacement with the FP64 operator (the
                // direction is kept)
                launch_fold(s, x, (const int*)flags);
                if ((rc = launch_true_residual(s, b, x, (const int*)flags, it, 1, 1))) return rc;
            }
            // z = M r ; beta ; rz
            if ((rc = mixed_vcycle(s, v, r, flexible, 0, !replace && !stdCycle, flags))) return rc;
        }
        double rn2Now = 0.0;
        T3D_HIP(hipMemcpyAsync(s->hostFlags, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipMemcpyAsync(&rn2Now, S + S_RN2, sizeof(double), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        // a non-finite residual, or one grown 1e6-fold from the start, cannot
        // recover (non-finite input, or an operator too inexact for the problem)
        if (!std::isfinite(rn2Now) || rn2Now > 1.0e12 * rn2First) {
            launch_fold(s, x, nullptr);
            T3D_HIP(hipStreamSynchronize(st));
            if (iters) *iters = it1;
            set_error("t3d_pcg_mixed: the residual diverged (non-finite or grown 1e6-fold) by iteration " +
                      std::to_string(it1));
            return T3D_ERR_DIVERGED;
        }
        if (s->timing) {   // launches of iterations that ran (after convergence the kernel returns at once)
            const int last = s->hostFlags[0] ? std::min(it1, s->hostFlags[1]) : it1;
            for (int it = it0; it < last; ++it) {
                float ms = 0.f;
                T3D_HIP(hipEventElapsedTime(&ms, s->tev[2 * (it - it0)], s->tev[2 * (it - it0) + 1]));
                s->timedMs += ms;
                ++s->timedLaunches;
            }
        }
        if (!s->hostFlags[0]) {
            // the next chunk ends two iterations after the convergence that the
            // mean decay so far predicts (iterations launched beyond it only
            // return at once), at least 4 and at most CHUNK iterations
            chunk = CHUNK;
            const double stop2 = stopHost * stopHost;
            if (rn2Now > 0.0 && rn2Now < rn2First && stop2 > 0.0 && stop2 < rn2Now) {
                const double perIt = std::log(rn2Now / rn2First) / double(it1);
                const double need = std::log(stop2 / rn2Now) / perIt;
                if (need < double(CHUNK)) chunk = std::max(4, int(std::ceil(need)) + 2);
            }
            it0 = it1;
            continue;
        }
        // the residual passed the test at iteration itDone: confirm with the
        // true residual b - K x (FP64); otherwise continue from there
        const int itDone = s->hostFlags[1] - 1;
        launch_fold(s, x, nullptr);
        if ((rc = launch_true_residual(s, b, x, nullptr, itDone, 0, 0))) return rc;
        double hs[2];
        T3D_HIP(hipMemcpyAsync(hs, S + S_STOP, sizeof(double), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipMemcpyAsync(hs + 1, S + S_RN2, sizeof(double), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        if (std::sqrt(hs[1]) < hs[0]) {
            result = itDone + 1;
            confirmed = true;
            break;
        }
        hipLaunchKernelGGL(k_mx_clear_done, dim3(1), dim3(1), 0, st, flags);
        if ((rc = mixed_vcycle(s, v, r, flexible, 0, false, flags))) return rc;
        it0 = itDone + 1;
    }
    if (!confirmed) launch_fold(s, x, nullptr);   // a confirmed solve has folded its last increment
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(st));
    if (iters) *iters = result;
    return T3D_OK;
}

}  // extern "C"
