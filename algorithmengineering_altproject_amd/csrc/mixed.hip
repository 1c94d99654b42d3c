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
//   * FP64 x, r, dot products, alpha and beta; beta either the reference's
//     (r.z)_new/(r.z)_old or the flexible (Polak-Ribiere) form
//     -alpha (z_new.Ap)/(r.z)_old, which tolerates the FP32 preconditioner
//     being only approximately linear;
//   * every `trueEvery` iterations the residual is recomputed as b - K x with
//     the FP64 operator (residual replacement; the direction is kept).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "t3d_internal.hpp"

namespace t3d {

// ---------------------------------------------------------------------------
// FP32 operator with fused direction update and p.Ap partials
// ---------------------------------------------------------------------------
namespace mx {
constexpr int TX = 16, TY = 8, TZ = 4;
constexpr int UX = TX + 2, UY = TY + 2, UZ = TZ + 2;
constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1;
constexpr int UN = UX * UY * UZ;
constexpr int EN = EX * EY * EZ;
}  // namespace mx

__device__ __forceinline__ int mx_corner(int dx, int dy, int dz) { return (dy ? (dx ? 1 : 0) : (dx ? 2 : 3)) + 4 * dz; }

// p_cur = z + beta*p_prev (z == nullptr: p_cur = p_prev); y = K p_cur;
// partial[block] = sum over the tile's nodes of p_cur.y (FP64)
__global__ void __launch_bounds__(512) k_kxu_f32_fused(Grid g, const float* __restrict__ Ke,
                                                       const float* __restrict__ Ecell,
                                                       const uint8_t* __restrict__ nodeFixed,
                                                       const float* __restrict__ z, const float* __restrict__ pprev,
                                                       const double* __restrict__ beta, float* __restrict__ pcur,
                                                       float* __restrict__ y, double* __restrict__ partial,
                                                       const int* __restrict__ done)
{
    using namespace mx;
    if (done && *done) return;
    __shared__ float sU[3 * UN];
    __shared__ float sE[EN];
    __shared__ double sRed[8];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * TZ;
    const int tid = threadIdx.x;
    const float b = (z && beta) ? float(*beta) : 0.0f;
    for (int i = tid; i < UN; i += 512) {
        const int lx = i % UX, ly = (i / UX) % UY, lz = i / (UX * UY);
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        float v0 = 0.f, v1 = 0.f, v2 = 0.f;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < g.nnx && Y < g.nny && Z < g.nnz) {
            const int64_t n = (int64_t(Z) * g.nny + Y) * g.nnx + X;
            v0 = pprev[3 * n]; v1 = pprev[3 * n + 1]; v2 = pprev[3 * n + 2];
            if (z) {
                v0 = fmaf(b, v0, z[3 * n]);
                v1 = fmaf(b, v1, z[3 * n + 1]);
                v2 = fmaf(b, v2, z[3 * n + 2]);
                // the tile interior owns the write of p_cur
                if (pcur && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && lz >= 1 && lz <= TZ) {
                    pcur[3 * n] = v0; pcur[3 * n + 1] = v1; pcur[3 * n + 2] = v2;
                }
            }
        }
        sU[i] = v0; sU[UN + i] = v1; sU[2 * UN + i] = v2;
    }
    for (int i = tid; i < EN; i += 512) {
        const int lx = i % EX, ly = (i / EX) % EY, lz = i / (EX * EY);
        const int X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
        float e = 0.f;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < g.resX && Y < g.resY && Z < g.resZ)
            e = Ecell[(int64_t(Z) * g.resY + Y) * g.resX + X];
        sE[i] = e;
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int px = w & 1, py = (w >> 1) & 1, pz = (w >> 2) & 1;
    const int lx = 2 * (lane & 7) + px, ly = 2 * ((lane >> 3) & 3) + py, lz = 2 * (lane >> 5) + pz;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 1
    for (int col = 0; col < 8; ++col) {
        const int dx = (px ^ col) & 1, dy = (py ^ (col >> 1) ^ 1) & 1, dz = (pz ^ (col >> 2)) & 1;
        const int cxl = lx - dx, cyl = ly - dy, czl = lz - dz;
        const float E = sE[((czl + 1) * EY + (cyl + 1)) * EX + (cxl + 1)];
        const int a = __builtin_amdgcn_readfirstlane(mx_corner(dx, dy, dz));
        const float* __restrict__ K0 = Ke + (3 * a) * 24;
        const int ub = ((czl + 1) * UY + (cyl + 1)) * UX + (cxl + 1);
        float f0 = 0.f, f1 = 0.f, f2 = 0.f;
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) {
            const int ox = (bb == 1 || bb == 2 || bb == 5 || bb == 6) ? 1 : 0;
            const int oy = ((bb & 3) <= 1) ? 1 : 0;
            const int oz = bb >> 2;
            const int idx = ub + (oz * UY + oy) * UX + ox;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float ue = sU[c * UN + idx];
                const int j = 3 * bb + c;
                f0 = fmaf(K0[j], ue, f0);
                f1 = fmaf(K0[24 + j], ue, f1);
                f2 = fmaf(K0[48 + j], ue, f2);
            }
        }
        const float Eu = (E <= 1.0e-16f) ? 0.f : E;
        a0 = fmaf(f0, Eu, a0);
        a1 = fmaf(f1, Eu, a1);
        a2 = fmaf(f2, Eu, a2);
    }
    const int X = x0 + lx, Y = y0 + ly, Z = z0 + lz;
    double dotp = 0.0;
    if (X < g.nnx && Y < g.nny && Z < g.nnz) {
        const int64_t n = (int64_t(Z) * g.nny + Y) * g.nnx + X;
        const unsigned fixed = nodeFixed[n];
        const float y0v = (fixed & 1u) ? 0.f : a0, y1v = (fixed & 2u) ? 0.f : a1, y2v = (fixed & 4u) ? 0.f : a2;
        y[3 * n] = y0v; y[3 * n + 1] = y1v; y[3 * n + 2] = y2v;
        const int ui = ((lz + 1) * UY + (ly + 1)) * UX + (lx + 1);
        dotp = double(sU[ui]) * y0v + double(sU[UN + ui]) * y1v + double(sU[2 * UN + ui]) * y2v;
    }
    if (partial) {
        for (int o = 32; o >= 1; o >>= 1) dotp += __shfl_xor(dotp, o, 64);
        if (lane == 0) sRed[w] = dotp;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int k = 0; k < 8; ++k) s += sRed[k];
            partial[(int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// deterministic block sums for the mixed solver (fixed tree per block, blocks
// reduced in index order)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum_256(double s, double* sh)
{
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < int(blockDim.x >> 6); ++k) t += sh[k];
    return t;   // valid in thread 0
}

// x += alpha p32 ; r -= alpha Ap32 ; partial r.r
__global__ void __launch_bounds__(256) k_mx_update_xr(int64_t n, const double* __restrict__ rz,
                                                      const double* __restrict__ denom,
                                                      const float* __restrict__ p, const float* __restrict__ Ap,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ alphaOut, double* __restrict__ partial,
                                                      const int* __restrict__ done)
{
    if (*done) return;
    __shared__ double sh[4];
    const double d = *denom;
    const double alpha = *rz / (1.0e-30 < d ? d : 1.0e-30);
    if (blockIdx.x == 0 && threadIdx.x == 0) *alphaOut = alpha;
    double s = 0.0;
    const int64_t base = int64_t(blockIdx.x) * blockDim.x * 4 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * blockDim.x;
        if (i < n) {
            x[i] = fma(alpha, double(p[i]), x[i]);
            const double v = fma(-alpha, double(Ap[i]), r[i]);
            r[i] = v;
            s = fma(v, v, s);
        }
    }
    const double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) k_mx_sum_partials(int64_t m, const double* __restrict__ in,
                                                         double* __restrict__ out, const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) s += in[i];
    const double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) *out = t;
}

// r = b - Ap (FP64) ; partial r.r
__global__ void __launch_bounds__(256) k_mx_true_residual(int64_t n, const double* __restrict__ b,
                                                          const double* __restrict__ Ap, double* __restrict__ r,
                                                          double* __restrict__ partial, const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[4];
    double s = 0.0;
    const int64_t base = int64_t(blockIdx.x) * blockDim.x * 4 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * blockDim.x;
        if (i < n) {
            const double v = b[i] - Ap[i];
            r[i] = v;
            s = fma(v, v, s);
        }
    }
    const double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// FP32 V-cycle kernels
// ---------------------------------------------------------------------------
struct LevelF {
    int resX, resY, resZ;
    int nnx, nny, nnz;
    int64_t nN;
};

template <class TIn>
__global__ void __launch_bounds__(256) k_restrict_f32(LevelF Lf, LevelF Lc, int span, const TIn* __restrict__ rf,
                                                      const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                                      const float* __restrict__ Dc, float* __restrict__ xc,
                                                      double* __restrict__ rcd, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lc.nN) return;
    const int cx = int(n % Lc.nnx), cy = int((n / Lc.nnx) % Lc.nny), cz = int(n / (int64_t(Lc.nnx) * Lc.nny));
    const int wSize = 2 * span - 1;
    const float inv = 1.0f / float(span);
    const int fx0 = cx * span, fy0 = cy * span, fz0 = cz * span;
    const int iz0 = (fz0 - span + 1 < 0) ? (span - 1 - fz0) : 0;
    const int iz1 = (fz0 + span - 1 >= Lf.nnz) ? (wSize - (fz0 + span - Lf.nnz)) : wSize;
    const int iy0 = (fy0 - span + 1 < 0) ? (span - 1 - fy0) : 0;
    const int iy1 = (fy0 + span - 1 >= Lf.nny) ? (wSize - (fy0 + span - Lf.nny)) : wSize;
    const int ix0 = (fx0 - span + 1 < 0) ? (span - 1 - fx0) : 0;
    const int ix1 = (fx0 + span - 1 >= Lf.nnx) ? (wSize - (fx0 + span - Lf.nnx)) : wSize;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, ws = 0.f;
    for (int iz = iz0; iz < iz1; ++iz) {
        const int dz = iz - (span - 1);
        const float wz = 1.0f - float(dz < 0 ? -dz : dz) * inv;
        const int64_t zo = int64_t(fz0 + dz) * Lf.nny * Lf.nnx;
        for (int iy = iy0; iy < iy1; ++iy) {
            const int dy = iy - (span - 1);
            const float wzy = wz * (1.0f - float(dy < 0 ? -dy : dy) * inv);
            const int64_t yo = zo + int64_t(fy0 + dy) * Lf.nnx;
            for (int ix = ix0; ix < ix1; ++ix) {
                const int dx = ix - (span - 1);
                const float w = wzy * (1.0f - float(dx < 0 ? -dx : dx) * inv);
                const int64_t fi = 3 * (yo + (fx0 + dx));
                v0 = fmaf(float(rf[fi]), w, v0);
                v1 = fmaf(float(rf[fi + 1]), w, v1);
                v2 = fmaf(float(rf[fi + 2]), w, v2);
                ws += w;
            }
        }
    }
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    if (ws > 1e-12f) {
        const float iw = 1.0f / ws;
        r0 = v0 * iw; r1 = v1 * iw; r2 = v2 * iw;
    }
    const unsigned m = cmask[n];
    if (m & 1u) r0 = 0.f;
    if (m & 2u) r1 = 0.f;
    if (m & 4u) r2 = 0.f;
    rc[3 * n] = r0; rc[3 * n + 1] = r1; rc[3 * n + 2] = r2;
    if (rcd) {   // FP64 copy for the coarsest Cholesky solve
        rcd[3 * n] = r0; rcd[3 * n + 1] = r1; rcd[3 * n + 2] = r2;
    }
    if (Dc) {
        xc[3 * n] = (m & 1u) ? 0.f : r0 / fmaxf(1.0e-30f, Dc[3 * n]);
        xc[3 * n + 1] = (m & 2u) ? 0.f : r1 / fmaxf(1.0e-30f, Dc[3 * n + 1]);
        xc[3 * n + 2] = (m & 4u) ? 0.f : r2 / fmaxf(1.0e-30f, Dc[3 * n + 2]);
    }
}

// x_f = mask ? 0 : 2t + P(x_c), t = w r/D ; at the finest level also the
// partials of r.z and z.Ap (FP64)
template <class TIn, bool FINEST>
__global__ void __launch_bounds__(256) k_prolong_smooth_f32(LevelF Lf, LevelF Lc, int span,
                                                            const float* __restrict__ xc, const TIn* __restrict__ rf,
                                                            const float* __restrict__ Df,
                                                            const uint8_t* __restrict__ fmask, float weight,
                                                            float* __restrict__ xf, const float* __restrict__ Ap,
                                                            double* __restrict__ partRZ, double* __restrict__ partZAp,
                                                            const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[4];
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double srz = 0.0, szap = 0.0;
    if (n < Lf.nN) {
        const int fx = int(n % Lf.nnx), fy = int((n / Lf.nnx) % Lf.nny), fz = int(n / (int64_t(Lf.nnx) * Lf.nny));
        int cx = fx / span, cy = fy / span, cz = fz / span;
        int rx = fx % span, ry = fy % span, rz = fz % span;
        if (cx >= Lc.resX) { cx = Lc.resX - 1; rx = span; }
        if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = span; }
        if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = span; }
        const float inv = 1.0f / float(span);
        const float wx1 = float(rx) * inv, wx0 = 1.f - wx1;
        const float wy1 = float(ry) * inv, wy0 = 1.f - wy1;
        const float wz1 = float(rz) * inv, wz0 = 1.f - wz1;
        const int64_t s = int64_t(Lc.nnx) * Lc.nny;
        const int64_t b00 = s * cz + int64_t(Lc.nnx) * cy + cx;
        const int64_t b10 = b00 + Lc.nnx, n00 = b00 + s, n10 = n00 + Lc.nnx;
        const unsigned m = fmask[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float lo = wy0 * (wx0 * xc[3 * b00 + c] + wx1 * xc[3 * b00 + 3 + c]) +
                             wy1 * (wx0 * xc[3 * b10 + c] + wx1 * xc[3 * b10 + 3 + c]);
            const float hi = wy0 * (wx0 * xc[3 * n00 + c] + wx1 * xc[3 * n00 + 3 + c]) +
                             wy1 * (wx0 * xc[3 * n10 + c] + wx1 * xc[3 * n10 + 3 + c]);
            const float v = wz0 * lo + wz1 * hi;
            float out = 0.f;
            if (!((m >> c) & 1u)) {
                const float t = weight * float(rf[3 * n + c]) / fmaxf(1.0e-30f, Df[3 * n + c]);
                out = 2.f * t + v;
            }
            xf[3 * n + c] = out;
            if (FINEST) {
                srz = fma(double(rf[3 * n + c]), double(out), srz);
                szap = fma(double(out), double(Ap[3 * n + c]), szap);
            }
        }
    }
    if (FINEST) {
        const double a = block_sum_256(srz, sh);
        __syncthreads();
        const double b = block_sum_256(szap, sh);
        if (threadIdx.x == 0) {
            partRZ[blockIdx.x] = a;
            partZAp[blockIdx.x] = b;
        }
    }
}

__global__ void k_d2f(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

// flexible: beta = -alpha (z.Ap) / max(1e-30, rz_old) (Polak-Ribiere);
// standard: beta = rz_new / max(1e-30, rz_old)
__global__ void k_mx_beta(const double* __restrict__ alpha, const double* __restrict__ zAp,
                          const double* __restrict__ rzNew, const double* __restrict__ rzOld, int flexible,
                          double* __restrict__ beta, const int* __restrict__ done)
{
    if (*done) return;
    const double o = *rzOld;
    const double num = flexible ? -(*alpha) * (*zAp) : *rzNew;
    *beta = num / (1.0e-30 < o ? o : 1.0e-30);
}

__global__ void k_mx_check(const double* __restrict__ rn2, const double* __restrict__ stop, int it,
                           int* __restrict__ flags)
{
    if (flags[0]) return;
    if (sqrt(*rn2) < *stop) {
        flags[0] = 1;
        flags[1] = it + 1;
    }
}

__global__ void k_mx_set_int(int* __restrict__ p, int v) { *p = v; }
__global__ void k_mx_set_double(double* __restrict__ p, double v) { *p = v; }

__global__ void k_pcg_stop_mixed(const double* __restrict__ normb2, double tol, double* __restrict__ stop)
{
    const double nb = sqrt(*normb2);
    *stop = (1.0e-30 < nb ? nb : 1.0e-30) * tol;
}

}  // namespace t3d

using namespace t3d;

// mixed-precision state attached to a multigrid object
struct t3d_mixed {
    t3d_problem* p = nullptr;
    t3d_mg* M = nullptr;
    float* Ke32 = nullptr;
    float* D32[8] = {};
    float* r32[8] = {};
    float* x32[8] = {};
    double* rc64 = nullptr;    // coarsest residual / solution in FP64 for Cholesky
    double* xc64 = nullptr;
    float *pA = nullptr, *pB = nullptr, *Ap = nullptr, *z = nullptr;
    double* part = nullptr;    // partial sums
    double* part2 = nullptr;
    double* sc = nullptr;      // scalars
    int* iflags = nullptr;     // [0] done [1] iters [2] restart
    int hostFlags[4];
    int64_t nTilesK = 0, nBlk = 0;
    int nlev = 0;
};

// internal accessors of the multigrid object (mg.hip)
namespace t3d {
struct MGView {
    int nlev;
    int span[8];
    int resX[8], resY[8], resZ[8];
    int64_t nN[8];
    const uint8_t* mask[8];
    const double* D[8];
    const double* L;
    const double* LT;
    int Nc;
    double weight;
};
int mg_view(const t3d_mg* g, MGView& v);
__global__ void k_chol_solve(int N, const double* __restrict__ L, const double* __restrict__ LT,
                             const double* __restrict__ b, const uint8_t* __restrict__ mask, double* __restrict__ x,
                             const int* __restrict__ done);
}  // namespace t3d

static LevelF levf(const MGView& v, int l)
{
    LevelF L;
    L.resX = v.resX[l]; L.resY = v.resY[l]; L.resZ = v.resZ[l];
    L.nnx = v.resX[l] + 1; L.nny = v.resY[l] + 1; L.nnz = v.resZ[l] + 1;
    L.nN = v.nN[l];
    return L;
}

extern "C" {

int t3d_mixed_destroy(t3d_mixed* s)
{
    if (!s) return T3D_OK;
    void* b[] = {s->Ke32, s->rc64, s->xc64, s->pA, s->pB, s->Ap, s->z, s->part, s->part2, s->sc, s->iflags};
    for (void* q : b)
        if (q) (void)hipFree(q);
    for (int l = 0; l < 8; ++l) {
        if (s->D32[l]) (void)hipFree(s->D32[l]);
        if (s->r32[l]) (void)hipFree(s->r32[l]);
        if (s->x32[l]) (void)hipFree(s->x32[l]);
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
    const int64_t nD = 3 * p->g.nN;
    const Grid& g = p->g;
    s->nTilesK = int64_t((g.nnx + mx::TX - 1) / mx::TX) * ((g.nny + mx::TY - 1) / mx::TY) *
                 ((g.nnz + mx::TZ - 1) / mx::TZ);
    s->nBlk = (nD + 1023) / 1024;
    const int64_t npart = std::max<int64_t>(std::max<int64_t>(s->nTilesK, s->nBlk), (g.nN + 255) / 256) + 16;
    if (hipMalloc(&s->Ke32, sizeof(float) * 576) != hipSuccess ||
        hipMemcpy(s->Ke32, K32, sizeof(float) * 576, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->pA, sizeof(float) * nD) != hipSuccess || hipMalloc(&s->pB, sizeof(float) * nD) != hipSuccess ||
        hipMalloc(&s->Ap, sizeof(float) * nD) != hipSuccess || hipMalloc(&s->z, sizeof(float) * nD) != hipSuccess ||
        hipMalloc(&s->part, sizeof(double) * npart) != hipSuccess ||
        hipMalloc(&s->part2, sizeof(double) * npart) != hipSuccess ||
        hipMalloc(&s->sc, sizeof(double) * 64) != hipSuccess || hipMalloc(&s->iflags, sizeof(int) * 4) != hipSuccess)
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
    *out = s;
    return T3D_OK;
}

// refresh the FP32 diagonals after t3d_mg_setup
int t3d_mixed_setup(t3d_mixed* s)
{
    T3D_CHECK(s, T3D_ERR_ARG, "t3d_mixed_setup: null");
    MGView v;
    int rc = mg_view(s->M, v);
    if (rc) return rc;
    for (int l = 0; l < v.nlev; ++l) {
        const int64_t n = 3 * v.nN[l];
        hipLaunchKernelGGL(k_d2f, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s->p->stream, n, v.D[l], s->D32[l]);
        T3D_HIP(hipGetLastError());
    }
    T3D_HIP(hipStreamSynchronize(s->p->stream));
    return T3D_OK;
}

}  // extern "C"

namespace t3d {

// z32 = M32(r), with r.z and z.Ap partials at the finest level
static int mixed_vcycle(t3d_mixed* s, const MGView& v, const double* r, const float* Ap, double* rzOut,
                        double* zApOut, const int* done)
{
    t3d_problem* p = s->p;
    hipStream_t st = p->stream;
    const int L = v.nlev - 1;
    const int B = 256;
    const bool chol = v.Nc > 0;
    T3D_CHECK(L >= 1, T3D_ERR_STATE, "mixed V-cycle needs at least two levels");
    for (int l = 0; l < L; ++l) {
        const bool last = (l + 1 == L);
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        const unsigned nb = unsigned((Lc.nN + B - 1) / B);
        const float* Dc = (last && !chol) ? s->D32[L] : nullptr;
        double* rcd = (last && chol) ? s->rc64 : nullptr;
        if (l == 0)
            hipLaunchKernelGGL(k_restrict_f32<double>, dim3(nb), dim3(B), 0, st, Lf, Lc, v.span[1], r, v.mask[1],
                               s->r32[1], Dc, s->x32[L], rcd, done);
        else
            hipLaunchKernelGGL(k_restrict_f32<float>, dim3(nb), dim3(B), 0, st, Lf, Lc, v.span[l + 1],
                               (const float*)s->r32[l], v.mask[l + 1], s->r32[l + 1], Dc, s->x32[L], rcd, done);
        T3D_HIP(hipGetLastError());
    }
    if (chol) {
        hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), 0, st, v.Nc, v.L, v.LT, s->rc64, v.mask[L], s->xc64,
                           done);
        const int64_t nc = 3 * v.nN[L];
        hipLaunchKernelGGL(k_d2f, dim3(unsigned((nc + 255) / 256)), dim3(256), 0, st, nc, s->xc64, s->x32[L]);
        T3D_HIP(hipGetLastError());
    }
    const float w = float(v.weight);
    for (int l = L - 1; l >= 0; --l) {
        const LevelF Lf = levf(v, l), Lc = levf(v, l + 1);
        const unsigned nb = unsigned((Lf.nN + B - 1) / B);
        if (l == 0) {
            hipLaunchKernelGGL((k_prolong_smooth_f32<double, true>), dim3(nb), dim3(B), 0, st, Lf, Lc, v.span[1],
                               s->x32[1], r, s->D32[0], v.mask[0], w, s->z, Ap, s->part, s->part2, done);
            T3D_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nb), s->part, rzOut, done);
            hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nb), s->part2, zApOut, done);
        } else {
            hipLaunchKernelGGL((k_prolong_smooth_f32<float, false>), dim3(nb), dim3(B), 0, st, Lf, Lc,
                               v.span[l + 1], s->x32[l + 1], (const float*)s->r32[l], s->D32[l], v.mask[l], w,
                               s->x32[l], nullptr, nullptr, nullptr, done);
        }
        T3D_HIP(hipGetLastError());
    }
    return T3D_OK;
}

// end of an iteration: z = M r (rz_new -> S[9], z.Ap -> S[4]); beta -> S[7]; rz = rz_new
static int mixed_tail(t3d_mixed* s, const MGView& v, const double* r, double* S, int flexible, const int* flags)
{
    hipStream_t st = s->p->stream;
    int rc = mixed_vcycle(s, v, r, s->Ap, S + 9, S + 4, flags);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mx_beta, dim3(1), dim3(1), 0, st, S + 6, S + 4, S + 9, S + 2, flexible, S + 7, flags);
    hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, st, S + 9, S + 2, flags);
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

}  // namespace t3d

extern "C" {

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
    const unsigned nblk = unsigned((n + 1023) / 1024);
    double* r = nullptr;
    double* Ap64 = nullptr;
    // FP64 work vectors of the problem (shared with the plain PCG)
    if (!p->wr) {
        T3D_HIP(hipMalloc(&p->wr, sizeof(double) * n));
        T3D_HIP(hipMalloc(&p->wz, sizeof(double) * n));
        T3D_HIP(hipMalloc(&p->wp, sizeof(double) * n));
        T3D_HIP(hipMalloc(&p->wAp, sizeof(double) * n));
        T3D_HIP(hipMalloc(&p->pcgFlags, sizeof(int) * 4));
        T3D_HIP(hipHostMalloc(&p->hostFlags, sizeof(int) * 4, hipHostMallocDefault));
    }
    r = p->wr;
    Ap64 = p->wAp;
    double* S = s->sc;   // [0] normb2 [1] stop [2] rz [3] rzNew [4] zAp [5] denom [6] alpha [7] beta [8] rn2
    int* flags = s->iflags;
    dim3 kgrid(unsigned((g.nnx + mx::TX - 1) / mx::TX), unsigned((g.nny + mx::TY - 1) / mx::TY),
               unsigned((g.nnz + mx::TZ - 1) / mx::TZ));
    const int64_t ntk = s->nTilesK;
    // r = b - K64 x ; rn2 ; z = M r ; rz
    if ((rc = launch_kxu_f64(p, x, Ap64))) return rc;
    hipLaunchKernelGGL(k_mx_set_int, dim3(1), dim3(1), 0, st, flags, 0);
    hipLaunchKernelGGL(k_mx_set_int, dim3(1), dim3(1), 0, st, flags + 1, -1);
    hipLaunchKernelGGL(k_mx_true_residual, dim3(nblk), dim3(256), 0, st, n, b, Ap64, r, s->part, flags);
    hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nblk), s->part, S + 8, flags);
    T3D_HIP(hipMemsetAsync(s->Ap, 0, sizeof(float) * n, st));
    if ((rc = mixed_vcycle(s, v, r, s->Ap, S + 2, S + 4, flags))) return rc;
    if ((rc = canon_dot_dev(p, b, b, n, S + 0))) return rc;
    hipLaunchKernelGGL(k_pcg_stop_mixed, dim3(1), dim3(1), 0, st, S + 0, tol, S + 1);
    T3D_HIP(hipMemsetAsync(s->pA, 0, sizeof(float) * n, st));
    hipLaunchKernelGGL(k_mx_set_double, dim3(1), dim3(1), 0, st, S + 7, 0.0);   // beta_0 = 0
    // iteration it reads p_prev and writes p_cur; the two buffers alternate
    auto pPrevOf = [&](int it) { return (it & 1) ? s->pB : s->pA; };
    auto pCurOf = [&](int it) { return (it & 1) ? s->pA : s->pB; };
    const int CHUNK = 32;
    int result = maxIt;
    int it0 = 0;
    while (it0 < maxIt) {
        const int it1 = std::min(maxIt, it0 + CHUNK);
        for (int it = it0; it < it1; ++it) {
            // p_cur = z + beta p_prev ; Ap = K32 p_cur ; denom = p_cur.Ap
            hipLaunchKernelGGL(k_kxu_f32_fused, kgrid, dim3(512), 0, st, g, s->Ke32, p->EcellF, p->nodeFixed,
                               s->z, pPrevOf(it), S + 7, pCurOf(it), s->Ap, s->part, flags);
            hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, ntk, s->part, S + 5, flags);
            // alpha = rz / denom ; x += alpha p ; r -= alpha Ap ; rn2
            hipLaunchKernelGGL(k_mx_update_xr, dim3(nblk), dim3(256), 0, st, n, S + 2, S + 5, pCurOf(it), s->Ap, x,
                               r, S + 6, s->part, flags);
            hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nblk), s->part, S + 8, flags);
            if (trueEvery > 0 && ((it + 1) % trueEvery) == 0) {
                // residual replacement with the FP64 operator (the direction is kept)
                if ((rc = launch_kxu_f64(p, x, Ap64, flags))) return rc;
                hipLaunchKernelGGL(k_mx_true_residual, dim3(nblk), dim3(256), 0, st, n, b, Ap64, r, s->part, flags);
                hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nblk), s->part, S + 8,
                                   flags);
            }
            hipLaunchKernelGGL(k_mx_check, dim3(1), dim3(1), 0, st, S + 8, S + 1, it, flags);
            if ((rc = mixed_tail(s, v, r, S, flexible, flags))) return rc;
        }
        T3D_HIP(hipMemcpyAsync(s->hostFlags, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        if (!s->hostFlags[0]) {
            it0 = it1;
            continue;
        }
        // the recursive residual passed the test at iteration itDone: confirm
        // with the true residual b - K x (FP64); otherwise continue from it
        const int itDone = s->hostFlags[1] - 1;
        if ((rc = launch_kxu_f64(p, x, Ap64))) return rc;
        hipLaunchKernelGGL(k_mx_true_residual, dim3(nblk), dim3(256), 0, st, n, b, Ap64, r, s->part, nullptr);
        hipLaunchKernelGGL(k_mx_sum_partials, dim3(1), dim3(256), 0, st, int64_t(nblk), s->part, S + 8, nullptr);
        double hs[2];
        T3D_HIP(hipMemcpyAsync(hs, S + 1, sizeof(double), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipMemcpyAsync(hs + 1, S + 8, sizeof(double), hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        if (std::sqrt(hs[1]) < hs[0]) {
            result = itDone + 1;
            break;
        }
        hipLaunchKernelGGL(k_mx_set_int, dim3(1), dim3(1), 0, st, flags, 0);
        hipLaunchKernelGGL(k_mx_set_int, dim3(1), dim3(1), 0, st, flags + 1, -1);
        if ((rc = mixed_tail(s, v, r, S, flexible, flags))) return rc;
        it0 = itDone + 1;
    }
    T3D_HIP(hipStreamSynchronize(st));
    if (iters) *iters = result;
    return T3D_OK;
}

}  // extern "C"
