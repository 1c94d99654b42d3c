// mg.hip -- geometric multigrid preconditioner on the device.
//
// Reference: make_diagonal_preconditioner_from_static and its V-cycle closure
// (cpp/src/multigrid/preconditioner.cpp:38-266), MG_BuildDiagonals
// (cpp/src/multigrid/diagonal.cpp:11-160), MG_Restrict_nodes_Vec3_Inner /
// MG_Prolongate_nodes_Vec3_Inner (cpp/src/multigrid/transfers.cpp:378-533),
// chol_solve_lower (cpp/src/multigrid/coarsest.cpp:92-122).
//
// All vectors use the full node-major AoS layout of their level.  Every
// accumulation keeps the reference's sequential order, so results are
// bit-identical to the reference's arithmetic:
//   * finest diagonal: per DOF, incident cells in colour order 0..7;
//   * coarse diagonals: per coarse DOF, incident coarse cells in colour order,
//     each contributing its (span+1)^3 embedded fine nodes in (iz,iy,ix) order;
//   * restriction: per coarse node, the clamped (2s-1)^3 stencil in (iz,iy,ix)
//     order, scaled by 1/wSum;
//   * prolongation: 8 trilinear terms in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mg_host.hpp"
#include "t3d_internal.hpp"
#include "chol.hpp"
#include "coarse.hpp"

namespace t3d {

struct LevelDev {
    int resX, resY, resZ;   // cells
    int nnx, nny, nnz;      // nodes
    int64_t nN;
};

static LevelDev level_dev(const MGLevelDims& d)
{
    LevelDev L;
    L.resX = d.resX; L.resY = d.resY; L.resZ = d.resZ;
    L.nnx = d.resX + 1; L.nny = d.resY + 1; L.nnz = d.resZ + 1;
    L.nN = d.nN;
    return L;
}

// local index a of the corner at cell-offset (dx,dy,dz); dy = 1 is the upper node row
__device__ __forceinline__ int corner_index(int dx, int dy, int dz)
{
    return (dy ? (dx ? 1 : 0) : (dx ? 2 : 3)) + 4 * dz;
}

// ---------------------------------------------------------------------------
// finest diagonal (diagonal.cpp:11-63 + BC at preconditioner level, :147-152)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_diag_finest(Grid g, const double* __restrict__ Ke,
                                                     const double* __restrict__ Ecell,
                                                     const uint8_t* __restrict__ fixedBits,
                                                     double* __restrict__ D)
{
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= g.nN) return;
    const int ix = int(n % g.nnx), iy = int((n / g.nnx) % g.nny), iz = int(n / (int64_t(g.nnx) * g.nny));
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int col = 0; col < 8; ++col) {
        const int dx = (ix ^ col) & 1, dy = (iy ^ (col >> 1) ^ 1) & 1, dz = (iz ^ (col >> 2)) & 1;
        const int cx = ix - dx, ty = iy - dy, cz = iz - dz;
        if (cx < 0 || ty < 0 || cz < 0 || cx >= g.resX || ty >= g.resY || cz >= g.resZ) continue;
        const double E = Ecell[(int64_t(cz) * g.resY + ty) * g.resX + cx];
        const int a = corner_index(dx, dy, dz);
        const int l = 3 * a;
        // void cells hold E = 0 and add +0.0, which leaves the sum unchanged
        d0 = d0 + Ke[(l + 0) * 25] * E;
        d1 = d1 + Ke[(l + 1) * 25] * E;
        d2 = d2 + Ke[(l + 2) * 25] * E;
    }
    const unsigned f = fixedBits[n];
    d0 = (d0 > 0.0) ? d0 : 1.0;
    d1 = (d1 > 0.0) ? d1 : 1.0;
    d2 = (d2 > 0.0) ? d2 : 1.0;
    D[3 * n + 0] = (f & 1u) ? 1.0 : d0;
    D[3 * n + 1] = (f & 2u) ? 1.0 : d1;
    D[3 * n + 2] = (f & 4u) ? 1.0 : d2;
}

// ---------------------------------------------------------------------------
// coarse diagonal by W^2-weighted averaging (diagonal.cpp:66-138)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_diag_coarsen(LevelDev Lf, LevelDev Lc, int span,
                                                      const double* __restrict__ W,
                                                      const double* __restrict__ Df,
                                                      const uint8_t* __restrict__ fmask,
                                                      const uint8_t* __restrict__ cmask,
                                                      double* __restrict__ Dc)
{
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lc.nN) return;
    const int X = int(n % Lc.nnx), Y = int((n / Lc.nnx) % Lc.nny), Z = int(n / (int64_t(Lc.nnx) * Lc.nny));
    const unsigned cm = cmask[n];
    const int grid = span + 1;
    double acc[3] = {0.0, 0.0, 0.0}, ws[3] = {0.0, 0.0, 0.0};
    for (int col = 0; col < 8; ++col) {
        const int cex = X - ((X ^ col) & 1);
        const int cez = Z - ((Z ^ (col >> 2)) & 1);
        // element rows cey = resY-1-cty with cty in {Y-1, Y}
        int cey = Lc.resY - Y;                    // cty = Y-1
        if ((cey & 1) != ((col >> 1) & 1)) cey = Lc.resY - 1 - Y;   // cty = Y
        if (cex < 0 || cez < 0 || cey < 0 || cex >= Lc.resX || cez >= Lc.resZ || cey >= Lc.resY) continue;
        const int cty = Lc.resY - 1 - cey;
        const int a = corner_index(X - cex, Y - cty, Z - cez);
        const int fx0 = cex * span, fy0 = (Lc.resY - cey) * span, fz0 = cez * span;
        for (int iz = 0; iz <= span; ++iz)
            for (int iy = 0; iy <= span; ++iy)
                for (int ix = 0; ix <= span; ++ix) {
                    const int64_t fn = (int64_t(fz0 + iz) * Lf.nny + (fy0 - iy)) * Lf.nnx + (fx0 + ix);
                    const double w = W[((iz * grid + iy) * grid + ix) * 8 + a];
                    const double w2 = w * w;
                    const unsigned fm = fmask[fn];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        if ((cm >> c) & 1u) continue;
                        if ((fm >> c) & 1u) continue;
                        acc[c] = acc[c] + w2 * Df[3 * fn + c];
                        ws[c] = ws[c] + w2;
                    }
                }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double v;
        if ((cm >> c) & 1u) v = 1.0;
        else if (ws[c] > 0) {
            v = acc[c] / ws[c];
            if (!(v > 0.0)) v = 1.0;
        } else v = 1.0;
        Dc[3 * n + c] = v;
    }
}

// ---------------------------------------------------------------------------
// restriction fine -> coarse, coarse mask applied; when Dc != nullptr the
// level is the coarsest one and the diagonal solve x = r / max(1e-30, D) is
// fused (preconditioner.cpp:195-227)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double hat_w(int i, int span, double inv)
{
    const int d = i - (span - 1);
    return 1.0 - double(d < 0 ? -d : d) * inv;
}

// norm = 0 (Galerkin form): the plain transpose of the prolongation, sum w r
__global__ void __launch_bounds__(256) k_restrict(LevelDev Lf, LevelDev Lc, int span,
                                                  const double* __restrict__ rf, const uint8_t* __restrict__ cmask,
                                                  double* __restrict__ rc, const double* __restrict__ Dc,
                                                  double* __restrict__ xc, const int* __restrict__ done, int norm)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lc.nN) return;
    const int cx = int(n % Lc.nnx), cy = int((n / Lc.nnx) % Lc.nny), cz = int(n / (int64_t(Lc.nnx) * Lc.nny));
    const int wSize = 2 * span - 1;
    const double inv = 1.0 / span;
    const int fx0 = cx * span, fy0 = cy * span, fz0 = cz * span;
    const int iz0 = (fz0 - span + 1 < 0) ? (span - 1 - fz0) : 0;
    const int iz1 = (fz0 + span - 1 >= Lf.nnz) ? (wSize - (fz0 + span - Lf.nnz)) : wSize;
    const int iy0 = (fy0 - span + 1 < 0) ? (span - 1 - fy0) : 0;
    const int iy1 = (fy0 + span - 1 >= Lf.nny) ? (wSize - (fy0 + span - Lf.nny)) : wSize;
    const int ix0 = (fx0 - span + 1 < 0) ? (span - 1 - fx0) : 0;
    const int ix1 = (fx0 + span - 1 >= Lf.nnx) ? (wSize - (fx0 + span - Lf.nnx)) : wSize;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, wsum = 0.0;
    for (int iz = iz0; iz < iz1; ++iz) {
        const double wz = hat_w(iz, span, inv);
        const int64_t zo = int64_t(fz0 + iz - (span - 1)) * Lf.nny * Lf.nnx;
        for (int iy = iy0; iy < iy1; ++iy) {
            const double wzy = wz * hat_w(iy, span, inv);
            const int64_t yo = zo + int64_t(fy0 + iy - (span - 1)) * Lf.nnx;
            for (int ix = ix0; ix < ix1; ++ix) {
                const double w = wzy * hat_w(ix, span, inv);
                const int64_t fi = 3 * (yo + (fx0 + ix - (span - 1)));
                v0 = v0 + rf[fi + 0] * w;
                v1 = v1 + rf[fi + 1] * w;
                v2 = v2 + rf[fi + 2] * w;
                wsum = wsum + w;
            }
        }
    }
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (!norm) {
        r0 = v0; r1 = v1; r2 = v2;
    } else if (wsum > 1e-12) {
        const double invW = 1.0 / wsum;
        r0 = v0 * invW; r1 = v1 * invW; r2 = v2 * invW;
    }
    const unsigned m = cmask[n];
    if (m & 1u) r0 = 0.0;
    if (m & 2u) r1 = 0.0;
    if (m & 4u) r2 = 0.0;
    rc[3 * n + 0] = r0; rc[3 * n + 1] = r1; rc[3 * n + 2] = r2;
    if (Dc) {
        const double a0 = Dc[3 * n + 0], a1 = Dc[3 * n + 1], a2 = Dc[3 * n + 2];
        xc[3 * n + 0] = (m & 1u) ? 0.0 : r0 / (1.0e-30 < a0 ? a0 : 1.0e-30);
        xc[3 * n + 1] = (m & 2u) ? 0.0 : r1 / (1.0e-30 < a1 ? a1 : 1.0e-30);
        xc[3 * n + 2] = (m & 4u) ? 0.0 : r2 / (1.0e-30 < a2 ? a2 : 1.0e-30);
    }
}

// The same restriction with one thread per coarse DOF (node n = t / 3,
// component c = t % 3) for the spans the hierarchy uses: neighbouring lanes
// read neighbouring doubles, and the loads of a fine row (2 SPAN - 1 nodes)
// are issued together before its terms are added.  Every coarse value and the
// weight sum are the same sequence of additions as in k_restrict.
template <int SPAN>
__global__ void __launch_bounds__(256) k_restrict_c(LevelDev Lf, LevelDev Lc, const double* __restrict__ rf,
                                                    const uint8_t* __restrict__ cmask, double* __restrict__ rc,
                                                    const double* __restrict__ Dc, double* __restrict__ xc,
                                                    const int* __restrict__ done, int norm)
{
    if (done && *done) return;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * Lc.nN) return;
    const int64_t n = t / 3;
    const int c = int(t - 3 * n);
    const int cx = int(n % Lc.nnx), cy = int((n / Lc.nnx) % Lc.nny), cz = int(n / (int64_t(Lc.nnx) * Lc.nny));
    constexpr int W = 2 * SPAN - 1;
    const double inv = 1.0 / SPAN;
    const int fx0 = cx * SPAN, fy0 = cy * SPAN, fz0 = cz * SPAN;
    const int iz0 = (fz0 - SPAN + 1 < 0) ? (SPAN - 1 - fz0) : 0;
    const int iz1 = (fz0 + SPAN - 1 >= Lf.nnz) ? (W - (fz0 + SPAN - Lf.nnz)) : W;
    const int iy0 = (fy0 - SPAN + 1 < 0) ? (SPAN - 1 - fy0) : 0;
    const int iy1 = (fy0 + SPAN - 1 >= Lf.nny) ? (W - (fy0 + SPAN - Lf.nny)) : W;
    const int ix0 = (fx0 - SPAN + 1 < 0) ? (SPAN - 1 - fx0) : 0;
    const int ix1 = (fx0 + SPAN - 1 >= Lf.nnx) ? (W - (fx0 + SPAN - Lf.nnx)) : W;
    double v = 0.0, wsum = 0.0;
    for (int iz = iz0; iz < iz1; ++iz) {
        const double wz = hat_w(iz, SPAN, inv);
        const int64_t zo = int64_t(fz0 + iz - (SPAN - 1)) * Lf.nny * Lf.nnx;
        for (int iy = iy0; iy < iy1; ++iy) {
            const double wzy = wz * hat_w(iy, SPAN, inv);
            const int64_t yo = zo + int64_t(fy0 + iy - (SPAN - 1)) * Lf.nnx;
            double f[W];
#pragma unroll
            for (int ix = 0; ix < W; ++ix)
                f[ix] = (ix >= ix0 && ix < ix1) ? rf[3 * (yo + (fx0 + ix - (SPAN - 1))) + c] : 0.0;
#pragma unroll
            for (int ix = 0; ix < W; ++ix)
                if (ix >= ix0 && ix < ix1) {
                    const double w = wzy * hat_w(ix, SPAN, inv);
                    v = v + f[ix] * w;
                    wsum = wsum + w;
                }
        }
    }
    double r = 0.0;
    if (!norm) r = v;
    else if (wsum > 1e-12) r = v * (1.0 / wsum);
    const unsigned m = cmask[n];
    if ((m >> c) & 1u) r = 0.0;
    rc[t] = r;
    if (Dc) {
        const double a = Dc[t];
        xc[t] = ((m >> c) & 1u) ? 0.0 : r / (1.0e-30 < a ? a : 1.0e-30);
    }
}

// k_restrict for level l -> l + 1, by coarse DOF for spans 2 and 4
static int launch_restrict(hipStream_t st, const LevelDev& Lf, const LevelDev& Lc, int span, const double* rf,
                           const uint8_t* cmask, double* rc, const double* Dc, double* xc, const int* done, int norm)
{
    const int B = 256;
    if (span == 2 || span == 4) {
        const dim3 grid(unsigned((3 * Lc.nN + B - 1) / B));
        if (span == 2)
            hipLaunchKernelGGL(k_restrict_c<2>, grid, dim3(B), 0, st, Lf, Lc, rf, cmask, rc, Dc, xc, done, norm);
        else
            hipLaunchKernelGGL(k_restrict_c<4>, grid, dim3(B), 0, st, Lf, Lc, rf, cmask, rc, Dc, xc, done, norm);
    } else {
        hipLaunchKernelGGL(k_restrict, dim3(unsigned((Lc.nN + B - 1) / B)), dim3(B), 0, st, Lf, Lc, span, rf, cmask,
                           rc, Dc, xc, done, norm);
    }
    T3D_HIP(hipGetLastError());
    return T3D_OK;
}

// ---------------------------------------------------------------------------
// prolongation + additive correction + post-smoothing + mask
// (transfers.cpp:378-445, preconditioner.cpp:230-253):
//   t = (w*r)/max(1e-30,D);  x = (t + P(xc)) + t;  masked -> 0
// ---------------------------------------------------------------------------
// the value of DOF (n, c): its trilinear prolongation from the coarse nodes
// and the smoothing of its residual
__device__ __forceinline__ double prolong_smooth_dof(const LevelDev& Lf, const LevelDev& Lc, int span,
                                                     const double* __restrict__ xc, const double* __restrict__ rf,
                                                     const double* __restrict__ Df, unsigned m, double weight,
                                                     int64_t n, int c)
{
    const int fx = int(n % Lf.nnx), fy = int((n / Lf.nnx) % Lf.nny), fz = int(n / (int64_t(Lf.nnx) * Lf.nny));
    int cx = fx / span, cy = fy / span, cz = fz / span;
    int rx = fx % span, ry = fy % span, rz = fz % span;
    if (cx >= Lc.resX) { cx = Lc.resX - 1; rx = span; }
    if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = span; }
    if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = span; }
    const double inv = 1.0 / span;
    const double wx1 = double(rx) * inv, wx0 = 1.0 - wx1;
    const double wy1 = double(ry) * inv, wy0 = 1.0 - wy1;
    const double wz1 = double(rz) * inv, wz0 = 1.0 - wz1;
    const int64_t b00 = (int64_t(Lc.nnx) * Lc.nny * cz + int64_t(Lc.nnx) * cy + cx);
    const int64_t b01 = b00 + 1, b10 = b00 + Lc.nnx, b11 = b10 + 1;
    const int64_t s = int64_t(Lc.nnx) * Lc.nny;
    const int64_t n00 = b00 + s, n01 = n00 + 1, n10 = n00 + Lc.nnx, n11 = n10 + 1;
    const double v = xc[3 * b00 + c] * wx0 * wy0 * wz0 + xc[3 * b01 + c] * wx1 * wy0 * wz0 +
                     xc[3 * b10 + c] * wx0 * wy1 * wz0 + xc[3 * b11 + c] * wx1 * wy1 * wz0 +
                     xc[3 * n00 + c] * wx0 * wy0 * wz1 + xc[3 * n01 + c] * wx1 * wy0 * wz1 +
                     xc[3 * n10 + c] * wx0 * wy1 * wz1 + xc[3 * n11 + c] * wx1 * wy1 * wz1;
    double out = 0.0;
    if (!((m >> c) & 1u)) {
        const double d = Df[3 * n + c];
        const double t = weight * rf[3 * n + c] / (1.0e-30 < d ? d : 1.0e-30);
        out = (t + v) + t;
    }
    return out;
}

__global__ void __launch_bounds__(256) k_prolong_smooth(LevelDev Lf, LevelDev Lc, int span,
                                                        const double* __restrict__ xc,
                                                        const double* __restrict__ rf,
                                                        const double* __restrict__ Df,
                                                        const uint8_t* __restrict__ fmask, double weight,
                                                        double* __restrict__ xf, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lf.nN) return;
    const unsigned m = fmask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) xf[3 * n + c] = prolong_smooth_dof(Lf, Lc, span, xc, rf, Df, m, weight, n, c);
}

// The finest level's k_prolong_smooth evaluated per DOF in the canonical tile
// geometry (tile t holds DOFs 1024 t + 256 k + lane, k = 0..3), with the
// tile's partial of r.z summed as k_canon_dot_tiles sums it: the V-cycle's
// last pass and the PCG's dot product in one read of r.  Per DOF, for the
// spans the hierarchy uses and grids below 2^32 nodes (32-bit coordinates;
// the launcher checks): the same operations as prolong_smooth_dof.
template <int SPAN>
__global__ void __launch_bounds__(256) k_prolong_smooth_rz(LevelDev Lf, LevelDev Lc, const double* __restrict__ xc,
                                                           const double* __restrict__ rf,
                                                           const double* __restrict__ Df,
                                                           const uint8_t* __restrict__ fmask, double weight,
                                                           double* __restrict__ xf, double* __restrict__ partial,
                                                           const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[CANON_LANES];
    const int64_t nD = 3 * Lf.nN;
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    const uint32_t nnx = uint32_t(Lf.nnx), nny = uint32_t(Lf.nny);
    const double inv = 1.0 / SPAN;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < nD) {
            const uint32_t n = uint32_t(i / 3);
            const int c = int(i - 3 * int64_t(n));
            const uint32_t q = n / nnx;
            const int fx = int(n - q * nnx), fy = int(q % nny), fz = int(q / nny);
            int cx = fx / SPAN, cy = fy / SPAN, cz = fz / SPAN;
            int rx = fx % SPAN, ry = fy % SPAN, rz = fz % SPAN;
            if (cx >= Lc.resX) { cx = Lc.resX - 1; rx = SPAN; }
            if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = SPAN; }
            if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = SPAN; }
            const double wx1 = double(rx) * inv, wx0 = 1.0 - wx1;
            const double wy1 = double(ry) * inv, wy0 = 1.0 - wy1;
            const double wz1 = double(rz) * inv, wz0 = 1.0 - wz1;
            const int64_t b00 = (int64_t(Lc.nnx) * Lc.nny * cz + int64_t(Lc.nnx) * cy + cx);
            const int64_t b01 = b00 + 1, b10 = b00 + Lc.nnx, b11 = b10 + 1;
            const int64_t sz = int64_t(Lc.nnx) * Lc.nny;
            const int64_t n00 = b00 + sz, n01 = n00 + 1, n10 = n00 + Lc.nnx, n11 = n10 + 1;
            const double v = xc[3 * b00 + c] * wx0 * wy0 * wz0 + xc[3 * b01 + c] * wx1 * wy0 * wz0 +
                             xc[3 * b10 + c] * wx0 * wy1 * wz0 + xc[3 * b11 + c] * wx1 * wy1 * wz0 +
                             xc[3 * n00 + c] * wx0 * wy0 * wz1 + xc[3 * n01 + c] * wx1 * wy0 * wz1 +
                             xc[3 * n10 + c] * wx0 * wy1 * wz1 + xc[3 * n11 + c] * wx1 * wy1 * wz1;
            double out = 0.0;
            if (!((fmask[n] >> c) & 1u)) {
                const double d = Df[i];
                const double t = weight * rf[i] / (1.0e-30 < d ? d : 1.0e-30);
                out = (t + v) + t;
            }
            xf[i] = out;
            s = s + rf[i] * out;
        }
    }
    const double r = block_tree_256(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// single-level hierarchy: z = mask ? 0 : r / max(1e-30, D)
__global__ void k_diag_solve(int64_t nN, const double* __restrict__ r, const double* __restrict__ D,
                             const uint8_t* __restrict__ mask, double* __restrict__ x, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= nN) return;
    const unsigned m = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double d = D[3 * n + c];
        x[3 * n + c] = ((m >> c) & 1u) ? 0.0 : r[3 * n + c] / (1.0e-30 < d ? d : 1.0e-30);
    }
}

// coarsest Cholesky solve (chol.hpp)
__global__ void __launch_bounds__(CHOL_MAX) k_chol_solve(int N, const double* __restrict__ L,
                                                         const double* __restrict__ LT,
                                                         const double* __restrict__ b,
                                                         const uint8_t* __restrict__ mask,
                                                         double* __restrict__ x, const int* __restrict__ done)
{
    if (done && *done) return;
    chol_solve_body(N, L, LT, b, mask, x);
}

// ---------------------------------------------------------------------------
// Galerkin hierarchy (TOP3D_XL.m:945-1090, Solving_AssembleFEAstencil).  The
// level-(l+1) operator is P^T A_l P with P the trilinear prolongation of the
// hierarchy; as P is local to a coarse cell, it is formed cell by cell:
//   K_c = sum over the cells f covered by coarse cell c of T_f^T K_f T_f,
// T_f (24 x 24) interpolating c's corner DOFs to f's corner DOFs.  The finest
// operator is A_0 = M K M + (I - M): a cell with fixed DOFs drops their rows
// and columns and carries 1/nAdj on their diagonal, nAdj the solid cells at
// the node (Solving_ApplyBConEleStiffMat_B, TOP3D_XL.m:1450-1463).  Coarse
// DOFs with a zero diagonal (covering only void) are masked.
// Element matrices are row-major 24 x 24 in local DOF order 3a + c.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void corner_offsets(int a, int& dx, int& dy, int& dz)
{
    const int q = a & 3;
    dx = (q == 1 || q == 2) ? 1 : 0;
    dy = (q <= 1) ? 1 : 0;
    dz = a >> 2;
}

// solid cells among the (up to) eight around node (ix, iy, iz) of the finest grid
__device__ __forceinline__ int solid_cells_at(const Grid& g, const int32_t* __restrict__ cellToCompact, int ix, int iy,
                                              int iz)
{
    int n = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = ix - dx, cy = iy - dy, cz = iz - dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= g.resX || cy >= g.resY || cz >= g.resZ) continue;
                if (cellToCompact[(int64_t(cz) * g.resY + cy) * g.resX + cx] >= 0) ++n;
            }
    return n;
}

// acc[q] += (T_f^T A T_f)[t + 256 q] for the matrix A staged in sA (row-major
// 24 x 24); W: child weights W[a * 8 + A]; sH: LDS scratch
__device__ __forceinline__ void gal_accumulate(const double* sA, double* sH, const double* __restrict__ W, double acc[3])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = t + 256 * q;
        if (e >= 576) break;
        const int ii = e / 24, J = e - 24 * (e / 24), AJ = J / 3, cJ = J - 3 * AJ;
        double h = 0.0;
        for (int b = 0; b < 8; ++b) h += sA[ii * 24 + 3 * b + cJ] * W[b * 8 + AJ];
        sH[e] = h;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = t + 256 * q;
        if (e >= 576) break;
        const int I = e / 24, J = e - 24 * (e / 24), AI = I / 3, cI = I - 3 * AI;
        double v = 0.0;
        for (int a = 0; a < 8; ++a) v += W[a * 8 + AI] * sH[(3 * a + cI) * 24 + J];
        acc[q] += v;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// MATLAB's 'Standard' V-cycle (TOP3D_XL.m:783-803) in the Galerkin form:
//   down:  x_l = w r_l / D_l;  r_{l+1} = P^T (r_l - A_l x_l)
//   up:    x_l += P x_{l+1};   x_l += w (r_l - A_l x_l) / D_l
// every result masked.  A_0 is the finest K*u (fixed DOFs 0); A_l (l >= 1)
// is applied from the level's element matrices by k_gal_apply.
// ---------------------------------------------------------------------------
// op 0: x = w r / D;  op 1: x = r - y;  op 2: x += w (r - y) / D  (masked -> 0)
__global__ void k_std_point(int op, int64_t nN, const double* __restrict__ r, const double* __restrict__ y,
                            const double* __restrict__ D, const uint8_t* __restrict__ mask, double weight,
                            double* __restrict__ x, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= nN) return;
    const unsigned m = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t i = 3 * n + c;
        double out = 0.0;
        if (!((m >> c) & 1u)) {
            if (op == 1) {
                out = r[i] - y[i];
            } else {
                const double d = D[i];
                const double dd = 1.0e-30 < d ? d : 1.0e-30;
                out = (op == 0) ? weight * r[i] / dd : x[i] + weight * (r[i] - y[i]) / dd;
            }
        }
        x[i] = out;
    }
}

// xf += P xc on the fine level (trilinear, as k_prolong_smooth), masked -> 0
__global__ void __launch_bounds__(256) k_prolong_add(LevelDev Lf, LevelDev Lc, int span,
                                                     const double* __restrict__ xc,
                                                     const uint8_t* __restrict__ fmask, double* __restrict__ xf,
                                                     const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= Lf.nN) return;
    const int fx = int(n % Lf.nnx), fy = int((n / Lf.nnx) % Lf.nny), fz = int(n / (int64_t(Lf.nnx) * Lf.nny));
    int cx = fx / span, cy = fy / span, cz = fz / span;
    int rx = fx % span, ry = fy % span, rz = fz % span;
    if (cx >= Lc.resX) { cx = Lc.resX - 1; rx = span; }
    if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = span; }
    if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = span; }
    const double inv = 1.0 / span;
    const double wx1 = double(rx) * inv, wx0 = 1.0 - wx1;
    const double wy1 = double(ry) * inv, wy0 = 1.0 - wy1;
    const double wz1 = double(rz) * inv, wz0 = 1.0 - wz1;
    const int64_t b00 = (int64_t(Lc.nnx) * Lc.nny * cz + int64_t(Lc.nnx) * cy + cx);
    const int64_t b01 = b00 + 1, b10 = b00 + Lc.nnx, b11 = b10 + 1;
    const int64_t s = int64_t(Lc.nnx) * Lc.nny;
    const int64_t n00 = b00 + s, n01 = n00 + 1, n10 = n00 + Lc.nnx, n11 = n10 + 1;
    const unsigned m = fmask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v = xc[3 * b00 + c] * wx0 * wy0 * wz0 + xc[3 * b01 + c] * wx1 * wy0 * wz0 +
                         xc[3 * b10 + c] * wx0 * wy1 * wz0 + xc[3 * b11 + c] * wx1 * wy1 * wz0 +
                         xc[3 * n00 + c] * wx0 * wy0 * wz1 + xc[3 * n01 + c] * wx1 * wy0 * wz1 +
                         xc[3 * n10 + c] * wx0 * wy1 * wz1 + xc[3 * n11 + c] * wx1 * wy1 * wz1;
        xf[3 * n + c] = ((m >> c) & 1u) ? 0.0 : xf[3 * n + c] + v;
    }
}

// y = A_l x from the level's element matrices (row-major 24 x 24 per cell, local
// DOF 3a + c): one thread per node gathers the rows of its corner in each of
// the (up to) eight cells around it
__global__ void __launch_bounds__(256) k_gal_apply(LevelDev L, const double* __restrict__ Kel,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= L.nN) return;
    const int ix = int(n % L.nnx), iy = int((n / L.nnx) % L.nny), iz = int(n / (int64_t(L.nnx) * L.nny));
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = ix - dx, cy = iy - dy, cz = iz - dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= L.resX || cy >= L.resY || cz >= L.resZ) continue;
                // corner a of the node in the cell (corner_offsets(a) == (dx, dy, dz))
                const int a = 4 * dz + (dy ? (dx ? 1 : 0) : (dx ? 2 : 3));
                const double* K = Kel + ((int64_t(cz) * L.resY + cy) * L.resX + cx) * 576 + 3 * a * 24;
                for (int b = 0; b < 8; ++b) {
                    int ox, oy, oz;
                    corner_offsets(b, ox, oy, oz);
                    const int64_t nb = (int64_t(cz + oz) * L.nny + (cy + oy)) * L.nnx + (cx + ox);
                    const double u0 = x[3 * nb], u1 = x[3 * nb + 1], u2 = x[3 * nb + 2];
                    acc0 += K[3 * b] * u0 + K[3 * b + 1] * u1 + K[3 * b + 2] * u2;
                    acc1 += K[24 + 3 * b] * u0 + K[24 + 3 * b + 1] * u1 + K[24 + 3 * b + 2] * u2;
                    acc2 += K[48 + 3 * b] * u0 + K[48 + 3 * b + 1] * u1 + K[48 + 3 * b + 2] * u2;
                }
            }
    y[3 * n] = acc0;
    y[3 * n + 1] = acc1;
    y[3 * n + 2] = acc2;
}

// Level 1 as a product.  A level-1 cell's matrix is sum_f E_f G_f over its
// s^3 finest cells f, G_f = T_f^T Ke T_f shared by all cells; with the
// effective moduli Eeff[cell][f] (E_f, or 0 for void cells, moduli <= 1e-16
// and cells with fixed DOFs) the whole level is Kel = Eeff G, an
// [ncell x s^3] by [s^3 x 576] FP64 product; k_gal_fix then adds the cells
// with fixed DOFs (a static list): for those finest cells it forms
// T_f^T (E M Ke M + D_fix) T_f through LDS.
__global__ void k_gal_eff(Grid g, LevelDev Lc, int s, const double* __restrict__ Ecell,
                          const uint8_t* __restrict__ nodeFixed, const int32_t* __restrict__ cellToCompact,
                          double* __restrict__ Eeff)
{
    const int nch = s * s * s;
    const int64_t n = int64_t(Lc.resX) * Lc.resY * Lc.resZ * nch;
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int64_t ce = idx / nch;
        const int f = int(idx - ce * nch);
        const int cx = int(ce % Lc.resX), cy = int((ce / Lc.resX) % Lc.resY),
                  cz = int(ce / (int64_t(Lc.resX) * Lc.resY));
        const int i = f % s, j = (f / s) % s, k = f / (s * s);
        const int fx = cx * s + i, fy = cy * s + j, fz = cz * s + k;
        const int64_t fc = (int64_t(fz) * g.resY + fy) * g.resX + fx;
        double E = 0.0;
        if (cellToCompact[fc] >= 0) {
            unsigned fixedAny = 0;
            for (int a = 0; a < 8; ++a) {
                int dx, dy, dz;
                corner_offsets(a, dx, dy, dz);
                fixedAny |= nodeFixed[(int64_t(fz + dz) * g.nny + (fy + dy)) * g.nnx + (fx + dx)];
            }
            const double e = Ecell[fc];
            if (!fixedAny) E = (e <= 1.0e-16) ? 0.0 : e;   // fea.cpp:743-744
        }
        Eeff[idx] = E;
    }
}

// Kel[ncell][576] = Eeff[ncell][nch] G[nch][576]: 64 cells (blockIdx.x) x 64
// entries (blockIdx.y) per 256-thread workgroup, 4 x 4 per thread, nch in LDS
// chunks of 16
__global__ void __launch_bounds__(256) k_gal_gemm(int64_t ncell, int nch, const double* __restrict__ Eeff,
                                                  const double* __restrict__ G, double* __restrict__ Kel)
{
    __shared__ double Sa[16][65], Sb[16][65];
    const int64_t c0 = int64_t(blockIdx.x) * 64;
    const int e0 = int(blockIdx.y) * 64;
    const int t = int(threadIdx.x);
    const int ri = (t / 16) * 4, rj = (t % 16) * 4;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < nch; k0 += 16) {
        for (int q = t; q < 16 * 64; q += 256) {
            const int r = q / 16, kk = q % 16;        // Eeff rows: 16 consecutive children of a cell
            const bool in = c0 + r < ncell && k0 + kk < nch;
            Sa[kk][r] = in ? Eeff[(c0 + r) * nch + k0 + kk] : 0.0;
            const int kb = q / 64, c = q % 64;        // G rows: 64 consecutive entries
            Sb[kb][c] = (k0 + kb < nch && e0 + c < 576) ? G[size_t(k0 + kb) * 576 + e0 + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(Sa[kk][ri + a], Sb[kk][rj + b], acc[a][b]);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t c = c0 + ri + a;
        if (c >= ncell) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (e0 + rj + b < 576) Kel[c * 576 + e0 + rj + b] = acc[a][b];
    }
}

// the level-1 cells with fixed DOFs (fixCells): add T_f^T (E M Ke M + D_fix) T_f
// of their finest cells with fixed DOFs, the terms k_gal_eff left out
__global__ void __launch_bounds__(256) k_gal_fix(Grid g, LevelDev Lc, int s, const int32_t* __restrict__ fixCells,
                                                 const double* __restrict__ Wch, const double* __restrict__ Ke,
                                                 const double* __restrict__ Ecell,
                                                 const uint8_t* __restrict__ nodeFixed,
                                                 const int32_t* __restrict__ cellToCompact, double* __restrict__ Kel)
{
    const int64_t ce = fixCells[blockIdx.x];
    const int cx = int(ce % Lc.resX), cy = int((ce / Lc.resX) % Lc.resY), cz = int(ce / (int64_t(Lc.resX) * Lc.resY));
    __shared__ double sA[576], sH[576];
    const int t = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < s; ++k)
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < s; ++i) {
                const int f = (k * s + j) * s + i;
                const int fx = cx * s + i, fy = cy * s + j, fz = cz * s + k;
                const int64_t fc = (int64_t(fz) * g.resY + fy) * g.resX + fx;
                if (cellToCompact[fc] < 0) continue;   // uniform in the block
                unsigned fixedAny = 0;
                for (int a = 0; a < 8; ++a) {
                    int dx, dy, dz;
                    corner_offsets(a, dx, dy, dz);
                    fixedAny |= nodeFixed[(int64_t(fz + dz) * g.nny + (fy + dy)) * g.nnx + (fx + dx)];
                }
                if (!fixedAny) continue;
                double E = Ecell[fc];
                if (E <= 1.0e-16) E = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const int e = t + 256 * q;
                    if (e >= 576) break;
                    const int ii = e / 24, jj = e - 24 * (e / 24);
                    int dx, dy, dz;
                    corner_offsets(ii / 3, dx, dy, dz);
                    const int nx = fx + dx, ny = fy + dy, nz = fz + dz;
                    const bool fi = (nodeFixed[(int64_t(nz) * g.nny + ny) * g.nnx + nx] >> (ii % 3)) & 1u;
                    int ex, ey, ez;
                    corner_offsets(jj / 3, ex, ey, ez);
                    const bool fj =
                        (nodeFixed[(int64_t(fz + ez) * g.nny + (fy + ey)) * g.nnx + (fx + ex)] >> (jj % 3)) & 1u;
                    double v = (fi || fj) ? 0.0 : E * Ke[e];
                    if (ii == jj && fi) v = 1.0 / double(solid_cells_at(g, cellToCompact, nx, ny, nz));
                    sA[e] = v;
                }
                __syncthreads();
                gal_accumulate(sA, sH, Wch + size_t(f) * 64, acc);
            }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = t + 256 * q;
        if (e < 576) Kel[size_t(ce) * 576 + e] += acc[q];
    }
}

// level-(l+1) element matrices from the level-l ones (l >= 1)
__global__ void __launch_bounds__(256) k_gal_coarse(LevelDev Lf, LevelDev Lc, int s, const double* __restrict__ Wch,
                                                    const double* __restrict__ Kf, double* __restrict__ Kc)
{
    const int64_t ce = blockIdx.x;
    const int cx = int(ce % Lc.resX), cy = int((ce / Lc.resX) % Lc.resY), cz = int(ce / (int64_t(Lc.resX) * Lc.resY));
    __shared__ double sA[576], sH[576];
    const int t = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < s; ++k)
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < s; ++i) {
                const int f = (k * s + j) * s + i;
                const int64_t fc = (int64_t(cz * s + k) * Lf.resY + (cy * s + j)) * Lf.resX + (cx * s + i);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const int e = t + 256 * q;
                    if (e < 576) sA[e] = Kf[size_t(fc) * 576 + e];
                }
                __syncthreads();
                gal_accumulate(sA, sH, Wch + size_t(f) * 64, acc);
            }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = t + 256 * q;
        if (e < 576) Kc[size_t(ce) * 576 + e] = acc[q];
    }
}

// diagonal of a Galerkin level from its element matrices (cells in a fixed
// order); DOFs with a zero diagonal are masked and get D = 1
__global__ void __launch_bounds__(256) k_gal_diag(LevelDev L, const double* __restrict__ Kel, double* __restrict__ D,
                                                  uint8_t* __restrict__ mask)
{
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= L.nN) return;
    const int ix = int(n % L.nnx), iy = int((n / L.nnx) % L.nny), iz = int(n / (int64_t(L.nnx) * L.nny));
    double d[3] = {0.0, 0.0, 0.0};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = ix - dx, cy = iy - dy, cz = iz - dz;
                if (cx < 0 || cy < 0 || cz < 0 || cx >= L.resX || cy >= L.resY || cz >= L.resZ) continue;
                const int a = corner_index(dx, dy, dz);
                const double* K = Kel + size_t((int64_t(cz) * L.resY + cy) * L.resX + cx) * 576;
                for (int c = 0; c < 3; ++c) d[c] += K[(3 * a + c) * 25];
            }
    unsigned bits = 0;
    for (int c = 0; c < 3; ++c) {
        if (!(d[c] > 0.0)) {
            bits |= 1u << c;
            d[c] = 1.0;
        }
        D[3 * n + c] = d[c];
    }
    mask[n] = uint8_t(bits);
}

// ---------------------------------------------------------------------------
// Dense right-looking blocked Cholesky A = L L^T in FP64 on the device (the
// Galerkin form's coarsest matrix, row-major N x N, lower triangle used and
// overwritten by L).  Per panel of CB columns at k0: the diagonal block is
// factored by one workgroup in LDS, the rows below are solved against it, and
// the trailing lower triangle is updated by 64 x 64 tiles.
// ---------------------------------------------------------------------------
constexpr int CB = 32;

__global__ void __launch_bounds__(256) k_chol_diag(int N, double* __restrict__ A, int k0, int* __restrict__ fail)
{
    __shared__ double a[CB][CB + 1];
    const int nb = min(CB, N - k0);
    for (int e = threadIdx.x; e < CB * CB; e += 256) {
        const int i = e / CB, j = e % CB;
        a[i][j] = (i < nb && j < nb && j <= i) ? A[size_t(k0 + i) * N + k0 + j] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const double d = a[j][j];
        __syncthreads();
        if (!(d > 0.0)) {
            if (threadIdx.x == 0) *fail = 1;
            return;   // uniform: every thread read the same d
        }
        const double s = sqrt(d);
        if (threadIdx.x < CB) {
            const int i = threadIdx.x;
            if (i == j) a[j][j] = s;
            else if (i > j && i < nb) a[i][j] /= s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CB * CB; e += 256) {
            const int i = e / CB, l = e % CB;
            if (i > j && l > j && l <= i && i < nb) a[i][l] -= a[i][j] * a[l][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < CB * CB; e += 256) {
        const int i = e / CB, j = e % CB;
        if (i < nb && j < nb && j <= i) A[size_t(k0 + i) * N + k0 + j] = a[i][j];
    }
}

// rows i >= k0 + CB of the panel: L[i][k0:k0+CB] = A[i][k0:k0+CB] L_kk^-T
__global__ void __launch_bounds__(256) k_chol_panel(int N, double* __restrict__ A, int k0,
                                                    const int* __restrict__ fail)
{
    if (*fail) return;
    __shared__ double l[CB][CB + 1];
    for (int e = threadIdx.x; e < CB * CB; e += 256) {
        const int i = e / CB, j = e % CB;
        l[i][j] = (j <= i) ? A[size_t(k0 + i) * N + k0 + j] : 0.0;
    }
    __syncthreads();
    const int i = k0 + CB + int(blockIdx.x) * 256 + int(threadIdx.x);
    if (i >= N) return;
    double x[CB];
    double* row = A + size_t(i) * N + k0;
#pragma unroll
    for (int j = 0; j < CB; ++j) x[j] = row[j];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        double v = x[j];
#pragma unroll
        for (int t = 0; t < j; ++t) v -= x[t] * l[j][t];
        x[j] = v / l[j][j];
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) row[j] = x[j];
}

// trailing update of the lower triangle: A[i][j] -= sum_t L[i][k0+t] L[j][k0+t]
// for k0 + CB <= j <= i < N, 64 x 64 tiles, 4 x 4 outputs per thread
__global__ void __launch_bounds__(256) k_chol_update(int N, double* __restrict__ A, int k0,
                                                     const int* __restrict__ fail)
{
    if (*fail) return;
    const int b0 = k0 + CB;
    const int ti = int(blockIdx.y), tj = int(blockIdx.x);
    if (tj > ti) return;   // upper tiles
    const int i0 = b0 + 64 * ti, j0 = b0 + 64 * tj;
    __shared__ double Pi[64][CB + 1], Pj[64][CB + 1];
    for (int e = threadIdx.x; e < 64 * CB; e += 256) {
        const int r = e / CB, t = e % CB;
        Pi[r][t] = (i0 + r < N) ? A[size_t(i0 + r) * N + k0 + t] : 0.0;
        Pj[r][t] = (j0 + r < N) ? A[size_t(j0 + r) * N + k0 + t] : 0.0;
    }
    __syncthreads();
    const int ri = (int(threadIdx.x) / 16) * 4, rj = (int(threadIdx.x) % 16) * 4;
    double acc[4][4] = {};
    for (int t = 0; t < CB; ++t)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += Pi[ri + a][t] * Pj[rj + b][t];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ri + a, j = j0 + rj + b;
            if (i < N && j <= i) A[size_t(i) * N + j] -= acc[a][b];
        }
}

// coarsest solve of the Galerkin form through the inverse factor (coarse.hpp)
__global__ void __launch_bounds__(256) k_mg_gram(int N, const double* __restrict__ Li, double* __restrict__ Ap)
{
    int I, J;
    coarse_tile_ij(blockIdx.x, I, J);
    gram_body(N, Li, Ap, I, J);
}

__global__ void __launch_bounds__(256) k_mg_coarse_tiles(int N, const double* __restrict__ Ap,
                                                         const double* __restrict__ v, double* __restrict__ P,
                                                         const int* __restrict__ done)
{
    if (done && *done) return;
    coarse_tiles_loop(coarse_tiles(N), [=](unsigned e) { return CoarseTile{N, Ap, v, P, e, true}; });
}

__global__ void __launch_bounds__(CT) k_mg_coarse_sum(int N, const double* __restrict__ P,
                                                      const uint8_t* __restrict__ mask, double* __restrict__ out,
                                                      const int* __restrict__ done)
{
    if (done && *done) return;
    coarse_sum_body(N, P, mask, out, nullptr, int(blockIdx.x));
}

}  // namespace t3d

using namespace t3d;

struct t3d_mg {
    t3d_problem* p = nullptr;
    HostMG H;
    int nlev = 0;
    LevelDev ld[8];
    uint8_t* mask[8] = {};
    double* D[8] = {};
    double* r[8] = {};
    double* x[8] = {};
    double* W[8] = {};
    double* L = nullptr;
    double* LT = nullptr;
    int Nc = 0;
    int cholOk = 1;
    bool setupDone = false;
    std::vector<double> Kc;   // last assembled coarsest matrix (host, for inspection)
    // Galerkin form (mode 1)
    int mode = 0;             // 0: the reference's V-cycle; 1: Galerkin hierarchy
    double* Kel[8] = {};      // element matrices of levels >= 1, [cells][576]
    double* Wch[8] = {};      // child interpolation weights of levels >= 1, [span^3][64]
    double* Gch = nullptr;    // T_f^T Ke T_f per finest cell of a level-1 cell, [span^3][576]
    double* Eeff = nullptr;   // level-1 effective moduli, [cells][span^3]
    int32_t* fixCells = nullptr;   // level-1 cells covering finest cells with fixed DOFs
    int nFix = 0;
    int vtype = 0;            // Galerkin V-cycle: 0 'Adapted', 1 'Standard'
    double* ys[8] = {};       // 'Standard': A_l x and residuals per level l < L
    double* Li = nullptr;     // inverse coarsest factor L^-1
    double* Ainv = nullptr;   // Li^T Li: the coarsest inverse of the FP64 and mixed V-cycles (packed lower tiles)
    double* cP = nullptr;     // partial products of the packed coarsest product
    double* yc = nullptr;     // device flag of the factorisation (after N doubles)
};

namespace t3d {

// x = Ainv r over the packed tiles of the coarsest inverse, masked (Galerkin form)
static void launch_mg_coarse(t3d_mg* g, const double* r, double* x, const int* done)
{
    hipStream_t st = g->p->stream;
    const int L = g->nlev - 1;
    hipLaunchKernelGGL(k_mg_coarse_tiles, dim3(coarse_tile_blocks(coarse_tiles(g->Nc))), dim3(256), 0, st, g->Nc,
                       (const double*)g->Ainv,
                       r, g->cP, done);
    hipLaunchKernelGGL(k_mg_coarse_sum, dim3(unsigned(coarse_nt(g->Nc))), dim3(CT), 0, st, g->Nc,
                       (const double*)g->cP, (const uint8_t*)g->mask[L], x, done);
}

// the 'Standard' V-cycle of the Galerkin form (TOP3D_XL.m:783-803); z holds x_0
static int mg_apply_standard(t3d_mg* g, const double* r0, double* z, const int* done)
{
    t3d_problem* p = g->p;
    hipStream_t st = p->stream;
    const int L = g->nlev - 1;
    const int B = 256;
    auto grid = [&](int l) { return dim3(unsigned((g->ld[l].nN + B - 1) / B)); };
    auto apply = [&](int l, const double* x, double* y) -> int {
        if (l == 0) return launch_kxu_f64(p, x, y, done);
        hipLaunchKernelGGL(k_gal_apply, grid(l), dim3(B), 0, st, g->ld[l], (const double*)g->Kel[l], x, y, done);
        return hipGetLastError() == hipSuccess ? T3D_OK : T3D_ERR_HIP;
    };
    const double w = g->H.weight;
    int rc;
    for (int l = 0; l < L; ++l) {
        const double* rl = (l == 0) ? r0 : g->r[l];
        double* xl = (l == 0) ? z : g->x[l];
        hipLaunchKernelGGL(k_std_point, grid(l), dim3(B), 0, st, 0, g->ld[l].nN, rl, (const double*)nullptr,
                           (const double*)g->D[l], (const uint8_t*)g->mask[l], w, xl, done);
        if ((rc = apply(l, xl, g->ys[l]))) return rc;
        hipLaunchKernelGGL(k_std_point, grid(l), dim3(B), 0, st, 1, g->ld[l].nN, rl, (const double*)g->ys[l],
                           (const double*)nullptr, (const uint8_t*)g->mask[l], w, g->ys[l], done);
        int rc = launch_restrict(st, g->ld[l], g->ld[l + 1], g->H.lv[l + 1].span, (const double*)g->ys[l],
                                 (const uint8_t*)g->mask[l + 1], g->r[l + 1], nullptr, g->x[L], done, 0);
        if (rc) return rc;
    }
    launch_mg_coarse(g, g->r[L], g->x[L], done);
    T3D_HIP(hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const double* rl = (l == 0) ? r0 : g->r[l];
        double* xl = (l == 0) ? z : g->x[l];
        hipLaunchKernelGGL(k_prolong_add, grid(l), dim3(B), 0, st, g->ld[l], g->ld[l + 1], g->H.lv[l + 1].span,
                           (const double*)g->x[l + 1], (const uint8_t*)g->mask[l], xl, done);
        if ((rc = apply(l, xl, g->ys[l]))) return rc;
        hipLaunchKernelGGL(k_std_point, grid(l), dim3(B), 0, st, 2, g->ld[l].nN, rl, (const double*)g->ys[l],
                           (const double*)g->D[l], (const uint8_t*)g->mask[l], w, xl, done);
        T3D_HIP(hipGetLastError());
    }
    return T3D_OK;
}

int mg_apply_dev(t3d_mg* g, const double* r0, double* z, const int* done, double* rzDst)
{
    t3d_problem* p = g->p;
    hipStream_t st = p->stream;
    const int L = g->nlev - 1;
    const int B = 256;
    if (L == 0 || (g->mode == 1 && g->vtype == 1)) {
        if (L == 0) {
            hipLaunchKernelGGL(k_diag_solve, dim3(unsigned((g->ld[0].nN + B - 1) / B)), dim3(B), 0, st, g->ld[0].nN,
                               r0, g->D[0], g->mask[0], z, done);
            T3D_HIP(hipGetLastError());
        } else {
            int rc = mg_apply_standard(g, r0, z, done);
            if (rc) return rc;
        }
        return rzDst ? canon_dot_dev(p, r0, z, 3 * g->ld[0].nN, rzDst, done) : T3D_OK;
    }
    const bool chol = g->Nc > 0;
    for (int l = 0; l < L; ++l) {
        const double* rf = (l == 0) ? r0 : g->r[l];
        const bool last = (l + 1 == L);
        int rc = launch_restrict(st, g->ld[l], g->ld[l + 1], g->H.lv[l + 1].span, rf, g->mask[l + 1], g->r[l + 1],
                                 (last && !chol) ? g->D[L] : nullptr, g->x[L], done, g->mode == 0 ? 1 : 0);
        if (rc) return rc;
    }
    if (g->mode == 1) {   // x = Ainv r = Li^T Li r, masked
        launch_mg_coarse(g, g->r[L], g->x[L], done);
        T3D_HIP(hipGetLastError());
    } else if (chol) {
        hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(CHOL_MAX), 0, st, g->Nc, g->L, g->LT, g->r[L], g->mask[L],
                           g->x[L], done);
        T3D_HIP(hipGetLastError());
    }
    for (int l = L - 1; l >= 0; --l) {
        const double* rf = (l == 0) ? r0 : g->r[l];
        double* xo = (l == 0) ? z : g->x[l];
        const int span0 = g->H.lv[1].span;
        if (l == 0 && rzDst && (span0 == 2 || span0 == 4) && g->ld[0].nN < (int64_t(1) << 32)) {
            // with the r.z partials, in the canonical geometry
            const int64_t nD = 3 * g->ld[0].nN;
            int rc = ensure_red(p, nD);
            if (rc) return rc;
            const int64_t tiles = (nD + CANON_TILE - 1) / CANON_TILE;
            if (span0 == 2)
                hipLaunchKernelGGL(k_prolong_smooth_rz<2>, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, st, g->ld[0],
                                   g->ld[1], g->x[1], rf, g->D[0], g->mask[0], g->H.weight, xo, p->red, done);
            else
                hipLaunchKernelGGL(k_prolong_smooth_rz<4>, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, st, g->ld[0],
                                   g->ld[1], g->x[1], rf, g->D[0], g->mask[0], g->H.weight, xo, p->red, done);
            T3D_HIP(hipGetLastError());
            return canon_finish(p, tiles, rzDst, done);
        }
        hipLaunchKernelGGL(k_prolong_smooth, dim3(unsigned((g->ld[l].nN + B - 1) / B)), dim3(B), 0, st, g->ld[l],
                           g->ld[l + 1], g->H.lv[l + 1].span, g->x[l + 1], rf, g->D[l], g->mask[l], g->H.weight, xo,
                           done);
        T3D_HIP(hipGetLastError());
    }
    return rzDst ? canon_dot_dev(p, r0, z, 3 * g->ld[0].nN, rzDst, done) : T3D_OK;
}

}  // namespace t3d

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
    int norm;   // 1: the reference's normalised restriction; 0: the Galerkin form's P^T
    const double* Ainv;  // coarsest inverse Li^T Li when the multigrid keeps it
    int vtype;           // 0 'Adapted', 1 'Standard' (Galerkin form)
    const double* Kel[8];   // Galerkin element matrices of levels >= 1 ([cells][576]), else null
};
int mg_view(const t3d_mg* g, MGView& v)
{
    T3D_CHECK(g && g->setupDone, T3D_ERR_STATE, "multigrid object not set up");
    v.nlev = g->nlev;
    for (int l = 0; l < g->nlev; ++l) {
        v.span[l] = g->H.lv[l].span;
        v.resX[l] = g->H.lv[l].resX; v.resY[l] = g->H.lv[l].resY; v.resZ[l] = g->H.lv[l].resZ;
        v.nN[l] = g->H.lv[l].nN;
        v.mask[l] = g->mask[l];
        v.D[l] = g->D[l];
    }
    v.L = g->L;
    v.LT = g->LT;
    v.Nc = g->Nc;
    v.weight = g->H.weight;
    v.norm = g->mode == 0 ? 1 : 0;
    v.Ainv = g->mode == 1 ? g->Ainv : nullptr;
    v.vtype = g->vtype;
    for (int l = 0; l < 8; ++l) v.Kel[l] = (g->mode == 1 && l >= 1 && l < g->nlev) ? g->Kel[l] : nullptr;
    return T3D_OK;
}
}  // namespace t3d

namespace t3d {

// child interpolation weights of a span-s coarse cell:
// W[f * 64 + a * 8 + A] = trilinear weight of coarse corner A at corner a of
// the covered cell f = (k s + j) s + i (TOP3D_XL.m:1881-1980 tabulates the same
// weights per embedded node)
static std::vector<double> child_weights(int s)
{
    auto off = [](int a, int& dx, int& dy, int& dz) {
        const int q = a & 3;
        dx = (q == 1 || q == 2) ? 1 : 0;
        dy = (q <= 1) ? 1 : 0;
        dz = a >> 2;
    };
    std::vector<double> W(size_t(s) * s * s * 64);
    for (int k = 0; k < s; ++k)
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < s; ++i) {
                const int f = (k * s + j) * s + i;
                for (int a = 0; a < 8; ++a) {
                    int dx, dy, dz;
                    off(a, dx, dy, dz);
                    const double px = double(i + dx) / s, py = double(j + dy) / s, pz = double(k + dz) / s;
                    for (int A = 0; A < 8; ++A) {
                        int ex, ey, ez;
                        off(A, ex, ey, ez);
                        W[size_t(f) * 64 + a * 8 + A] = (ex ? px : 1.0 - px) * (ey ? py : 1.0 - py) * (ez ? pz : 1.0 - pz);
                    }
                }
            }
    return W;
}

// G_f = T_f^T Ke T_f for each covered cell f
static std::vector<double> child_products(const std::vector<double>& W, const double* Ke, int nch)
{
    std::vector<double> G(size_t(nch) * 576, 0.0);
    for (int f = 0; f < nch; ++f) {
        const double* w = &W[size_t(f) * 64];
        std::vector<double> H(576, 0.0);   // H = Ke T_f
        for (int ii = 0; ii < 24; ++ii)
            for (int J = 0; J < 24; ++J) {
                double h = 0.0;
                for (int b = 0; b < 8; ++b) h += Ke[ii * 24 + 3 * b + J % 3] * w[b * 8 + J / 3];
                H[ii * 24 + J] = h;
            }
        for (int I = 0; I < 24; ++I)
            for (int J = 0; J < 24; ++J) {
                double v = 0.0;
                for (int a = 0; a < 8; ++a) v += w[a * 8 + I / 3] * H[(3 * a + I % 3) * 24 + J];
                G[size_t(f) * 576 + I * 24 + J] = v;
            }
    }
    return G;
}

// level-1 cells (structured index, increasing) that cover a solid finest cell
// with a fixed DOF: the cells around each fixed node
static std::vector<int32_t> fixed_level1_cells(const t3d_problem* p, const MGLevelDims& c)
{
    const HostModel& m = p->host;
    const Grid& g = p->g;
    const int s = c.span;
    std::vector<uint8_t> solid(size_t(g.ncell), 0), mark(size_t(c.resX) * c.resY * c.resZ, 0);
    for (int e = 0; e < m.numElements; ++e) solid[size_t(cell_of_full(m.eleMapBack[e], m.resX, m.resY))] = 1;
    for (int64_t n = 0; n < g.nN; ++n) {
        if (m.isFreeDOF[3 * n] && m.isFreeDOF[3 * n + 1] && m.isFreeDOF[3 * n + 2]) continue;
        const int ix = int(n % g.nnx), iy = int((n / g.nnx) % g.nny), iz = int(n / (int64_t(g.nnx) * g.nny));
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int fx = ix - dx, fy = iy - dy, fz = iz - dz;
                    if (fx < 0 || fy < 0 || fz < 0 || fx >= g.resX || fy >= g.resY || fz >= g.resZ) continue;
                    if (!solid[(size_t(fz) * g.resY + fy) * g.resX + fx]) continue;
                    mark[(size_t(fz / s) * c.resY + fy / s) * c.resX + fx / s] = 1;
                }
    }
    std::vector<int32_t> out;
    for (size_t i = 0; i < mark.size(); ++i)
        if (mark[i]) out.push_back(int32_t(i));
    return out;
}

// Galerkin set-up: element matrices of every coarse level, their diagonals and
// masks, the coarsest matrix (assembled here in cell order), its Cholesky
// factor and inverse factor
static int mg_galerkin_setup(t3d_mg* g)
{
    t3d_problem* p = g->p;
    hipStream_t st = p->stream;
    const int B = 256;
    const int L = g->nlev - 1;
    T3D_CHECK(L >= 1, T3D_ERR_STATE, "Galerkin multigrid needs at least two levels");
    for (int l = 1; l <= L; ++l) {
        const int sp = g->H.lv[l].span;
        const size_t nch = size_t(sp) * sp * sp;
        const int64_t ncell = int64_t(g->H.lv[l].resX) * g->H.lv[l].resY * g->H.lv[l].resZ;
        if (!g->Kel[l]) {
            const std::vector<double> W = child_weights(sp);
            T3D_HIP(hipMalloc(&g->Kel[l], sizeof(double) * 576 * size_t(ncell)));
            T3D_HIP(hipMalloc(&g->Wch[l], sizeof(double) * W.size()));
            T3D_HIP(hipMemcpy(g->Wch[l], W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice));
            if (l == 1) {
                const std::vector<double> G = child_products(W, p->host.Ke.data(), int(nch));
                T3D_HIP(hipMalloc(&g->Gch, sizeof(double) * G.size()));
                T3D_HIP(hipMemcpy(g->Gch, G.data(), sizeof(double) * G.size(), hipMemcpyHostToDevice));
                T3D_HIP(hipMalloc(&g->Eeff, sizeof(double) * nch * size_t(ncell)));
                const std::vector<int32_t> fix = fixed_level1_cells(p, g->H.lv[1]);
                g->nFix = int(fix.size());
                if (g->nFix) {
                    T3D_HIP(hipMalloc(&g->fixCells, sizeof(int32_t) * fix.size()));
                    T3D_HIP(hipMemcpy(g->fixCells, fix.data(), sizeof(int32_t) * fix.size(), hipMemcpyHostToDevice));
                }
            }
        }
        if (l == 1) {
            hipLaunchKernelGGL(k_gal_eff, dim3(4096), dim3(256), 0, st, p->g, g->ld[1], sp, (const double*)p->Ecell,
                               (const uint8_t*)p->nodeFixed, (const int32_t*)p->cellToCompact, g->Eeff);
            hipLaunchKernelGGL(k_gal_gemm, dim3(unsigned((ncell + 63) / 64), 9), dim3(256), 0, st, ncell, int(nch),
                               (const double*)g->Eeff, (const double*)g->Gch, g->Kel[1]);
            if (g->nFix)
                hipLaunchKernelGGL(k_gal_fix, dim3(unsigned(g->nFix)), dim3(256), 0, st, p->g, g->ld[1], sp,
                                   (const int32_t*)g->fixCells, (const double*)g->Wch[1], (const double*)p->Ke,
                                   (const double*)p->Ecell, (const uint8_t*)p->nodeFixed,
                                   (const int32_t*)p->cellToCompact, g->Kel[1]);
        } else
            hipLaunchKernelGGL(k_gal_coarse, dim3(unsigned(ncell)), dim3(256), 0, st, g->ld[l - 1], g->ld[l], sp,
                               (const double*)g->Wch[l], (const double*)g->Kel[l - 1], g->Kel[l]);
        T3D_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gal_diag, dim3(unsigned((g->ld[l].nN + B - 1) / B)), dim3(B), 0, st, g->ld[l],
                           (const double*)g->Kel[l], g->D[l], g->mask[l]);
        T3D_HIP(hipGetLastError());
    }
    // coarsest matrix, assembled on the host from the coarsest element matrices
    const MGLevelDims& c = g->H.lv[L];
    const int64_t ncell = int64_t(c.resX) * c.resY * c.resZ;
    const int nnx = c.resX + 1, nny = c.resY + 1;
    const int N = int(3 * c.nN);
    std::vector<double> Ke(576 * size_t(ncell));
    std::vector<uint8_t> mk(c.nN);
    T3D_HIP(hipMemcpyAsync(Ke.data(), g->Kel[L], sizeof(double) * Ke.size(), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipMemcpyAsync(mk.data(), g->mask[L], c.nN, hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));
    std::vector<double> A(size_t(N) * N, 0.0);
    for (int64_t ce = 0; ce < ncell; ++ce) {
        const int cx = int(ce % c.resX), cy = int((ce / c.resX) % c.resY), cz = int(ce / (int64_t(c.resX) * c.resY));
        int node[8];
        for (int a = 0; a < 8; ++a) {
            const int q = a & 3;
            const int dx = (q == 1 || q == 2) ? 1 : 0, dy = (q <= 1) ? 1 : 0, dz = a >> 2;
            node[a] = ((cz + dz) * nny + (cy + dy)) * nnx + (cx + dx);
        }
        const double* K = &Ke[size_t(ce) * 576];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j)
                A[size_t(3 * node[i / 3] + i % 3) * N + 3 * node[j / 3] + j % 3] += K[i * 24 + j];
    }
    for (int i = 0; i < N; ++i)
        if ((mk[i / 3] >> (i % 3)) & 1u) A[size_t(i) * N + i] = 1.0;   // a DOF covering only void
    g->Kc = A;
    // factor on the device (upper triangle zeroed first), then the inverse factor
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) A[size_t(i) * N + j] = 0.0;
    const size_t n2 = size_t(N) * N;
    if (!g->L) T3D_HIP(hipMalloc(&g->L, sizeof(double) * n2));
    if (!g->Li) {
        T3D_HIP(hipMalloc(&g->Li, sizeof(double) * n2));
        T3D_HIP(hipMalloc(&g->Ainv, sizeof(double) * coarse_packed_doubles(N)));
        T3D_HIP(hipMalloc(&g->cP, sizeof(double) * coarse_partial_doubles(N)));
        T3D_HIP(hipMalloc(&g->yc, sizeof(double) * (size_t(N) + 1)));
    }
    int* fail = reinterpret_cast<int*>(g->yc + N);
    T3D_HIP(hipMemcpyAsync(g->L, A.data(), sizeof(double) * n2, hipMemcpyHostToDevice, st));
    T3D_HIP(hipMemsetAsync(fail, 0, sizeof(int), st));
    for (int k0 = 0; k0 < N; k0 += CB) {
        hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, st, N, g->L, k0, fail);
        const int rest = N - k0 - CB;
        if (rest > 0) {
            hipLaunchKernelGGL(k_chol_panel, dim3(unsigned((rest + 255) / 256)), dim3(256), 0, st, N, g->L, k0,
                               (const int*)fail);
            const unsigned tiles = unsigned((rest + 63) / 64);
            hipLaunchKernelGGL(k_chol_update, dim3(tiles, tiles), dim3(256), 0, st, N, g->L, k0, (const int*)fail);
        }
    }
    T3D_HIP(hipGetLastError());
    int hfail = 0;
    T3D_HIP(hipMemcpyAsync(&hfail, fail, sizeof(int), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));   // A is a host buffer of this scope
    if (hfail) {
        g->cholOk = 0;
        set_error("Galerkin multigrid: the coarsest matrix is not positive definite");
        return T3D_ERR_NOT_SPD;
    }
    T3D_HIP(launch_tri_inv(st, N, g->L, g->Li));
    hipLaunchKernelGGL(k_mg_gram, dim3(coarse_tiles(N)), dim3(256), 0, st, N, (const double*)g->Li, g->Ainv);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(st));
    g->Nc = N;
    return T3D_OK;
}

}  // namespace t3d

extern "C" {

int t3d_mg_destroy(t3d_mg* g)
{
    if (!g) return T3D_OK;
    for (int l = 0; l < 8; ++l) {
        if (g->mask[l]) (void)hipFree(g->mask[l]);
        if (g->D[l]) (void)hipFree(g->D[l]);
        if (g->r[l]) (void)hipFree(g->r[l]);
        if (g->x[l]) (void)hipFree(g->x[l]);
        if (g->W[l]) (void)hipFree(g->W[l]);
    }
    for (int l = 0; l < 8; ++l) {
        if (g->Kel[l]) (void)hipFree(g->Kel[l]);
        if (g->Wch[l]) (void)hipFree(g->Wch[l]);
    }
    if (g->Gch) (void)hipFree(g->Gch);
    if (g->Eeff) (void)hipFree(g->Eeff);
    for (int l = 0; l < 8; ++l)
        if (g->ys[l]) (void)hipFree(g->ys[l]);
    if (g->fixCells) (void)hipFree(g->fixCells);
    if (g->Li) (void)hipFree(g->Li);
    if (g->Ainv) (void)hipFree(g->Ainv);
    if (g->cP) (void)hipFree(g->cP);
    if (g->yc) (void)hipFree(g->yc);
    if (g->L) (void)hipFree(g->L);
    if (g->LT) (void)hipFree(g->LT);
    delete g;
    return T3D_OK;
}

int t3d_mg_create(t3d_problem* p, int nonDyadic, int maxLevels, float weight, float Emin, int NlimitDofs,
                  t3d_mg** out)
{
    T3D_CHECK(p && out && maxLevels >= 1 && maxLevels <= 8 && NlimitDofs <= CHOL_MAX, T3D_ERR_ARG,
              "t3d_mg_create: bad argument (maxLevels 1..8, NlimitDofs <= 1024)");
    *out = nullptr;
    t3d_mg* g = new t3d_mg();
    g->p = p;
    mg_build_static(p->host, nonDyadic != 0, maxLevels, weight, Emin, NlimitDofs, g->H);
    g->nlev = int(g->H.lv.size());
    auto fail = [&](const char* m) {
        set_error(m);
        t3d_mg_destroy(g);
        return T3D_ERR_HIP;
    };
    for (int l = 0; l < g->nlev; ++l) {
        g->ld[l] = level_dev(g->H.lv[l]);
        const int64_t nN = g->H.lv[l].nN;
        std::vector<uint8_t> bits(nN);
        for (int64_t n = 0; n < nN; ++n)
            bits[n] = uint8_t(g->H.masks[l][3 * n] | (g->H.masks[l][3 * n + 1] << 1) | (g->H.masks[l][3 * n + 2] << 2));
        if (hipMalloc(&g->mask[l], nN) != hipSuccess ||
            hipMemcpy(g->mask[l], bits.data(), nN, hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc(&g->D[l], sizeof(double) * 3 * nN) != hipSuccess)
            return fail("t3d_mg_create: device allocation failed");
        if (l > 0) {
            const auto& W = g->H.W[l];
            if (hipMalloc(&g->r[l], sizeof(double) * 3 * nN) != hipSuccess ||
                hipMalloc(&g->x[l], sizeof(double) * 3 * nN) != hipSuccess ||
                hipMalloc(&g->W[l], sizeof(double) * W.size()) != hipSuccess ||
                hipMemcpy(g->W[l], W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice) != hipSuccess)
                return fail("t3d_mg_create: device allocation failed");
        }
    }
    // the set-up copies and fills above ran on the null stream
    T3D_HIP(hipDeviceSynchronize());
    *out = g;
    return T3D_OK;
}

// level dims: out[4*l..] = resX resY resZ span; *nlev, *Nc (0 = diagonal coarsest)
int t3d_mg_info(const t3d_mg* g, int* nlev, int* dims, int* Nc, int* cholOk)
{
    T3D_CHECK(g, T3D_ERR_ARG, "t3d_mg_info: null");
    if (nlev) *nlev = g->nlev;
    if (dims)
        for (int l = 0; l < g->nlev; ++l) {
            dims[4 * l] = g->H.lv[l].resX; dims[4 * l + 1] = g->H.lv[l].resY;
            dims[4 * l + 2] = g->H.lv[l].resZ; dims[4 * l + 3] = g->H.lv[l].span;
        }
    if (Nc) *Nc = g->Nc;
    if (cholOk) *cholOk = g->cholOk;
    return T3D_OK;
}

int t3d_mg_mask(const t3d_mg* g, int level, uint8_t* out)
{
    T3D_CHECK(g && out && level >= 0 && level < g->nlev, T3D_ERR_ARG, "t3d_mg_mask: bad argument");
    if (g->mode == 1 && level > 0) {   // set up with the element matrices
        T3D_CHECK(g->setupDone, T3D_ERR_STATE, "t3d_mg_mask: t3d_mg_setup not called");
        const int64_t nN = g->H.lv[level].nN;
        std::vector<uint8_t> bits(nN);
        T3D_HIP(hipDeviceSynchronize());
        T3D_HIP(hipMemcpy(bits.data(), g->mask[level], nN, hipMemcpyDeviceToHost));
        for (int64_t n = 0; n < nN; ++n)
            for (int c = 0; c < 3; ++c) out[3 * n + c] = (bits[n] >> c) & 1u;
        return T3D_OK;
    }
    std::copy(g->H.masks[level].begin(), g->H.masks[level].end(), out);
    return T3D_OK;
}

// multigrid form (before t3d_mg_setup): 0 the reference's V-cycle, 1 the
// Galerkin hierarchy of TOP3D_XL.m ('Adapted' V-cycle, TOP3D_XL.m:772-807)
int t3d_mg_set_mode(t3d_mg* g, int mode)
{
    T3D_CHECK(g && (mode == 0 || mode == 1), T3D_ERR_ARG, "t3d_mg_set_mode: mode must be 0 or 1");
    T3D_CHECK(mode == 0 || g->nlev > 1, T3D_ERR_ARG, "t3d_mg_set_mode: the Galerkin form needs two levels");
    if (mode == g->mode) return T3D_OK;
    if (mode == 0) {   // restore the reference's restricted masks
        for (int l = 1; l < g->nlev; ++l) {
            const int64_t nN = g->H.lv[l].nN;
            std::vector<uint8_t> bits(nN);
            for (int64_t n = 0; n < nN; ++n)
                bits[n] = uint8_t(g->H.masks[l][3 * n] | (g->H.masks[l][3 * n + 1] << 1) |
                                  (g->H.masks[l][3 * n + 2] << 2));
            T3D_HIP(hipMemcpy(g->mask[l], bits.data(), nN, hipMemcpyHostToDevice));
        }
        T3D_HIP(hipDeviceSynchronize());
    }
    g->mode = mode;
    if (mode == 0) g->vtype = 0;
    g->setupDone = false;
    return T3D_OK;
}

int t3d_mg_set_vcycle(t3d_mg* g, int type)
{
    T3D_CHECK(g && (type == 0 || type == 1), T3D_ERR_ARG, "t3d_mg_set_vcycle: type must be 0 or 1");
    T3D_CHECK(type == 0 || g->mode == 1, T3D_ERR_STATE,
              "t3d_mg_set_vcycle: the 'Standard' V-cycle needs the Galerkin form (t3d_mg_set_mode(g, 1))");
    if (type == 1)
        for (int l = 0; l + 1 < g->nlev; ++l)
            if (!g->ys[l]) T3D_HIP(hipMalloc(&g->ys[l], sizeof(double) * 3 * size_t(g->H.lv[l].nN)));
    g->vtype = type;
    return T3D_OK;
}

int t3d_mg_diag(const t3d_mg* g, int level, double* out)
{
    T3D_CHECK(g && out && level >= 0 && level < g->nlev, T3D_ERR_ARG, "t3d_mg_diag: bad argument");
    T3D_CHECK(g->setupDone, T3D_ERR_STATE, "t3d_mg_diag: t3d_mg_setup not called");
    T3D_HIP(hipDeviceSynchronize());
    T3D_HIP(hipMemcpy(out, g->D[level], sizeof(double) * 3 * g->H.lv[level].nN, hipMemcpyDeviceToHost));
    return T3D_OK;
}

// coarsest matrix before factorisation (N x N) and factor (N x N, if SPD)
int t3d_mg_coarsest(const t3d_mg* g, double* Kc, double* Lf)
{
    T3D_CHECK(g && g->setupDone, T3D_ERR_STATE, "t3d_mg_coarsest: not set up");
    const int N = int(3 * g->H.lv.back().nN);
    if (Kc && !g->Kc.empty()) std::copy(g->Kc.begin(), g->Kc.end(), Kc);
    T3D_HIP(hipDeviceSynchronize());
    if (Lf && g->Nc > 0) T3D_HIP(hipMemcpy(Lf, g->L, sizeof(double) * size_t(N) * N, hipMemcpyDeviceToHost));
    return T3D_OK;
}

int t3d_mg_setup(t3d_mg* g, const double* E_dev)
{
    T3D_CHECK(g && E_dev, T3D_ERR_ARG, "t3d_mg_setup: null argument");
    t3d_problem* p = g->p;
    int rc = t3d_set_modulus(p, E_dev);
    if (rc) return rc;
    hipStream_t st = p->stream;
    const int B = 256;
    hipLaunchKernelGGL(k_diag_finest, dim3(unsigned((p->g.nN + B - 1) / B)), dim3(B), 0, st, p->g, p->Ke, p->Ecell,
                       g->mask[0], g->D[0]);
    T3D_HIP(hipGetLastError());
    if (g->mode == 1) {
        g->setupDone = false;
        g->Nc = 0;
        g->cholOk = 1;
        rc = mg_galerkin_setup(g);
        if (rc) return rc;
        g->setupDone = true;
        return T3D_OK;
    }
    for (int l = 1; l < g->nlev; ++l) {
        hipLaunchKernelGGL(k_diag_coarsen, dim3(unsigned((g->ld[l].nN + B - 1) / B)), dim3(B), 0, st, g->ld[l - 1],
                           g->ld[l], g->H.lv[l].span, g->W[l], g->D[l - 1], g->mask[l - 1], g->mask[l], g->D[l]);
        T3D_HIP(hipGetLastError());
    }
    // coarsest direct solve (preconditioner.cpp:48-79): assembled and factored on the host
    g->Nc = 0;
    g->cholOk = 1;
    g->Kc.clear();
    const int L = g->nlev - 1;
    const int N = int(3 * g->H.lv[L].nN);
    if (g->nlev > 1 && N <= g->H.NlimitDofs) {
        std::vector<double> E(p->ne);
        T3D_HIP(hipMemcpyAsync(E.data(), E_dev, sizeof(double) * p->ne, hipMemcpyDeviceToHost, st));
        T3D_HIP(hipStreamSynchronize(st));
        std::vector<double> A;
        mg_assemble_coarsest(p->host, g->H, E.data(), A);
        g->Kc = A;
        if (chol_factor(A, N)) {
            std::vector<double> AT(size_t(N) * N);
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) AT[size_t(i) * N + j] = A[size_t(j) * N + i];
            // (the Galerkin form allocates L of the same size but no LT)
            if (!g->L) T3D_HIP(hipMalloc(&g->L, sizeof(double) * size_t(N) * N));
            if (!g->LT) T3D_HIP(hipMalloc(&g->LT, sizeof(double) * size_t(N) * N));
            T3D_HIP(hipMemcpyAsync(g->L, A.data(), sizeof(double) * size_t(N) * N, hipMemcpyHostToDevice, st));
            T3D_HIP(hipMemcpyAsync(g->LT, AT.data(), sizeof(double) * size_t(N) * N, hipMemcpyHostToDevice, st));
            T3D_HIP(hipStreamSynchronize(st));   // A and AT are host buffers of this scope
            g->Nc = N;
        } else {
            g->cholOk = 0;   // diagonal fallback, as the reference (preconditioner.cpp:76-77)
        }
    }
    T3D_HIP(hipStreamSynchronize(st));
    g->setupDone = true;
    return T3D_OK;
}

int t3d_mg_apply(t3d_mg* g, const double* r_dev, double* z_dev)
{
    T3D_CHECK(g && r_dev && z_dev, T3D_ERR_ARG, "t3d_mg_apply: null argument");
    T3D_CHECK(g->setupDone, T3D_ERR_STATE, "t3d_mg_apply: t3d_mg_setup not called");
    int rc = mg_apply_dev(g, r_dev, z_dev, nullptr);
    if (rc) return rc;
    T3D_HIP(hipStreamSynchronize(g->p->stream));
    return T3D_OK;
}

}  // extern "C"
