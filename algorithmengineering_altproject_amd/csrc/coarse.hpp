// coarse.hpp -- the coarsest-level solve of the mixed-precision V-cycle
// through the inverse Cholesky factor, in FP64.
//
// The reference solves L L^T x = b by forward and backward substitution
// (coarsest.cpp:92-122), a chain of N dependent divisions that one workgroup
// must walk row by row.  The mixed solver instead keeps Ainv = Li^T Li with
// Li = L^-1 (both computed on the device once per multigrid set-up) and applies
//     x = Ainv b                    (x masked on output, as the substitution)
// as one dense matrix-vector product spread over many workgroups.  Ainv is
// symmetric, so only its lower triangle is kept, in 64 x 64 tiles (tile
// (I, J), J <= I, at index I (I + 1) / 2 + J, row-major inside, zero beyond
// N): each tile is read once and gives two partial products, A_IJ x_J for row
// block I and (I > J) A_IJ^T x_I for row block J, written to a partial array;
// a second pass sums each row block's partials in a fixed order.  Half the
// bytes of the full matrix, in two launches.  Same factor, same FP64
// arithmetic class; the rounding differs from the substitution at the level
// of the matrix's condition number times 2^-53.  The FP64 plain path keeps
// the reference's substitution (chol.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "chol.hpp"

namespace t3d {

constexpr int CT = 64;   // tile edge of the packed coarsest inverse

// row blocks of an N x N matrix in tiles of CT
__host__ __device__ inline int coarse_nt(int N) { return (N + CT - 1) / CT; }
// tiles of the packed lower triangle
__host__ __device__ inline unsigned coarse_tiles(int N)
{
    const unsigned nT = unsigned(coarse_nt(N));
    return nT * (nT + 1) / 2;
}
// doubles of the packed lower triangle, and of the partial products (nT x nT x CT)
inline size_t coarse_packed_doubles(int N) { return size_t(coarse_tiles(N)) * CT * CT; }
inline size_t coarse_partial_doubles(int N) { return size_t(coarse_nt(N)) * size_t(coarse_nt(N)) * CT; }

// (I, J), J <= I, of packed tile index t = I (I + 1) / 2 + J
__device__ __forceinline__ void coarse_tile_ij(unsigned t, int& I, int& J)
{
    int i = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
    while (unsigned(i) * unsigned(i + 1) / 2 > t) --i;
    while (unsigned(i + 1) * unsigned(i + 2) / 2 <= t) ++i;
    I = i;
    J = int(t - unsigned(i) * unsigned(i + 1) / 2);
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
// diagonal), tile (ti, tj), tj <= ti, of the packed lower triangle: 64 x 64
// entries per 256-thread workgroup, 4 x 4 per thread, k from the tile's first
// nonzero row in chunks of 32 staged in LDS; entries beyond N are written as 0.
__device__ __forceinline__ void gram_body(int N, const double* __restrict__ Li, double* __restrict__ Ap, int ti,
                                          int tj)
{
    __shared__ double Sa[32][65], Sb[32][65];
    const int i0 = CT * ti, j0 = CT * tj;
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
    double* T = Ap + (size_t(ti) * size_t(ti + 1) / 2 + size_t(tj)) * (CT * CT);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) T[(ri + a) * CT + rj + b] = acc[a][b];
}

// First pass of out = Ainv v over the packed tiles: tile (I, J) writes
// P[I][J] = A_IJ v_J and, for I > J, P[J][I] = A_IJ^T v_I, each a CT-vector
// at P + (row block * nT + col) * CT.  Each sum has a fixed order: the row
// products as four 16-term chains added pairwise across lanes, the column
// products as four 16-term chains added pairwise through LDS.
struct CoarseTile {
    int N;
    const double* Ap;   // packed lower tiles
    const double* v;
    double* P;
    unsigned tile;      // packed tile index
    bool on;            // false: the tile's problem is finished (nothing read or written)
};

// 256-thread workgroups, each walking tiles e = blockIdx.x, + gridDim.x, ...
// of a list of n (entry(e) describes tile e): a tile is staged in LDS from
// registers while the next tile's loads are in flight.
template <class F>
__device__ __forceinline__ void coarse_tiles_loop(unsigned n, F entry)
{
    __shared__ double sA[CT][CT + 1];
    __shared__ double sv[2][CT];
    __shared__ double sq[4][CT];
    const int t = int(threadIdx.x);
    unsigned e = blockIdx.x;
    if (e >= n) return;
    CoarseTile c = entry(e);
    double R[16];
    auto load = [&](const CoarseTile& x) {
        const double* T = x.Ap + size_t(x.tile) * (CT * CT);
#pragma unroll
        for (int k = 0; k < 16; ++k) R[k] = x.on ? T[t + 256 * k] : 0.0;
    };
    load(c);
    for (;;) {
#pragma unroll
        for (int k = 0; k < 16; ++k) sA[t / CT + 4 * k][t % CT] = R[k];
        int I, J;
        coarse_tile_ij(c.tile, I, J);
        if (t < CT) {
            const int j = CT * J + t;
            sv[0][t] = (c.on && j < c.N) ? c.v[j] : 0.0;
        } else if (t < 2 * CT) {
            const int i = CT * I + t - CT;
            sv[1][t - CT] = (c.on && i < c.N) ? c.v[i] : 0.0;
        }
        const unsigned en = e + gridDim.x;
        const bool more = en < n;
        CoarseTile cn = c;
        if (more) {   // the next tile's loads, in flight during this tile's arithmetic
            cn = entry(en);
            load(cn);
        }
        __syncthreads();
        const int nT = coarse_nt(c.N);
        {   // row r = t / 4, columns 16 q .. 16 q + 15 (q = t % 4)
            const int r = t / 4, q = t % 4;
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) a = fma(sA[r][16 * q + k], sv[0][16 * q + k], a);
            a += __shfl_xor(a, 1, 64);
            a += __shfl_xor(a, 2, 64);
            if (c.on && q == 0) c.P[(size_t(I) * nT + J) * CT + r] = a;
        }
        {   // column cc = t % 64, rows 16 h .. 16 h + 15 (h = t / 64)
            const int cc = t % CT, h = t / CT;
            double b = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) b = fma(sA[16 * h + k][cc], sv[1][16 * h + k], b);
            sq[h][cc] = b;
        }
        __syncthreads();
        if (c.on && I != J && t < CT) c.P[(size_t(J) * nT + I) * CT + t] = (sq[0][t] + sq[1][t]) + (sq[2][t] + sq[3][t]);
        __syncthreads();   // sA, sv, sq are restaged for the next tile
        if (!more) break;
        e = en;
        c = cn;
    }
}

// workgroups of the first pass (each loops over tiles; 4 per CU fit the LDS)
inline unsigned coarse_tile_blocks(unsigned n) { return n < 1024u ? n : 1024u; }

// Second pass: out[i] = sum over the row block's nT partials in order (J = 0,
// 1, ...), zero for masked DOFs, for the CT rows of block I (CT threads).
// With out32, an FP32 copy is written too.
__device__ __forceinline__ void coarse_sum_body(int N, const double* __restrict__ P, const uint8_t* __restrict__ mask,
                                                double* __restrict__ out, float* __restrict__ out32, int I)
{
    const int nT = coarse_nt(N);
    const int r = int(threadIdx.x);
    const int i = CT * I + r;
    if (r >= CT || i >= N) return;
    const double* Pi = P + size_t(I) * nT * CT + r;
    double a = 0.0;
    int J = 0;
    for (; J + 8 <= nT; J += 8) {   // eight loads in flight, added in order
        double w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = Pi[size_t(J + u) * CT];
#pragma unroll
        for (int u = 0; u < 8; ++u) a += w[u];
    }
    for (; J < nT; ++J) a += Pi[size_t(J) * CT];
    if (mask && ((mask[i / 3] >> (i % 3)) & 1u)) a = 0.0;
    out[i] = a;
    if (out32) out32[i] = float(a);
}

}  // namespace t3d
