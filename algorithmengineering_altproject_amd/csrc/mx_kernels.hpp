// mx_kernels.hpp -- device bodies of the mixed-precision solver's kernels.
// Each body takes its block coordinates explicitly (bx_, by_, bz_, the
// partial slot slot_ and the block count gdx_), so the same code serves the
// single-problem kernels (mixed.hip, block coordinates from the launch) and
// the batched kernels (batch.hip, block coordinates from a block map).
#pragma once
#include <hip/hip_runtime.h>

#include "t3d_internal.hpp"

namespace t3d {

// scalar slots of the mixed solver
enum {
    S_NORMB2 = 0, S_STOP = 1, S_RZ = 2, S_ZAP = 4, S_PAP = 5, S_ALPHA = 6, S_BETA = 7, S_RN2 = 8, S_RZNEW = 9,
    // the x increment of an iteration, dx += alpha p, is added by the next
    // iteration's operator kernel (which reads that p as p_prev) or by the
    // fold: S_ALPHA_DX is the alpha still to apply (0 after a fold), S_PLAST
    // which buffer (1: pA, 0: pB) holds the last direction
    S_ALPHA_DX = 10, S_PLAST = 11
};

// ---------------------------------------------------------------------------
// block reductions
// ---------------------------------------------------------------------------
// sum over a 256-thread block; the result is valid in thread 0; sh[4]
__device__ __forceinline__ double block_sum_256(double s, double* sh)
{
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return t;
}

// sum over an NT-thread block (NT a multiple of 64, at most 1024); valid in
// thread 0; sh[NT / 64]
template <int NT>
__device__ __forceinline__ double block_sum_n(double s, double* sh)
{
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        if (NT == 256) {
            t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        } else {
            for (int k = 0; k < NT / 64; ++k) t += sh[k];
        }
    }
    return t;
}

// thread t's sequential sum of partial[t], partial[t + 1024], ... (eight
// loads in flight, added in index order)
__device__ __forceinline__ double thread_partials_1024(const double* __restrict__ partial, unsigned m)
{
    double s = 0.0;
    unsigned i = threadIdx.x;
    for (; i + 7u * 1024u < m; i += 8u * 1024u) {
        double w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = partial[i + 1024u * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += w[u];
    }
    for (; i < m; i += 1024) s += partial[i];
    return s;
}

// sum of partial[0..m) in a fixed order by a 1024-thread block (valid in thread 0)
__device__ __forceinline__ double sum_partials_1024(const double* __restrict__ partial, unsigned m, double* sh)
{
    double s = thread_partials_1024(partial, m);
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < 16; ++k) t += sh[k];
    return t;
}

__device__ __forceinline__ unsigned linear_block()
{
    return (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

// ---------------------------------------------------------------------------
// FP32 operator.  Tile of 32 x 8 x NZ nodes per 256-thread block; the
// (NZ+2)-plane node halo and the (NZ+1)-plane cell halo are staged in LDS.
// Each thread owns one (x, y) node column and sweeps it in z with a rolling
// register window of the 3x3x3 neighbourhood, so every LDS value is read once
// per column; the 8 incident cells and Ke's corner rows are compile-time
// indices (Ke through scalar loads).
// Fused form: the staged values are p = z + beta*p_prev (the tile interior
// writes p), and the block's p.(K p) partial feeds alpha = rz / max(1e-30, p.Ap).
// ---------------------------------------------------------------------------
namespace kx {
constexpr int TX = 32, TY = 8;
template <int NZ>
struct Tile {
    static constexpr int UX = TX + 2, UY = TY + 2, UZ = NZ + 2, UN = UX * UY * UZ;
    static constexpr int EX = TX + 1, EY = TY + 1, EZ = NZ + 1, EN = EX * EY * EZ;
};
}  // namespace kx

__device__ __forceinline__ constexpr int corner_of(int dx, int dy, int dz)
{
    return (dy ? (dx ? 1 : 0) : (dx ? 2 : 3)) + 4 * dz;
}



// ---------------------------------------------------------------------------
// FP32 operator, Walsh form.  In the Walsh-Hadamard basis of a voxel's eight
// nodes, H[m][b] = (-1)^(m . o(b)) with o(b) the node's offset in the voxel,
// the element matrix K^ = (H x I3) Ke (H x I3)^T has 45 nonzeros (the constant
// modes carry no energy and the linear, bilinear and trilinear modes decouple
// under the 2x2x2 Gauss rule), and Ke u = (H x I3)^T (K^/64) (H x I3) u.
// Per voxel that is 45 FMAs plus butterflies instead of 576 FMAs.
//
// A 256-thread block owns 31 x 7 node columns and sweeps NZ node planes in z;
// thread (tx, ty) owns the voxel column at (x0-1+tx, y0-1+ty).  The 2D Walsh
// transform of a node plane's 2x2 face is shared by the voxel below and the
// voxel above; the z-transform pairs consecutive faces.  Each voxel's face
// forces are accumulated per node plane in Walsh space (top face of the voxel
// below + bottom face of the voxel above), transformed back to the four
// corners, and summed over the four voxel columns that share a node through
// LDS, in a fixed order.
// ---------------------------------------------------------------------------
namespace wk {
constexpr int CX = 32, CY = 8;           // voxel columns (threads)
constexpr int NX = CX - 1, NY = CY - 1;  // owned node columns
constexpr int PX = CX + 1, PY = CY + 1;  // staged node plane
constexpr int PN = PX * PY;
constexpr int NNZ = 45;
// nonzero structure of K^, index (4 mz + 2 my + mx) * 3 + c; verified on the
// host against the element matrix when the operator is created
__device__ constexpr unsigned char OUT[NNZ] = {3,  3,  3,  4,  4,  5,  5,  6,  6,  7,  7,  7,  8,  8,  9,
                                               9,  10, 10, 11, 11, 11, 12, 12, 13, 13, 14, 14, 14, 15, 15,
                                               16, 16, 16, 17, 17, 18, 18, 18, 19, 19, 20, 20, 21, 22, 23};
__device__ constexpr unsigned char IN[NNZ] = {3,  7,  14, 4,  6,  5,  12, 4,  6,  3,  7,  14, 8,  13, 9,
                                              20, 10, 17, 11, 16, 18, 5,  12, 8,  13, 3,  7,  14, 15, 19,
                                              11, 16, 18, 10, 17, 11, 16, 18, 15, 19, 9,  20, 21, 22, 23};
}  // namespace wk

// 2D Walsh butterfly of the four corners a[oy][ox] -> F[my][mx] (self-inverse up to 4)
__device__ __forceinline__ void walsh2(float a00, float a10, float a01, float a11, float F[2][2])
{
    const float s0 = a00 + a10, d0 = a00 - a10, s1 = a01 + a11, d1 = a01 - a11;
    F[0][0] = s0 + s1;
    F[0][1] = d0 + d1;
    F[1][0] = s0 - s1;
    F[1][1] = d0 - d1;
}

// Raw buffer resource over `bytes` bytes at `base` (gfx9 resource word 3).
// Loads at offsets beyond the range return 0 and stores there are dropped;
// the Walsh-form operator addresses its grid edges this way instead of with
// branches, and keeps 32-bit offsets instead of 64-bit addresses.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes), 0x00020000);
}
constexpr uint32_t BUF_OOB = 0x80000000u;   // beyond every range used with buf_rsrc
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// ---------------------------------------------------------------------------
// MFMA form of the same operator (split bf16): a wavefront contracts two rows
// of 32 voxel columns, F (32 x 32) = Ke (24 x 24, padded) U (24 x 32), on the
// matrix cores as v_mfma_f32_32x32x16_bf16 with compensated products
// accumulated in FP32:
//   PIECES = 2 (bf16 x3): Ke = Kh + Kl, U = Uh + Ul (round to nearest even),
//     F ~ Kh Uh + Kh Ul + Kl Uh;
//   PIECES = 3 (bf16 x6): three pieces each, F ~ KhUh + KhUm + KmUh + KhUl +
//     KmUm + KlUh (24-bit operands, FP32-level accuracy).
// U is each voxel's corner values less their per-component mean (a rigid
// translation, which Ke maps to zero), so the operand rounding follows the
// variation of u over the voxel, not u itself.
// Operand layouts of the instruction on gfx950: A[i][k]: lane l holds row
// i = l % 32, columns k = 8 (l / 32) + e; B[k][j]: lane l holds k = 8 (l / 32) + e
// of voxel column j = l % 32; D[i][j]: lane l register r holds row
// i = 8 (r / 4) + 4 (l / 32) + r % 4 of column j = l % 32.  Input DOFs are
// component-major (k = 8 c + b for local node b); output rows are ordered so
// that each lane receives all eight nodes of its components: lanes l < 32
// components 0 and 1 (registers 8 j + b, j = c), lanes l >= 32 component 2
// (registers 0..7); rows beyond the 24 DOFs are zero.  Each lane therefore
// keeps the top-face forces of its voxel in registers for the layer above,
// as the Walsh form does.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

// node offset o(b) of local node b (fea.cpp node order); node 4 + m lies above node m
__host__ __device__ constexpr int node_off_x(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__host__ __device__ constexpr int node_off_y(int b) { return ((b & 3) <= 1) ? 1 : 0; }

// output DOF (component c, local node b) of MFMA row i, or c = -1 for a padding row
__host__ __device__ constexpr int mfma_row_c(int i)
{
    return ((i / 4) % 2 == 0) ? (i / 8) / 2 : ((i / 8) < 2 ? 2 : -1);
}
__host__ __device__ constexpr int mfma_row_b(int i) { return 4 * ((i / 8) % 2) + i % 4; }

// CYT: voxel rows of a tile (the workgroup is 32 x CYT threads).  PIECES = 0:
// the Walsh form (KW); PIECES = 2 or 3: the MFMA form (KA: the element
// matrix's bf16 pieces in the A layout, [K step][piece][lane][8]), which needs
// CYT = 16 (eight wavefronts of two voxel rows each).
template <int NZ, bool FUSED, int CYT = wk::CY, int PIECES = 0>
__device__ __forceinline__ void kxw32_body(Grid g, const float* __restrict__ KW, const float* __restrict__ Ecell,
                                               const uint8_t* __restrict__ nodeFixed, const float* __restrict__ u,
                                               const float* __restrict__ z, float* __restrict__ pcur,
                                               float* __restrict__ y, double* __restrict__ partial,
                                               const double* __restrict__ S, const int* __restrict__ done,
                                               float* __restrict__ dxv, int pid,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_,
    const unsigned short* __restrict__ KA = nullptr)
{
    using wk::IN;
    using wk::NNZ;
    using wk::OUT;
    constexpr int CX = wk::CX, CY = CYT;
    constexpr int NX = CX - 1, NY = CY - 1;
    constexpr int PX = CX + 1, PY = CY + 1, PN = PX * PY;
    constexpr int NT = CX * CY;
    constexpr int SPT = (3 * PN + NT - 1) / NT;   // staging slots per thread
    constexpr int SPN = 3 * PN + 64;              // staged plane + 64 write-only dummy slots
    static_assert(PIECES == 0 || CY == 2 * (NT / 64), "MFMA form: two voxel rows per wavefront");
    if (done && *done) return;   // the fused form always has a flag; the operator has one inside a solve
    __shared__ float sP[3][SPN];       // three staged node planes, [buffer][component * PN + node]
    __shared__ float sC[12][NT];       // face-corner forces [(2 oy + ox) * 3 + c][voxel column]
    __shared__ double sh[NT / 64];
    const int tid = threadIdx.x;
    const int tx = tid % CX, ty = tid / CX;
    // hardware tile order: measured faster here than the XCD-grouped order
    // (which halves the L2 misses but leaves z-neighbours, which share halo
    // planes, far apart in time)
    const int x0 = bx_ * NX, y0 = by_ * NY, z0 = bz_ * NZ;
    const float beta = FUSED ? float(S[S_BETA]) : 0.0f;
    const float alphaDx = FUSED ? float(S[S_ALPHA_DX]) : 0.0f;
    if (FUSED && slot_ == 0 && tid == 0) const_cast<double*>(S)[S_PLAST] = double(pid);
    const int cx = x0 - 1 + tx, cy = y0 - 1 + ty;   // this thread's voxel column (Walsh form)
    const bool colIn = cx >= 0 && cx < g.resX && cy >= 0 && cy < g.resY;
    const int X = x0 + tx, Y = y0 + ty;             // the node this thread gathers
    const bool ownNode = tx < NX && ty < NY && X < g.nnx && Y < g.nny;
    // MFMA form: lane (jx, kh) of wavefront wv takes voxel columns (jx, 2 wv + nb)
    const int lane = tid & 63, wv = tid >> 6, jx = lane & 31, kh = lane >> 5;
    // Node planes are staged in LDS; rows of PX nodes are 3*PX contiguous
    // floats in memory.  Each thread owns SPT slots of a staged plane, whose
    // byte offsets within a node plane do not depend on z: they are computed
    // once (BUF_OOB outside the grid, so those slots load 0).  A plane is
    // fetched into registers one step ahead (fetch) and written to LDS after
    // the arithmetic (commit), so its loads are in flight meanwhile; the fused
    // form forms p = z + beta * p_prev at the commit and stores the owned
    // nodes of p, and adds the previous iteration's alpha p_prev to the x
    // increment dx of the owned nodes.  The voxel modulus and the node's
    // fixed-DOF bits are also loaded one step ahead, so no step waits on a
    // load it has just issued.
    uint32_t voff[SPT];        // byte offset of slot j within a node plane
    uint32_t vown[SPT];        // the same for owned nodes (p store), BUF_OOB otherwise
    int slds[SPT];             // LDS index c * PN + row * PX + lx of slot j (a dummy slot if none)
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
        voff[j] = in ? uint32_t(4 * (3 * (Ys * g.nnx + Xs) + c)) : BUF_OOB;
        vown[j] = (in && lx >= 1 && lx <= NX && row >= 1 && row <= NY) ? voff[j] : BUF_OOB;
    }
    const int64_t plane = int64_t(3) * g.nnx * g.nny;
    const uint32_t planeB = uint32_t(4 * plane);
    const int64_t layer = int64_t(g.resX) * g.resY;
    // byte offsets of the moduli of the voxel columns this thread computes
    // (Walsh: its own; MFMA: (jx, 2 wv) and (jx, 2 wv + 1))
    uint32_t eoff[PIECES ? 2 : 1];
    if constexpr (PIECES == 0) {
        eoff[0] = colIn ? uint32_t(4 * (cy * g.resX + cx)) : BUF_OOB;
    } else {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int mx = x0 - 1 + jx, my = y0 - 1 + 2 * wv + nb;
            eoff[nb] = (mx >= 0 && mx < g.resX && my >= 0 && my < g.resY) ? uint32_t(4 * (my * g.resX + mx)) : BUF_OOB;
        }
    }
    constexpr int NE = PIECES ? 2 : 1;
    const uint32_t noff = ownNode ? uint32_t(Y * g.nnx + X) : BUF_OOB;
    auto plane_rsrc = [&](const float* base, int Z) {
        const bool zin = Z >= 0 && Z < g.nnz;
        return buf_rsrc(base + (zin ? Z * plane : 0), zin ? planeB : 0u);
    };
    float fu[SPT], fz[SPT], fd[SPT];
    auto fetch = [&](int Z) {
        const __amdgpu_buffer_rsrc_t ru = plane_rsrc(u, Z);
#pragma unroll
        for (int j = 0; j < SPT; ++j) fu[j] = buf_ld(ru, voff[j]);
        if (FUSED) {
            const __amdgpu_buffer_rsrc_t rz = plane_rsrc(z, Z);
#pragma unroll
            for (int j = 0; j < SPT; ++j) fz[j] = buf_ld(rz, voff[j]);
            // dx of the owned nodes, in the planes this block owns
            const bool own = Z >= z0 && Z < z0 + NZ && Z < g.nnz;
            const __amdgpu_buffer_rsrc_t rd = buf_rsrc(own ? dxv + Z * plane : dxv, own ? planeB : 0u);
#pragma unroll
            for (int j = 0; j < SPT; ++j) fd[j] = buf_ld(rd, vown[j]);
        }
    };
    auto commit = [&](int Z, int q) {
        const bool own = Z >= z0 && Z < z0 + NZ && Z < g.nnz;
        const __amdgpu_buffer_rsrc_t rp = buf_rsrc(FUSED && own ? pcur + Z * plane : pcur, FUSED && own ? planeB : 0u);
        const __amdgpu_buffer_rsrc_t rd = buf_rsrc(FUSED && own ? dxv + Z * plane : dxv, FUSED && own ? planeB : 0u);
        float* sq = &sP[q][0];
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            float v = fu[j];
            if (FUSED) {
                buf_st(fmaf(alphaDx, v, fd[j]), rd, vown[j]);
                v = fmaf(beta, v, fz[j]);
                buf_st(v, rp, vown[j]);
            }
            sq[slds[j]] = v;
        }
    };
    // voxel moduli of layer cz (0 outside the grid), unclamped
    auto modulus = [&](int cz, float (&Ev)[NE]) {
        const bool in = cz >= 0 && cz < g.resZ;
        const __amdgpu_buffer_rsrc_t re = buf_rsrc(Ecell + (in ? cz * layer : 0), in ? uint32_t(4 * layer) : 0u);
#pragma unroll
        for (int k = 0; k < NE; ++k) Ev[k] = buf_ld(re, eoff[k]);
    };
    // fixed-DOF bits of this thread's node in plane Z
    auto fixed_bits = [&](int Z) {
        const bool zin = Z >= 0 && Z < g.nnz;
        const int64_t pn = int64_t(g.nnx) * g.nny;
        return unsigned(__builtin_amdgcn_raw_buffer_load_b8(
            buf_rsrc(nodeFixed + (zin ? Z * pn : 0), zin ? uint32_t(pn) : 0u), noff, 0, 0));
    };
    // 2D Walsh transform of this voxel column's face in buffer q: F[c][my][mx]
    auto face = [&](int q, float F[3][2][2]) {
        const int b = ty * PX + tx;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            walsh2(sP[q][c * PN + b], sP[q][c * PN + b + 1], sP[q][c * PN + b + PX], sP[q][c * PN + b + PX + 1],
                   F[c]);
    };
    // MFMA form: the element matrix pieces (A operands) and, per voxel row
    // nb and component slot j (j = c for lanes < 32, j = 0 <-> c = 2 above),
    // the E-scaled top-face forces of the layer below
    bf16x8 A[2][PIECES ? PIECES : 1];
    float mcarry[2][2][4];
    if constexpr (PIECES > 0) {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int pc = 0; pc < PIECES; ++pc) {
                u16x8 h;
#pragma unroll
                for (int e = 0; e < 8; ++e) h[e] = KA[((ks * PIECES + pc) * 64 + lane) * 8 + e];
                A[ks][pc] = __builtin_bit_cast(bf16x8, h);
            }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int m = 0; m < 4; ++m) mcarry[nb][j][m] = 0.f;
    }
    float Fp[3][2][2], Fn[3][2][2], carry[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) (&carry[c][0][0])[k] = 0.f;
    fetch(z0 - 1);
    commit(z0 - 1, 0);
    fetch(z0);
    commit(z0, 1);
    float Enext[NE];
    modulus(z0 - 1, Enext);
    unsigned fxNext = 0u;
    __syncthreads();
    if constexpr (PIECES == 0) face(0, Fp);
    double dot = 0.0;
    const int nsteps = min(NZ, g.nnz - z0) + 1;
    for (int s = 0; s < nsteps; ++s) {
        const int qn = (s + 1) % 3, qo = s % 3;
        const bool more = s + 1 < nsteps;
        const int cz = z0 - 1 + s;
        // voxel layer cz and the fixed bits of node (X, Y, cz), loaded last step
        float E[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) E[k] = Enext[k] <= 1.0e-16f ? 0.f : Enext[k];   // fea.cpp:743-744
        const unsigned fx = fxNext;
        if (more) {   // loads of the next step, in flight during this one
            fetch(z0 + s + 1);
            modulus(cz + 1, Enext);
            fxNext = fixed_bits(cz + 1);
        }
        if constexpr (PIECES == 0) {
            face(qn, Fn);
            float uh[24], vh[24];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int my = 0; my < 2; ++my)
#pragma unroll
                    for (int mx = 0; mx < 2; ++mx) {
                        uh[(2 * my + mx) * 3 + c] = E[0] * (Fp[c][my][mx] + Fn[c][my][mx]);
                        uh[(4 + 2 * my + mx) * 3 + c] = E[0] * (Fp[c][my][mx] - Fn[c][my][mx]);
                    }
#pragma unroll
            for (int i = 0; i < 24; ++i) vh[i] = 0.f;
#pragma unroll
            for (int k = 0; k < NNZ; ++k) vh[OUT[k]] = fmaf(KW[k], uh[IN[k]], vh[OUT[k]]);
            // faces: bottom (oz = 0) = vh[mz=0] + vh[mz=1] closes plane cz; top carries to cz + 1
            float A4[3][2][2];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int my = 0; my < 2; ++my)
#pragma unroll
                    for (int mx = 0; mx < 2; ++mx) {
                        const float v0 = vh[(2 * my + mx) * 3 + c], v1 = vh[(4 + 2 * my + mx) * 3 + c];
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
                    float C4[2][2];
                    walsh2(A4[c][0][0], A4[c][0][1], A4[c][1][0], A4[c][1][1], C4);
                    // walsh2 returns [my][mx] slots; as the inverse they are corners [oy][ox]
                    sC[0 * 3 + c][tid] = C4[0][0];
                    sC[1 * 3 + c][tid] = C4[0][1];
                    sC[2 * 3 + c][tid] = C4[1][0];
                    sC[3 * 3 + c][tid] = C4[1][1];
                }
            }
        } else {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int tyv = 2 * wv + nb;   // voxel row of this MFMA tile
                f32x16 acc = {};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    // component of this lane's 8 input DOFs: ks 0 -> kh, ks 1 -> 2 (lanes 32..63: zero padding)
                    const int c = ks == 0 ? kh : 2;
                    const bool live = ks == 0 || kh == 0;
                    float uv[8];
                    float mean = 0.f;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        uv[b] = live ? sP[(b >> 2) ? qn : qo][c * PN + (tyv + node_off_y(b)) * PX + jx + node_off_x(b)]
                                     : 0.f;
                        mean += uv[b];
                    }
                    mean *= 0.125f;
                    u16x8 Bp[PIECES];
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const float v = uv[b] - mean;
                        const unsigned short h = bf16_rne(v);
                        Bp[0][b] = h;
                        const float r1 = v - bf16_val(h);
                        const unsigned short m = bf16_rne(r1);
                        Bp[1][b] = m;
                        if (PIECES == 3) Bp[PIECES - 1][b] = bf16_rne(r1 - bf16_val(m));
                    }
                    const bf16x8 B0 = __builtin_bit_cast(bf16x8, Bp[0]), B1 = __builtin_bit_cast(bf16x8, Bp[1]);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][0], B0, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][0], B1, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][1], B0, acc, 0, 0, 0);
                    if (PIECES == 3) {
                        const bf16x8 B2 = __builtin_bit_cast(bf16x8, Bp[PIECES - 1]);
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][0], B2, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][1], B1, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[ks][PIECES - 1], B0, acc, 0, 0, 0);
                    }
                }
                // registers 8 j + m: bottom node m, 8 j + 4 + m: top node 4 + m, of
                // component j (lanes < 32) or 2 (lanes >= 32, j = 0)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (kh == 1 && j == 1) continue;
                    const int c = kh == 0 ? j : 2;
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const int slot = (2 * node_off_y(m) + node_off_x(m)) * 3 + c;
                        sC[slot][tyv * CX + jx] = fmaf(E[nb], acc[8 * j + m], mcarry[nb][j][m]);
                        mcarry[nb][j][m] = E[nb] * acc[8 * j + 4 + m];
                    }
                }
            }
        }
        if (s > 0) {   // plane Z = cz: the node gather (plane z0 - 1 belongs to the block below)
            const int Z = cz;
            __syncthreads();
            if (ownNode) {
                // node (X, Y) collects corner (ox, oy) of voxel column (X - ox, Y - oy)
                const __amdgpu_buffer_rsrc_t ry = plane_rsrc(y, Z);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float c00 = sC[3 * 3 + c][ty * CX + tx];              // column (X-1, Y-1), corner (1,1)
                    const float c10 = sC[2 * 3 + c][ty * CX + tx + 1];          // column (X,   Y-1), corner (0,1)
                    const float c01 = sC[1 * 3 + c][(ty + 1) * CX + tx];        // column (X-1, Y),   corner (1,0)
                    const float c11 = sC[0 * 3 + c][(ty + 1) * CX + tx + 1];    // column (X,   Y),   corner (0,0)
                    const float yv = (c00 + c10) + (c01 + c11);
                    const float v = ((fx >> c) & 1u) ? 0.f : yv;
                    buf_st(v, ry, 4 * (3 * noff + c));
                    if (FUSED) dot = fma(double(sP[qo][c * PN + (ty + 1) * PX + tx + 1]), double(v), dot);
                }
            }
        }
        if (more) commit(z0 + s + 1, (s + 2) % 3);
        __syncthreads();
    }
    if (FUSED) {
        const double t = block_sum_n<CX * CY>(dot, sh);
        if (tid == 0) partial[slot_] = t;
    }
}



// alpha = rz / max(1e-30, p.Ap)
__device__ __forceinline__ void fin_alpha_body(const double* __restrict__ partial, unsigned m,
                                                    double* __restrict__ S, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (*done) return;
    __shared__ double sh[16];
    const double pAp = sum_partials_1024(partial, m, sh);
    if (threadIdx.x == 0) {
        S[S_PAP] = pAp;
        S[S_ALPHA] = S[S_RZ] / (1.0e-30 < pAp ? pAp : 1.0e-30);
        S[S_ALPHA_DX] = S[S_ALPHA];
    }
}

// ---------------------------------------------------------------------------
// FP64 vector kernels (grid-stride, fixed grid; the last block finishes)
// ---------------------------------------------------------------------------
constexpr int VB = 2048;   // blocks of the vector kernels

// rn2 = r.r ; if check: the convergence test of iteration it
__device__ __forceinline__ void fin_rn2_body(const double* __restrict__ partial, unsigned m,
                                                  double* __restrict__ S, int* __restrict__ flags, int it, int check,
                                                  int guard,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (guard && flags[0]) return;
    __shared__ double sh[16];
    const double rn2 = sum_partials_1024(partial, m, sh);
    if (threadIdx.x == 0) {
        S[S_RN2] = rn2;
        if (check && !flags[0] && sqrt(rn2) < S[S_STOP]) {
            flags[0] = 1;
            flags[1] = it + 1;
        }
    }
}









// ---------------------------------------------------------------------------
// FP32 V-cycle kernels.  3D launches (x in blocks of 64 threads, y and z in
// the grid) so node coordinates need no division.
// ---------------------------------------------------------------------------
struct LevelF {
    int resX, resY, resZ;
    int nnx, nny, nnz;
    int64_t nN;
};

// restriction pass 1 (x): t[(fz*nny+fy)*cnx+cx] = sum_dx W(dx) rf[fz,fy,cx*s+dx] (unnormalised)
template <class TIn>
__device__ __forceinline__ void restrict_x_body(LevelF Lf, int cnx, int span, const TIn* __restrict__ rf,
                                                   float* __restrict__ t, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (done && *done) return;
    const int cx = bx_ * 64 + threadIdx.x;
    if (cx >= cnx) return;
    const int fy = by_, fz = bz_;
    const float inv = 1.0f / float(span);
    const int64_t row = (int64_t(fz) * Lf.nny + fy) * Lf.nnx;
    const int f0 = cx * span;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f;
    for (int d = -(span - 1); d <= span - 1; ++d) {
        const int fx = f0 + d;
        if (fx < 0 || fx >= Lf.nnx) continue;
        const float w = 1.0f - float(d < 0 ? -d : d) * inv;
        const int64_t i = 3 * (row + fx);
        v0 = fmaf(w, float(rf[i]), v0);
        v1 = fmaf(w, float(rf[i + 1]), v1);
        v2 = fmaf(w, float(rf[i + 2]), v2);
    }
    const int64_t o = 3 * ((int64_t(fz) * Lf.nny + fy) * cnx + cx);
    t[o] = v0;
    t[o + 1] = v1;
    t[o + 2] = v2;
}

// the end of the restriction of coarse node (cx, cy, cz): normalisation by
// the in-range weight sum (norm = 1), coarse mask, stores
__device__ __forceinline__ void restrict_finish(LevelF Lc, int cx, int cy, int cz, float v0, float v1, float v2,
                                                float wxs, float wyz, const uint8_t* __restrict__ cmask,
                                                float* __restrict__ rc, const float* __restrict__ Dc,
                                                float* __restrict__ xc, double* __restrict__ rcd, int norm)
{
    const float ws = wxs * wyz;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    if (!norm) {   // Galerkin form: the plain transpose of the prolongation
        r0 = v0;
        r1 = v1;
        r2 = v2;
    } else if (ws > 1e-12f) {
        const float iw = 1.0f / ws;
        r0 = v0 * iw;
        r1 = v1 * iw;
        r2 = v2 * iw;
    }
    const int64_t n = (int64_t(cz) * Lc.nny + cy) * Lc.nnx + cx;
    const unsigned m = cmask[n];
    if (m & 1u) r0 = 0.f;
    if (m & 2u) r1 = 0.f;
    if (m & 4u) r2 = 0.f;
    rc[3 * n] = r0;
    rc[3 * n + 1] = r1;
    rc[3 * n + 2] = r2;
    if (rcd) {
        rcd[3 * n] = r0;
        rcd[3 * n + 1] = r1;
        rcd[3 * n + 2] = r2;
    }
    if (Dc) {
        xc[3 * n] = (m & 1u) ? 0.f : r0 * Dc[3 * n];
        xc[3 * n + 1] = (m & 2u) ? 0.f : r1 * Dc[3 * n + 1];
        xc[3 * n + 2] = (m & 4u) ? 0.f : r2 * Dc[3 * n + 2];
    }
}

// restriction pass 2 (y, z), normalisation by the in-range weight sum, coarse
// mask; at the coarsest level optionally the diagonal solve xc = rc / D or an
// FP64 copy of rc for the Cholesky solve (transfers.cpp:449-533).  SP > 0: the
// span as a compile-time constant, taps unrolled with buffer loads (a tap
// outside the fine grid loads 0 with weight 0, which leaves every sum exactly
// as the skipping loop of SP = 0 leaves it), so all of a thread's loads are in
// flight together.
template <int SP>
__device__ __forceinline__ void restrict_yz_t(LevelF Lf, LevelF Lc, int span, const float* __restrict__ t,
                                                    const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                                    const float* __restrict__ Dc, float* __restrict__ xc,
                                                    double* __restrict__ rcd, int norm, unsigned bx_, unsigned by_,
                                                    unsigned bz_)
{
    const int cx = bx_ * 64 + threadIdx.x;
    if (cx >= Lc.nnx) return;
    const int cy = by_, cz = bz_;
    const float inv = 1.0f / float(span);
    const int cnx = Lc.nnx;
    float wxs = 0.f;
    for (int d = -(span - 1); d <= span - 1; ++d) {
        const int fx = cx * span + d;
        if (fx >= 0 && fx < Lf.nnx) wxs += 1.0f - float(d < 0 ? -d : d) * inv;
    }
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, wyz = 0.f;
    constexpr int NTAP = SP > 0 ? 2 * SP - 1 : 1;
    if constexpr (SP > 0) {
        const __amdgpu_buffer_rsrc_t rt =
            buf_rsrc(t, uint32_t(4 * 3 * int64_t(Lf.nnz) * Lf.nny * cnx));
#pragma unroll
        for (int dz = -(SP - 1); dz <= SP - 1; ++dz) {
            const int fz = cz * SP + dz;
            const bool zin = fz >= 0 && fz < Lf.nnz;
            const float wz = 1.0f - float(dz < 0 ? -dz : dz) / float(SP);
            float a0[NTAP], a1[NTAP], a2[NTAP], ww[NTAP];
#pragma unroll
            for (int dy = -(SP - 1); dy <= SP - 1; ++dy) {
                const int fy = cy * SP + dy;
                const bool in = zin && fy >= 0 && fy < Lf.nny;
                const uint32_t off = in ? uint32_t(4 * 3 * ((fz * Lf.nny + fy) * cnx + cx)) : BUF_OOB;
                const int q = dy + SP - 1;
                a0[q] = buf_ld(rt, off);
                a1[q] = buf_ld(rt, off + 4);
                a2[q] = buf_ld(rt, off + 8);
                ww[q] = in ? wz * (1.0f - float(dy < 0 ? -dy : dy) / float(SP)) : 0.f;
            }
#pragma unroll
            for (int q = 0; q < NTAP; ++q) {
                v0 = fmaf(ww[q], a0[q], v0);
                v1 = fmaf(ww[q], a1[q], v1);
                v2 = fmaf(ww[q], a2[q], v2);
                wyz += ww[q];
            }
        }
    } else {
        for (int dz = -(span - 1); dz <= span - 1; ++dz) {
            const int fz = cz * span + dz;
            if (fz < 0 || fz >= Lf.nnz) continue;
            const float wz = 1.0f - float(dz < 0 ? -dz : dz) * inv;
            for (int dy = -(span - 1); dy <= span - 1; ++dy) {
                const int fy = cy * span + dy;
                if (fy < 0 || fy >= Lf.nny) continue;
                const float w = wz * (1.0f - float(dy < 0 ? -dy : dy) * inv);
                const int64_t i = 3 * ((int64_t(fz) * Lf.nny + fy) * cnx + cx);
                v0 = fmaf(w, t[i], v0);
                v1 = fmaf(w, t[i + 1], v1);
                v2 = fmaf(w, t[i + 2], v2);
                wyz += w;
            }
        }
    }
    restrict_finish(Lc, cx, cy, cz, v0, v1, v2, wxs, wyz, cmask, rc, Dc, xc, rcd, norm);
}


__device__ __forceinline__ void restrict_yz_body(LevelF Lf, LevelF Lc, int span, const float* __restrict__ t,
                                                    const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                                    const float* __restrict__ Dc, float* __restrict__ xc,
                                                    double* __restrict__ rcd, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_, int norm = 1)
{
    if (done && *done) return;
    if (span == 4)
        restrict_yz_t<4>(Lf, Lc, span, t, cmask, rc, Dc, xc, rcd, norm, bx_, by_, bz_);
    else if (span == 2)
        restrict_yz_t<2>(Lf, Lc, span, t, cmask, rc, Dc, xc, rcd, norm, bx_, by_, bz_);
    else
        restrict_yz_t<0>(Lf, Lc, span, t, cmask, rc, Dc, xc, rcd, norm, bx_, by_, bz_);
}

// Restriction of an FP32 level with span SP (2 or 4) in one pass: for every
// (y, z) tap, the x-pass value that restrict_x_body computes at (cx, fy, fz),
// then the (y, z) sum of restrict_yz_t<SP> -- the same operations in the same
// order, so the result is bit-identical to the two passes, without their
// intermediate array and second launch.  For the small coarse levels, whose
// passes take a launch's latency each.
template <int SP>
__device__ __forceinline__ void restrict_c32_t(LevelF Lf, LevelF Lc, const float* __restrict__ rf,
                                               const uint8_t* __restrict__ cmask, float* __restrict__ rc,
                                               const float* __restrict__ Dc, float* __restrict__ xc,
                                               double* __restrict__ rcd, int norm, unsigned bx_, unsigned by_,
                                               unsigned bz_)
{
    const int cx = bx_ * 64 + threadIdx.x;
    if (cx >= Lc.nnx) return;
    const int cy = by_, cz = bz_;
    const float inv = 1.0f / float(SP);
    float wxs = 0.f;
    for (int d = -(SP - 1); d <= SP - 1; ++d) {
        const int fx = cx * SP + d;
        if (fx >= 0 && fx < Lf.nnx) wxs += 1.0f - float(d < 0 ? -d : d) * inv;
    }
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, wyz = 0.f;
#pragma unroll
    for (int dz = -(SP - 1); dz <= SP - 1; ++dz) {
        const int fz = cz * SP + dz;
        const bool zin = fz >= 0 && fz < Lf.nnz;
        const float wz = 1.0f - float(dz < 0 ? -dz : dz) / float(SP);
#pragma unroll
        for (int dy = -(SP - 1); dy <= SP - 1; ++dy) {
            const int fy = cy * SP + dy;
            const bool in = zin && fy >= 0 && fy < Lf.nny;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
            if (in) {   // restrict_x_body at (cx, fy, fz)
                const int64_t row = (int64_t(fz) * Lf.nny + fy) * Lf.nnx;
#pragma unroll
                for (int d = -(SP - 1); d <= SP - 1; ++d) {
                    const int fx = cx * SP + d;
                    if (fx < 0 || fx >= Lf.nnx) continue;
                    const float w = 1.0f - float(d < 0 ? -d : d) * inv;
                    const int64_t i = 3 * (row + fx);
                    a0 = fmaf(w, rf[i], a0);
                    a1 = fmaf(w, rf[i + 1], a1);
                    a2 = fmaf(w, rf[i + 2], a2);
                }
            }
            const float ww = in ? wz * (1.0f - float(dy < 0 ? -dy : dy) / float(SP)) : 0.f;
            v0 = fmaf(ww, a0, v0);
            v1 = fmaf(ww, a1, v1);
            v2 = fmaf(ww, a2, v2);
            wyz += ww;
        }
    }
    restrict_finish(Lc, cx, cy, cz, v0, v1, v2, wxs, wyz, cmask, rc, Dc, xc, rcd, norm);
}

// trilinear prolongation of xc at fine node (fx, fy, fz), clamped at the
// upper faces (transfers.cpp:378-445)
struct Prolong {
    int cx, cy, cz;
    float wx0, wx1, wy0, wy1, wz0, wz1;
    __device__ Prolong(const LevelF& Lc, int span, int fx, int fy, int fz)
    {
        cx = fx / span;
        cy = fy / span;
        cz = fz / span;
        int rx = fx - cx * span, ry = fy - cy * span, rz = fz - cz * span;
        if (cx >= Lc.resX) { cx = Lc.resX - 1; rx = span; }
        if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = span; }
        if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = span; }
        const float inv = 1.0f / float(span);
        wx1 = float(rx) * inv; wx0 = 1.f - wx1;
        wy1 = float(ry) * inv; wy0 = 1.f - wy1;
        wz1 = float(rz) * inv; wz0 = 1.f - wz1;
    }
    __device__ void eval(const LevelF& Lc, const float* __restrict__ xc, float out[3]) const
    {
        const int64_t s = int64_t(Lc.nnx) * Lc.nny;
        const int64_t b00 = s * cz + int64_t(Lc.nnx) * cy + cx;
        const int64_t b10 = b00 + Lc.nnx, n00 = b00 + s, n10 = n00 + Lc.nnx;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float lo = wy0 * (wx0 * xc[3 * b00 + c] + wx1 * xc[3 * b00 + 3 + c]) +
                             wy1 * (wx0 * xc[3 * b10 + c] + wx1 * xc[3 * b10 + 3 + c]);
            const float hi = wy0 * (wx0 * xc[3 * n00 + c] + wx1 * xc[3 * n00 + 3 + c]) +
                             wy1 * (wx0 * xc[3 * n10 + c] + wx1 * xc[3 * n10 + 3 + c]);
            out[c] = wz0 * lo + wz1 * hi;
        }
    }
};

// coarse levels: x_f = mask ? 0 : 2t + P(x_c), t = w r/D  (transfers.cpp:378-445,
// preconditioner.cpp:230-253)
__device__ __forceinline__ void prolong_smooth32_body(LevelF Lf, LevelF Lc, int span, const float* __restrict__ xc,
                                                         const float* __restrict__ rf, const float* __restrict__ Df,
                                                         const uint8_t* __restrict__ fmask, float weight,
                                                         float* __restrict__ xf, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (done && *done) return;
    const int fx = bx_ * 64 + threadIdx.x;
    if (fx >= Lf.nnx) return;
    const int fy = by_, fz = bz_;
    const int64_t n = (int64_t(fz) * Lf.nny + fy) * Lf.nnx + fx;
    const Prolong P(Lc, span, fx, fy, fz);
    float v[3];
    P.eval(Lc, xc, v);
    const unsigned m = fmask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float out = 0.f;
        if (!((m >> c) & 1u)) {
            const float t = weight * rf[3 * n + c] * Df[3 * n + c];
            out = 2.f * t + v[c];
        }
        xf[3 * n + c] = out;
    }
}

// ---------------------------------------------------------------------------
// Finest-level row kernels.  A node row (fixed fy, fz) is 3*nnx contiguous
// values; a block takes RPB consecutive rows, so every load and store is a
// contiguous run (one value per lane), and the x-direction work of the
// transfers is done from LDS.
// ---------------------------------------------------------------------------
constexpr int RPB = 4;

// x-pass of the restriction for the block's rows staged in sR (FP32):
// t[row][cx][c] = sum_d W(d) sR[row][cx*span + d][c]
__device__ __forceinline__ void restrict_x_rows(const float* sR, int L, int nnx, int nr, int64_t row0, int cnx,
                                                int span, float* __restrict__ t)
{
    const int Lc = 3 * cnx;
    const float inv = 1.0f / float(span);
    for (int rr = 0; rr < nr; ++rr)
        for (int k = threadIdx.x; k < Lc; k += 256) {
            const int cxx = k / 3, c = k - 3 * cxx;
            float acc = 0.f;
            for (int d = -(span - 1); d <= span - 1; ++d) {
                const int fx = cxx * span + d;
                if (fx < 0 || fx >= nnx) continue;
                acc = fmaf(1.0f - float(d < 0 ? -d : d) * inv, sR[rr * L + 3 * fx + c], acc);
            }
            t[(row0 + rr) * Lc + k] = acc;
        }
}

// r -= alpha Ap (FP32 storage, FP64 alpha) ; partial r.r (the x increment
// dx += alpha p is applied by the next operator launch or the fold) ;
// optionally (t != nullptr) the x-pass of the restriction of the new r
template <int R>
__device__ __forceinline__ void xr_rows_body(int nnx, int NR, const float* __restrict__ Ap,
                                                 float* __restrict__ r, double* __restrict__ partial,
                                                 const double* __restrict__ S, int cnx, int span,
                                                 float* __restrict__ t, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (*done) return;
    extern __shared__ float sR[];
    __shared__ double sh[4];
    const int L = 3 * nnx;
    const int64_t row0 = int64_t(bx_) * R;
    const int nr = min(R, int(NR - row0));
    const double alpha = S[S_ALPHA];
    double s = 0.0;
    // the block's rows are one contiguous range starting at an even index
    // (R is even), so float2 accesses are aligned; where both arrays' range
    // starts on 16 bytes (R L a multiple of 4, or an even block), float4
    // accesses, two per thread and array issued before their arithmetic
    const int64_t base = row0 * L;
    const int ne = nr * L;
    int e0 = 0;
    if (((reinterpret_cast<uintptr_t>(Ap + base) | reinterpret_cast<uintptr_t>(r + base)) & 15) == 0) {
        const int n4 = ne / 4;
        const float4* A4 = reinterpret_cast<const float4*>(Ap + base);
        float4* R4 = reinterpret_cast<float4*>(r + base);
        for (int q0 = threadIdx.x; q0 < n4; q0 += 512) {
            float4 aa[2], rv[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int q = q0 + 256 * u;
                if (q < n4) {
                    aa[u] = A4[q];
                    rv[u] = R4[q];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int q = q0 + 256 * u;
                if (q < n4) {
                    const double v0 = fma(-alpha, double(aa[u].x), double(rv[u].x));
                    const double v1 = fma(-alpha, double(aa[u].y), double(rv[u].y));
                    const double v2 = fma(-alpha, double(aa[u].z), double(rv[u].z));
                    const double v3 = fma(-alpha, double(aa[u].w), double(rv[u].w));
                    const float4 o = make_float4(float(v0), float(v1), float(v2), float(v3));
                    R4[q] = o;
                    s = fma(v0, v0, s);
                    s = fma(v1, v1, s);
                    s = fma(v2, v2, s);
                    s = fma(v3, v3, s);
                    sR[4 * q] = o.x;
                    sR[4 * q + 1] = o.y;
                    sR[4 * q + 2] = o.z;
                    sR[4 * q + 3] = o.w;
                }
            }
        }
        e0 = 4 * n4;
    }
    for (int e = e0 + 2 * threadIdx.x; e < ne; e += 512) {
        const int64_t i = base + e;
        if (e + 1 < ne) {
            const float2 aa = *reinterpret_cast<const float2*>(Ap + i);
            float2 rv = *reinterpret_cast<const float2*>(r + i);
            const double v0 = fma(-alpha, double(aa.x), double(rv.x));
            const double v1 = fma(-alpha, double(aa.y), double(rv.y));
            rv.x = float(v0);
            rv.y = float(v1);
            *reinterpret_cast<float2*>(r + i) = rv;
            s = fma(v0, v0, s);
            s = fma(v1, v1, s);
            sR[e] = rv.x;
            sR[e + 1] = rv.y;
        } else {
            const double v = fma(-alpha, double(Ap[i]), double(r[i]));
            r[i] = float(v);
            s = fma(v, v, s);
            sR[e] = float(v);
        }
    }
    const double tot = block_sum_256(s, sh);
    if (threadIdx.x == 0) partial[bx_] = tot;
    if (t) {
        __syncthreads();
        restrict_x_rows(sR, L, nnx, nr, row0, cnx, span, t);
    }
}

// x-pass of the restriction of the finest-level residual (first V-cycle of a
// solve and after a residual replacement)
template <int R>
__device__ __forceinline__ void rx_rows_body(int nnx, int NR, const float* __restrict__ r, int cnx, int span,
                                                 float* __restrict__ t, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (done && *done) return;
    extern __shared__ float sR[];
    const int L = 3 * nnx;
    const int64_t row0 = int64_t(bx_) * R;
    const int nr = min(R, int(NR - row0));
    for (int rr = 0; rr < nr; ++rr)
        for (int k = threadIdx.x; k < L; k += 256) sR[rr * L + k] = r[(row0 + rr) * L + k];
    __syncthreads();
    restrict_x_rows(sR, L, nnx, nr, row0, cnx, span, t);
}

// finest level: z = mask ? 0 : 2t + P(x_c), t = w r/D,
// with the partials of r.z and z.Ap (two per block).  The y/z part of the
// trilinear prolongation is evaluated once per row and coarse x in LDS.
// TW: weight of the smoothing term t (2: the 'Adapted' cycle's pre- and
// post-smoothing from the same residual; 1: the pre-smoothed value of the
// 'Standard' cycle, whose post-smoothing follows with a fresh residual);
// PARTIALS: also the r.z and z.Ap partials
template <int R, int U = 8, int TW = 2, bool PARTIALS = true>
__device__ __forceinline__ void pro_rows_body(LevelF Lf, LevelF Lc, int span, const float* __restrict__ xc,
                                                  const float* __restrict__ rf, const float* __restrict__ Df,
                                                  const uint8_t* __restrict__ fmask, float weight,
                                                  float* __restrict__ xf, const float* __restrict__ Ap,
                                                  double* __restrict__ partial, const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (*done) return;
    extern __shared__ float sQ[];   // [R][3 * Lc.nnx]
    __shared__ double sh[4];
    const int NR = Lf.nny * Lf.nnz;
    const int L = 3 * Lf.nnx, Lq = 3 * Lc.nnx;
    const int64_t row0 = int64_t(bx_) * R;
    const int nr = min(R, int(NR - row0));
    const float inv = 1.0f / float(span);
    const int64_t cs = int64_t(Lc.nnx) * Lc.nny;
    for (int rr = 0; rr < nr; ++rr) {
        const int row = int(row0) + rr;
        const int fy = row % Lf.nny, fz = row / Lf.nny;
        int cy = fy / span, cz = fz / span;
        int ry = fy - cy * span, rz = fz - cz * span;
        if (cy >= Lc.resY) { cy = Lc.resY - 1; ry = span; }
        if (cz >= Lc.resZ) { cz = Lc.resZ - 1; rz = span; }
        const float wy1 = float(ry) * inv, wy0 = 1.f - wy1, wz1 = float(rz) * inv, wz0 = 1.f - wz1;
        const int64_t b0 = cs * cz + int64_t(Lc.nnx) * cy;
        for (int k = threadIdx.x; k < Lq; k += 256) {
            const int64_t q = 3 * b0 + k;
            const float lo = wy0 * xc[q] + wy1 * xc[q + Lq];
            const float hi = wy0 * xc[q + 3 * cs] + wy1 * xc[q + 3 * cs + Lq];
            sQ[rr * Lq + k] = wz0 * lo + wz1 * hi;
        }
    }
    __syncthreads();
    // Each thread owns the row positions k = k0 + tid + 256 u; their x-direction
    // interpolation (coarse column, weight) and component do not depend on the
    // row, so they are computed once per chunk of 256 U positions.  A row's
    // loads are buffer loads over exactly that row (positions k >= L load 0
    // and their stores are dropped), so they carry no branch and are all
    // issued before the row's arithmetic.  Within a chunk the partial sums run
    // over rows, then positions, as one pass over the block's rows.  U: chunks
    // of 256 positions per pass (a row of 3 nnx positions takes ceil(3 nnx /
    // (256 U)) passes; the chunk positions beyond the row load nothing).
    double srz = 0.0, szap = 0.0;
    for (int k0 = 0; k0 < L; k0 += 256 * U) {
        int qo[U], cc[U];
        uint32_t xs[U];
        float w1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + int(threadIdx.x) + u * 256;
            const int xx = k / 3, c = k - 3 * xx;
            int cxx = xx / span, rx = xx - cxx * span;
            if (cxx >= Lc.resX) { cxx = Lc.resX - 1; rx = span; }
            w1[u] = float(rx) * inv;
            qo[u] = 3 * cxx + c;
            cc[u] = c;
            xs[u] = uint32_t(xx);
        }
        for (int rr = 0; rr < nr; ++rr) {
            const int64_t nbase = (row0 + rr) * Lf.nnx;
            const float* q = sQ + rr * Lq;
            const uint32_t rowB = uint32_t(4 * L);
            const __amdgpu_buffer_rsrc_t rR = buf_rsrc(rf + 3 * nbase, rowB), rD = buf_rsrc(Df + 3 * nbase, rowB),
                                         rA = buf_rsrc(Ap + 3 * nbase, rowB), rX = buf_rsrc(xf + 3 * nbase, rowB),
                                         rM = buf_rsrc(fmask + nbase, uint32_t(Lf.nnx));
            float rv[U], dv[U], av[U];
            unsigned mv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t off = 4u * uint32_t(k0 + int(threadIdx.x) + u * 256);
                rv[u] = buf_ld(rR, off);
                dv[u] = buf_ld(rD, off);
                av[u] = PARTIALS ? buf_ld(rA, off) : 0.f;
                mv[u] = __builtin_amdgcn_raw_buffer_load_b8(rM, xs[u], 0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t off = 4u * uint32_t(k0 + int(threadIdx.x) + u * 256);
                const float v = (1.f - w1[u]) * q[qo[u]] + w1[u] * q[qo[u] + 3];
                float out = 0.f;
                if (!((mv[u] >> cc[u]) & 1u)) out = float(TW) * (weight * rv[u] * dv[u]) + v;
                buf_st(out, rX, off);
                if (PARTIALS) {   // beyond the row rv = av = 0: the partials gain exact zeros
                    srz = fma(double(rv[u]), double(out), srz);
                    szap = fma(double(out), double(av[u]), szap);
                }
            }
        }
    }
    if (PARTIALS) {
        const double a = block_sum_256(srz, sh);
        if (threadIdx.x == 0) partial[bx_] = a;
        const double b = block_sum_256(szap, sh);
        if (threadIdx.x == 0) partial[gdx_ + bx_] = b;
    }
}

// rz_new = r.z, zAp = z.Ap -> beta (0 on the first V-cycle of a solve); rz = rz_new
__device__ __forceinline__ void fin_beta_body(const double* __restrict__ partial, unsigned m,
                                                   double* __restrict__ S, int flexible, int first,
                                                   const int* __restrict__ done,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    if (*done) return;
    __shared__ double sh[16];
    const double rz = sum_partials_1024(partial, m, sh);
    const double zAp = sum_partials_1024(partial + m, m, sh);
    if (threadIdx.x == 0) {
        if (first) {
            S[S_BETA] = 0.0;
        } else {
            const double o = S[S_RZ];
            const double num = flexible ? -S[S_ALPHA] * zAp : rz;
            S[S_BETA] = num / (1.0e-30 < o ? o : 1.0e-30);
        }
        S[S_RZ] = rz;
        S[S_RZNEW] = rz;
        S[S_ZAP] = zAp;
    }
}



__device__ __forceinline__ void d2f_body(int64_t n, const double* __restrict__ a, float* __restrict__ b,
    unsigned bx_, unsigned by_, unsigned bz_, unsigned slot_, unsigned gdx_)
{
    const int64_t i = int64_t(bx_) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}


}  // namespace t3d
