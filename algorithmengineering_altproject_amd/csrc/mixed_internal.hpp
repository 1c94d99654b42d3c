// mixed_internal.hpp -- the mixed-precision solver's state and the helpers
// shared with the batched solver (batch.hip).
#pragma once
#include "coarse.hpp"
#include "mx_kernels.hpp"

// mixed-precision state attached to a multigrid object
struct t3d_mixed {
    t3d_problem* p = nullptr;
    t3d_mg* M = nullptr;
    float* Ke32 = nullptr;     // element matrix (node-gather form)
    float* KW = nullptr;       // the 45 nonzeros of the Walsh-form element matrix / 64
    double* KW64 = nullptr;    // the same in FP64 (the true residual)
    unsigned short* KA2 = nullptr;  // element matrix as 2-piece bf16 MFMA A operands
    unsigned short* KA3 = nullptr;  // ... and as 3-piece
    int form = 0;              // operator of the solver: 0 Walsh form, 2 MFMA bf16 x3, 3 MFMA bf16 x6
    int kwXcd = 0;             // tile order of the Walsh-form kernel: 0 hardware, 1 XCD-grouped
    float* D32[8] = {};        // FP32 inverse diagonals 1 / max(1e-30, D)
    float* r32[8] = {};
    float* x32[8] = {};
    float* tmp = nullptr;      // restriction pass-1 buffer
    double* rc64 = nullptr;    // coarsest residual / solution in FP64 for Cholesky
    double* xc64 = nullptr;
    double* Li = nullptr;      // inverse Cholesky factor L^-1 (row-major, lower; set-up only)
    double* Ainv = nullptr;    // Li^T Li, the coarsest inverse applied per V-cycle (packed lower tiles)
    bool ownAinv = true;       // false: Ainv is the multigrid object's (Galerkin form)
    double* cP = nullptr;      // partial products of the packed coarsest product
    size_t cPn = 0;            // doubles allocated at cP
    // 'Standard' V-cycle: FP32 Galerkin element matrices of levels 1 .. L-1
    // (entry-major: entry e of cell c at e * ncell + c) and per-level work
    // vectors A_l x_l / residuals (levels 0 .. L-1)
    float* Kel32[8] = {};
    size_t Kel32n[8] = {};
    float* ys32[8] = {};
    size_t ys32n[8] = {};
    double* yc64 = nullptr;    // Li * rc64
    int NcInv = 0;             // size of Li (0: diagonal coarsest level)
    float *pA = nullptr, *pB = nullptr, *Ap = nullptr, *z = nullptr;
    float* r = nullptr;        // recursive residual (FP32)
    float* dx = nullptr;       // FP32 increment of x since the last fold
    double* part = nullptr;    // block partials
    double* sc = nullptr;      // scalars
    int* iflags = nullptr;     // [0] done [1] iterations
    int hostFlags[4];
    int nlev = 0;
    // per-launch HIP-event timing of the fused operator kernel inside the solves
    bool timing = false;
    hipEvent_t tev[64] = {};
    double timedMs = 0.0;
    long long timedLaunches = 0;
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
    int norm;   // 1: the reference's normalised restriction; 0: the Galerkin form's P^T
    const double* Ainv;  // coarsest inverse Li^T Li when the multigrid keeps it
    int vtype;           // 0 'Adapted', 1 'Standard' (Galerkin form)
    const double* Kel[8];   // Galerkin element matrices of levels >= 1 ([cells][576]), else null
};
int mg_view(const t3d_mg* g, MGView& v);
__global__ void k_chol_solve(int N, const double* __restrict__ L, const double* __restrict__ LT,
                             const double* __restrict__ b, const uint8_t* __restrict__ mask, double* __restrict__ x,
                             const int* __restrict__ done);
}  // namespace t3d

inline t3d::LevelF levf(const t3d::MGView& v, int l)
{
    t3d::LevelF L;
    L.resX = v.resX[l]; L.resY = v.resY[l]; L.resZ = v.resZ[l];
    L.nnx = v.resX[l] + 1; L.nny = v.resY[l] + 1; L.nnz = v.resZ[l] + 1;
    L.nN = v.nN[l];
    return L;
}

namespace t3d {
constexpr int KW_NZ = 16;   // node planes per Walsh-form tile of the single-problem solver
constexpr int KW_CY = 16;   // voxel rows per Walsh-form tile (workgroup of 32 x KW_CY threads)
dim3 kww_grid(const Grid& g);   // tiles of the Walsh-form operator (32 x KW_CY voxel columns)
unsigned row_blocks(const LevelF& L);
// z32 = M32(r) with rz and beta (mixed.hip); tmpReady: the finest restriction's
// x-pass is already in s->tmp
int mixed_vcycle(t3d_mixed* s, const MGView& v, const float* r, int flexible, int first, bool tmpReady,
                 const int* done);
// single-problem steps of the batched solver (mixed.hip)
int mixed_true_residual(t3d_mixed* s, const double* b, const double* x);
int mixed_replace_residual(t3d_mixed* s, const double* b, double* x, int it);
int mixed_confirm(t3d_mixed* s, const double* b, double* x, double* rn2, double* stop);
int mixed_fold(t3d_mixed* s, double* x);
int mixed_resume(t3d_mixed* s, int itDone, int itNext, int flexible);
}  // namespace t3d
