// t3d_internal.hpp -- internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/t3d_capi.h"
#include "model.hpp"

namespace t3d {

// ---- error handling: every C-ABI entry returns an int status ----
void set_error(const std::string& msg);
#define T3D_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::t3d::set_error(std::string(#call) + ": " + hipGetErrorString(_e));              \
            return T3D_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)
#define T3D_CHECK(cond, code, msg)                \
    do {                                          \
        if (!(cond)) {                            \
            ::t3d::set_error(msg);                \
            return (code);                        \
        }                                         \
    } while (0)

// canonical reduction geometry (identical to the oracle's definition):
// tile = 256 lanes x 4 items; lane l of tile t sums items t*1024 + k*256 + l
constexpr int CANON_LANES = 256;
constexpr int CANON_PER_LANE = 4;
constexpr int CANON_TILE = CANON_LANES * CANON_PER_LANE;

// the canonical tile tree over one workgroup of CANON_LANES lanes: lane
// partials halved pairwise (lane l adds lane l + w), the result in every thread
__device__ __forceinline__ double block_tree_256(double s, double* sh)
{
    const int l = threadIdx.x;
    sh[l] = s;
    __syncthreads();
#pragma unroll
    for (int w = CANON_LANES / 2; w >= 64; w >>= 1) {
        if (l < w) sh[l] = sh[l] + sh[l + w];
        __syncthreads();
    }
    // last 6 halvings inside wave 0: lane l adds lane l+w
    if (l < 64) {
        double v = sh[l];
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1) {
            double o = __shfl_down(v, w, 64);
            v = v + o;   // only lanes l < w are consumed; they add exactly s[l+w]
        }
        if (l == 0) sh[0] = v;
    }
    __syncthreads();
    return sh[0];
}

struct Grid {
    int resX, resY, resZ;   // cells
    int nnx, nny, nnz;      // nodes per axis
    int64_t nN;             // nodes
    int64_t ncell;
};

// XCD-aware tile order for 3D tiled kernels.  Workgroups are dispatched to
// the 8 XCDs round-robin (hardware block b runs on XCD b % 8), and each XCD
// has its own L2.  This maps hardware block b to a logical tile so that XCD k
// works through the k-th contiguous range of tiles (x fastest, then y, z):
// neighbouring tiles, which re-read each other's halos, share an L2.
constexpr unsigned NUM_XCD = 8;
struct TileId {
    unsigned x, y, z;
};
__device__ __forceinline__ TileId xcd_tile()
{
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    const unsigned q = nb / NUM_XCD, rem = nb % NUM_XCD;
    const unsigned k = b % NUM_XCD, j = b / NUM_XCD;
    const unsigned l = k * q + (k < rem ? k : rem) + j;   // a bijection of [0, nb)
    TileId t;
    t.x = l % gridDim.x;
    t.y = (l / gridDim.x) % gridDim.y;
    t.z = l / (gridDim.x * gridDim.y);
    return t;
}

}  // namespace t3d

// device-side problem context
struct t3d_mg;
struct t3d_problem {
    t3d::HostModel host;
    t3d::Grid g;
    int ne = 0;
    int64_t nF = 0;
    int device = 0;                     // the HIP device the problem was created on
    hipStream_t stream = nullptr;
    // device arrays
    double* Ke = nullptr;               // [576]
    uint8_t* nodeFixed = nullptr;       // [nN] bit c set = DOF c fixed
    int32_t* compactToCell = nullptr;   // [ne]
    int32_t* cellToCompact = nullptr;   // [ncell], -1 = void
    int32_t* freeDofIndex = nullptr;    // [nF]
    double* Ecell = nullptr;            // [ncell] modulus per structured cell (0 = void)
    float* EcellF = nullptr;            // [ncell] fp32 copy for the single precision operator
    double* bFull = nullptr;            // [3*nN] restrict_to_free(F) in the full layout
    // reduction scratch
    double* red = nullptr;              // partial sums
    int64_t redCap = 0;
    double* scal = nullptr;             // device scalars
    // PCG workspace (full layout), allocated on first use
    double *wr = nullptr, *wz = nullptr, *wp = nullptr, *wp2 = nullptr, *wAp = nullptr;
    double* pcgSc = nullptr;            // per-iteration scalars
    int pcgScCap = 0;
    int* pcgFlags = nullptr;            // [0] done, [1] iterations
    int* hostFlags = nullptr;           // pinned mirror
    // voxel-surface STL scratch (stl.hip), allocated on first use
    uint8_t* stlIns = nullptr;          // [ncell] inside flags in the volume's order
    uint64_t* stlOff = nullptr;         // [blocks + 1] face offsets per block of 256 voxels
    // OC bisection state and chunk aggregates of its exact sums (oc.hip)
    void* ocBis = nullptr;
    void* ocAgg = nullptr;
    int64_t ocAggCap = 0;
};

namespace t3d {
int ensure_red(t3d_problem* p, int64_t items);
// a problem with the cuboid's boundary conditions (bc == nullptr) or the given lists
int create_problem(int nely, int nelx, int nelz, const uint8_t* solid, const BcLists* bc, t3d_problem** out);
// canonical sum of n items produced by kernel `first` into p->red; final value to dst (device)
int canon_finish(t3d_problem* p, int64_t ntiles, double* dst, const int* done = nullptr);
// the same over the tile partials buf[0..ntiles), buf[ntiles..2 ntiles) the second buffer
int canon_finish_buf(t3d_problem* p, double* buf, int64_t ntiles, double* dst, const int* done = nullptr);
int canon_dot_dev(t3d_problem* p, const double* a, const double* b, int64_t n, double* dst,
                  const int* done = nullptr);
int launch_kxu_f64(t3d_problem* p, const double* u, double* y, const int* done = nullptr);
// p = z + rzNew / max(1e-30, rzOld) * pPrev into pOut (!= pPrev), then y = K p
int launch_kxu_f64_fusep(t3d_problem* p, const double* z, const double* pPrev, const double* rzNew,
                         const double* rzOld, double* pOut, double* y, const int* done, double* x = nullptr,
                         const double* denomOld = nullptr);
__global__ void k_copy_scalar(const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ done);
// z = M r; with rzDst also r.z by the canonical tree into that device scalar
int mg_apply_dev(t3d_mg* g, const double* r, double* z, const int* done, double* rzDst = nullptr);
int compliance_dev(t3d_problem* p, const double* E, const double* U, double* ce, double* Cdst);
int cobj_dev(t3d_problem* p, const double* E, const double* ce, const double* Csolid, double* dst);
int launch_simp(t3d_problem* p, const float* x, float E0, float Emin, double* E);
// position-keyed checksum of x[0..n) to dst (device), as the oracle's o_x_check
int x_check_dev(t3d_problem* p, int64_t n, const float* x, unsigned long long* dst);
int launch_fill_f64(t3d_problem* p, int64_t n, double v, double* a);
int launch_fill_f32(t3d_problem* p, int64_t n, float v, float* a);
int launch_sens(t3d_problem* p, const float* x, const double* ce, const double* Csolid, float E0, float Emin,
                float penal, float* xdc);
int launch_dc_post(t3d_problem* p, const float* filt, const float* x, float* dc);
int launch_oc(t3d_problem* p, const float* x, const float* dc, float V0, float move, float* xnew, float* scal,
              const uint8_t* passive = nullptr);
}  // namespace t3d
