// problem.hip -- device problem context, memory helpers, layout conversions,
// modulus scatter and the canonical reductions.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "t3d_internal.hpp"

namespace t3d {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// canonical reduction kernels (definition shared with the oracle)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_canon_dot_tiles(const double* __restrict__ a, const double* __restrict__ b,
                                                         int64_t n, double* __restrict__ out,
                                                         const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[CANON_LANES];
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) s = s + a[i] * b[i];
    }
    const double r = block_tree_256(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256) k_canon_sum_tiles(const double* __restrict__ v, int64_t n,
                                                         double* __restrict__ out, const int* __restrict__ done)
{
    if (done && *done) return;
    __shared__ double sh[CANON_LANES];
    const int64_t base = int64_t(blockIdx.x) * CANON_TILE + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CANON_PER_LANE; ++k) {
        const int64_t i = base + int64_t(k) * CANON_LANES;
        if (i < n) s = s + v[i];
    }
    const double r = block_tree_256(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

__global__ void k_copy_scalar(const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ done)
{
    if (done && *done) return;
    *dst = *src;
}

int ensure_red(t3d_problem* p, int64_t items)
{
    const int64_t tiles = (items + CANON_TILE - 1) / CANON_TILE;
    const int64_t need = 2 * tiles + 16;
    if (need <= p->redCap) return T3D_OK;
    if (p->red) T3D_HIP(hipFree(p->red));
    p->red = nullptr;
    T3D_HIP(hipMalloc(&p->red, sizeof(double) * need));
    p->redCap = need;
    return T3D_OK;
}

// reduces p->red[0..ntiles) (first-level tile partials) to one value at dst
int canon_finish(t3d_problem* p, int64_t ntiles, double* dst, const int* done)
{
    return canon_finish_buf(p, p->red, ntiles, dst, done);
}

// the same over buf[0..ntiles), with buf[ntiles..2 ntiles) as the second buffer
int canon_finish_buf(t3d_problem* p, double* buf, int64_t ntiles, double* dst, const int* done)
{
    double* src = buf;
    double* alt = buf + ntiles;
    int64_t m = ntiles;
    if (m <= 1) {
        hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, p->stream, src, dst, done);
        T3D_HIP(hipGetLastError());
        return T3D_OK;
    }
    while (m > 1) {
        const int64_t m2 = (m + CANON_TILE - 1) / CANON_TILE;
        // the last level's single tile writes the result itself
        hipLaunchKernelGGL(k_canon_sum_tiles, dim3(unsigned(m2)), dim3(CANON_LANES), 0, p->stream, src, m,
                           m2 == 1 ? dst : alt, done);
        T3D_HIP(hipGetLastError());
        std::swap(src, alt);
        m = m2;
    }
    return T3D_OK;
}

int canon_dot_dev(t3d_problem* p, const double* a, const double* b, int64_t n, double* dst, const int* done)
{
    if (n <= 0) {
        T3D_HIP(hipMemsetAsync(dst, 0, sizeof(double), p->stream));
        return T3D_OK;
    }
    int rc = ensure_red(p, n);
    if (rc) return rc;
    const int64_t tiles = (n + CANON_TILE - 1) / CANON_TILE;
    hipLaunchKernelGGL(k_canon_dot_tiles, dim3(unsigned(tiles)), dim3(CANON_LANES), 0, p->stream, a, b, n, p->red,
                       done);
    T3D_HIP(hipGetLastError());
    return canon_finish(p, tiles, dst, done);
}

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
__global__ void k_free_to_full(const double* __restrict__ vf, const int32_t* __restrict__ idx, int64_t nF,
                               double* __restrict__ full)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nF) full[idx[i]] = vf[i];
}
__global__ void k_full_to_free(const double* __restrict__ full, const int32_t* __restrict__ idx, int64_t nF,
                               double* __restrict__ vf)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nF) vf[i] = full[idx[i]];
}
__global__ void k_scatter_modulus(const double* __restrict__ E, const int32_t* __restrict__ c2cell, int ne,
                                  double* __restrict__ Ecell, float* __restrict__ EcellF)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) {
        const int c = c2cell[e];
        Ecell[c] = E[e];
        EcellF[c] = float(E[e]);
    }
}

}  // namespace t3d

using namespace t3d;

static inline unsigned nblk(int64_t n, int b) { return unsigned((n + b - 1) / b); }

extern "C" {

const char* t3d_last_error(void) { return g_last_error.c_str(); }

int t3d_device_info(char* name, size_t name_len, int* num_cus, size_t* total_mem)
{
    int dev = 0;
    T3D_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    T3D_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) {
        std::strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    return T3D_OK;
}

int t3d_malloc(void** ptr, size_t bytes)
{
    T3D_CHECK(ptr, T3D_ERR_ARG, "t3d_malloc: null ptr");
    T3D_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return T3D_OK;
}
int t3d_free(void* ptr)
{
    if (ptr) T3D_HIP(hipFree(ptr));
    return T3D_OK;
}
// The copy/fill helpers complete before they return: the problem streams are
// non-blocking and do not order against work on the null stream, and a
// null-stream hipMemset or pageable H2D copy can still be in flight when
// hipMemset/hipMemcpy return.
int t3d_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    T3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}
int t3d_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    T3D_HIP(hipDeviceSynchronize());
    T3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return T3D_OK;
}
int t3d_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    T3D_HIP(hipDeviceSynchronize());
    T3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}
int t3d_memset(void* dst, int value, size_t bytes)
{
    T3D_HIP(hipDeviceSynchronize());
    T3D_HIP(hipMemset(dst, value, bytes));
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}
int t3d_synchronize(void)
{
    T3D_HIP(hipDeviceSynchronize());
    return T3D_OK;
}

int t3d_problem_destroy(t3d_problem* p)
{
    if (!p) return T3D_OK;
    void* bufs[] = {p->Ke, p->nodeFixed, p->compactToCell, p->cellToCompact, p->freeDofIndex,
                    p->Ecell, p->EcellF, p->bFull, p->red, p->scal, p->wr, p->wz, p->wp, p->wp2, p->wAp, p->pcgSc,
                    p->pcgFlags, p->stlIns, p->stlOff, p->ocBis, p->ocAgg};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->hostFlags) (void)hipHostFree(p->hostFlags);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return T3D_OK;
}

int t3d_voxel_ke(double nu, double cellSize, double* Ke)
{
    T3D_CHECK(Ke, T3D_ERR_ARG, "t3d_voxel_ke: null output");
    const std::array<double, 576> k = voxel_ke(nu, cellSize);
    std::memcpy(Ke, k.data(), sizeof(double) * 576);
    return T3D_OK;
}

int t3d_problem_create_cuboid(int nely, int nelx, int nelz, t3d_problem** out)
{
    return t3d_problem_create_voxels(nely, nelx, nelz, nullptr, out);
}

int t3d_problem_create_voxels(int nely, int nelx, int nelz, const uint8_t* solid, t3d_problem** out)
{
    return create_problem(nely, nelx, nelz, solid, nullptr, out);
}

int t3d_problem_create_bc(int nely, int nelx, int nelz, const uint8_t* solid, int nFix, const int32_t* fix,
                          int nLoad, const int32_t* loadNode, const double* loadF, t3d_problem** out)
{
    T3D_CHECK(nFix >= 0 && nLoad >= 0 && (nFix == 0 || fix) && (nLoad == 0 || (loadNode && loadF)), T3D_ERR_ARG,
              "t3d_problem_create_bc: bad fixation or load list");
    BcLists bc;
    bc.nFix = nFix;
    bc.fix = fix;
    bc.nLoad = nLoad;
    bc.loadNode = loadNode;
    bc.loadF = loadF;
    return create_problem(nely, nelx, nelz, solid, &bc, out);
}

int t3d_problem_create_model(const t3d_voxel_model* vm, t3d_problem** out)
{
    T3D_CHECK(vm && vm->solid, T3D_ERR_ARG, "t3d_problem_create_model: null model");
    return t3d_problem_create_bc(vm->nely, vm->nelx, vm->nelz, vm->solid, vm->nFix, vm->fix, vm->nLoad,
                                 vm->loadNode, vm->loadF, out);
}

}  // extern "C"

namespace t3d {
// the padded voxel model with the cuboid's boundary conditions (bc == nullptr)
// or the given fixation and load lists, and its device mirror
int create_problem(int nely, int nelx, int nelz, const uint8_t* solid, const BcLists* bc, t3d_problem** out)
{
    T3D_CHECK(out && nely > 0 && nelx > 0 && nelz > 0, T3D_ERR_ARG, "t3d_problem_create: bad arguments");
    *out = nullptr;
    if (solid) {
        size_t cnt = 0;
        for (size_t i = 0; i < size_t(nely) * nelx * nelz; ++i) cnt += solid[i] != 0;
        T3D_CHECK(cnt > 0, T3D_ERR_ARG, "t3d_problem_create: no solid voxel");
    }
    t3d_problem* p = new t3d_problem();
    HostModel& m = p->host;
    build_voxels(m, nely, nelx, nelz, solid);
    if (bc) {
        std::string err;
        if (apply_bc_lists(m, *bc, err)) {
            delete p;
            set_error("t3d_problem_create_bc: " + err);
            return T3D_ERR_ARG;
        }
    } else {
        apply_bcs(m);
    }
    Grid& g = p->g;
    g.resX = m.resX; g.resY = m.resY; g.resZ = m.resZ;
    g.nnx = m.resX + 1; g.nny = m.resY + 1; g.nnz = m.resZ + 1;
    g.nN = int64_t(g.nnx) * g.nny * g.nnz;
    g.ncell = int64_t(g.resX) * g.resY * g.resZ;
    if (3 * g.nN >= (int64_t(1) << 31)) {
        delete p;
        set_error("t3d_problem_create_voxels: more than 2^31 DOFs");
        return T3D_ERR_ARG;
    }
    p->ne = m.numElements;
    p->nF = int64_t(m.freeDofIndex.size());
    auto fail = [&](int rc) { t3d_problem_destroy(p); return rc; };
    if (hipGetDevice(&p->device) != hipSuccess) {
        set_error("hipGetDevice failed");
        return fail(T3D_ERR_HIP);
    }
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(T3D_ERR_HIP);
    }
    std::vector<uint8_t> nodeFixed(g.nN);
    for (int64_t n = 0; n < g.nN; ++n)
        nodeFixed[n] = uint8_t((m.isFreeDOF[3 * n] ? 0 : 1) | (m.isFreeDOF[3 * n + 1] ? 0 : 2) |
                               (m.isFreeDOF[3 * n + 2] ? 0 : 4));
    std::vector<int32_t> c2cell(p->ne), cell2c(g.ncell, -1);
    for (int e = 0; e < p->ne; ++e) {
        c2cell[e] = cell_of_full(m.eleMapBack[e], m.resX, m.resY);
        cell2c[c2cell[e]] = e;
    }
    std::vector<double> bFull(3 * g.nN, 0.0);
    for (int64_t n = 0; n < g.nN; ++n) {
        if (m.isFreeDOF[3 * n]) bFull[3 * n] = m.Fx[n];
        if (m.isFreeDOF[3 * n + 1]) bFull[3 * n + 1] = m.Fy[n];
        if (m.isFreeDOF[3 * n + 2]) bFull[3 * n + 2] = m.Fz[n];
    }
#define ALLOC_UP(ptr, host, count)                                                                   \
    do {                                                                                             \
        if (hipMalloc(&ptr, sizeof(*ptr) * size_t(count ? count : 1)) != hipSuccess ||               \
            hipMemcpy(ptr, host, sizeof(*ptr) * size_t(count), hipMemcpyHostToDevice) != hipSuccess) { \
            set_error("device upload failed: " #ptr);                                                \
            return fail(T3D_ERR_HIP);                                                                \
        }                                                                                            \
    } while (0)
    ALLOC_UP(p->Ke, m.Ke.data(), 576);
    ALLOC_UP(p->nodeFixed, nodeFixed.data(), g.nN);
    ALLOC_UP(p->compactToCell, c2cell.data(), p->ne);
    ALLOC_UP(p->cellToCompact, cell2c.data(), g.ncell);
    ALLOC_UP(p->freeDofIndex, m.freeDofIndex.data(), p->nF);
    ALLOC_UP(p->bFull, bFull.data(), 3 * g.nN);
#undef ALLOC_UP
    if (hipMalloc(&p->Ecell, sizeof(double) * g.ncell) != hipSuccess ||
        hipMalloc(&p->EcellF, sizeof(float) * g.ncell) != hipSuccess ||
        hipMalloc(&p->scal, sizeof(double) * 64) != hipSuccess ||
        hipMemset(p->Ecell, 0, sizeof(double) * g.ncell) != hipSuccess ||
        hipMemset(p->EcellF, 0, sizeof(float) * g.ncell) != hipSuccess) {
        set_error("device allocation failed");
        return fail(T3D_ERR_HIP);
    }
    // the set-up copies and fills above ran on the null stream
    T3D_HIP(hipDeviceSynchronize());
    *out = p;
    return T3D_OK;
}
}  // namespace t3d

extern "C" {

int t3d_problem_sizes(const t3d_problem* p, int64_t* s)
{
    T3D_CHECK(p && s, T3D_ERR_ARG, "t3d_problem_sizes: null argument");
    s[0] = p->g.resX; s[1] = p->g.resY; s[2] = p->g.resZ;
    s[3] = p->ne; s[4] = p->g.nN; s[5] = 3 * p->g.nN; s[6] = p->nF;
    s[7] = p->g.nnx; s[8] = p->g.nny; s[9] = p->g.nnz;
    return T3D_OK;
}

int t3d_problem_model(const t3d_problem* p, int32_t* eleMapBack, int32_t* eNodMat, int32_t* elemColor, double* Ke,
                      double* F, uint8_t* isFree, int32_t* freeDofIndex, int32_t* eDofFreeMat)
{
    T3D_CHECK(p, T3D_ERR_ARG, "t3d_problem_model: null problem");
    const HostModel& m = p->host;
    if (eleMapBack) std::memcpy(eleMapBack, m.eleMapBack.data(), sizeof(int32_t) * m.eleMapBack.size());
    if (eNodMat) std::memcpy(eNodMat, m.eNodMat.data(), sizeof(int32_t) * m.eNodMat.size());
    if (elemColor) std::memcpy(elemColor, m.elemColor.data(), sizeof(int32_t) * m.elemColor.size());
    if (Ke) std::memcpy(Ke, m.Ke.data(), sizeof(double) * 576);
    if (F)
        for (int n = 0; n < m.numNodes; ++n) { F[3 * n] = m.Fx[n]; F[3 * n + 1] = m.Fy[n]; F[3 * n + 2] = m.Fz[n]; }
    if (isFree) std::memcpy(isFree, m.isFreeDOF.data(), m.isFreeDOF.size());
    if (freeDofIndex) std::memcpy(freeDofIndex, m.freeDofIndex.data(), sizeof(int32_t) * m.freeDofIndex.size());
    if (eDofFreeMat) std::memcpy(eDofFreeMat, m.eDofFreeMat.data(), sizeof(int32_t) * m.eDofFreeMat.size());
    return T3D_OK;
}

int t3d_free_to_full(t3d_problem* p, const double* vf, double* full)
{
    T3D_CHECK(p && vf && full, T3D_ERR_ARG, "t3d_free_to_full: null argument");
    T3D_HIP(hipMemsetAsync(full, 0, sizeof(double) * 3 * p->g.nN, p->stream));
    hipLaunchKernelGGL(k_free_to_full, dim3(nblk(p->nF, 256)), dim3(256), 0, p->stream, vf, p->freeDofIndex, p->nF,
                       full);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_full_to_free(t3d_problem* p, const double* full, double* vf)
{
    T3D_CHECK(p && vf && full, T3D_ERR_ARG, "t3d_full_to_free: null argument");
    hipLaunchKernelGGL(k_full_to_free, dim3(nblk(p->nF, 256)), dim3(256), 0, p->stream, full, p->freeDofIndex, p->nF,
                       vf);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_load_vector(t3d_problem* p, double* b)
{
    T3D_CHECK(p && b, T3D_ERR_ARG, "t3d_load_vector: null argument");
    T3D_HIP(hipMemcpyAsync(b, p->bFull, sizeof(double) * 3 * p->g.nN, hipMemcpyDeviceToDevice, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_set_modulus(t3d_problem* p, const double* E)
{
    T3D_CHECK(p && E, T3D_ERR_ARG, "t3d_set_modulus: null argument");
    hipLaunchKernelGGL(k_scatter_modulus, dim3(nblk(p->ne, 256)), dim3(256), 0, p->stream, E, p->compactToCell,
                       p->ne, p->Ecell, p->EcellF);
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

int t3d_dot_f64(t3d_problem* p, const double* a, const double* b, int64_t n, double* out)
{
    T3D_CHECK(p && a && b && out && n >= 0, T3D_ERR_ARG, "t3d_dot_f64: bad argument");
    int rc = canon_dot_dev(p, a, b, n, p->scal);
    if (rc) return rc;
    T3D_HIP(hipMemcpyAsync(out, p->scal, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    T3D_HIP(hipStreamSynchronize(p->stream));
    return T3D_OK;
}

}  // extern "C"
