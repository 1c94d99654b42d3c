// stl.hip -- binary STL image of a design's voxel surface, built on the device.
//
// What it produces (byte for byte the file of cpp/src/io.cpp:80-87
// export_surface_stl):
//   * the density xPhys (compact element order) placed in the padded volume
//     vol[y + ny x + ny nx z] (ny = resY, nx = resX, nz = resZ, y the flipped
//     row index of the element's full index, fea.cpp:206-215), 0 on cells
//     without an element (io.cpp:81-83);
//   * a voxel is inside when vol >= iso; for every voxel in the order z, x, y
//     (outer to inner) and every direction -X, +X, -Y, +Y, -Z, +Z, an inside
//     voxel whose neighbour in that direction is outside or beyond the volume
//     contributes one square face of two triangles
//     (geometry_out/voxel_surface.hpp:12-93);
//   * the binary STL layout of geometry_out/export_stl.hpp:14-33: an 80-byte
//     header holding "Generated by TOP3D_XL" and zeros, the triangle count as
//     a little-endian uint32, then per triangle a zero normal, three vertices
//     (x, y, z floats) and a zero uint16 attribute -- 50 bytes.
//
// Device form: inside bytes for the whole volume (one pass over the cells,
// one over the elements), a count pass (faces per block of 2048 consecutive
// voxels in the volume's own order, 8 per thread), one workgroup's exclusive
// scan of the block counts, and an emit pass.  A face is 100 bytes, so face k
// starts at byte 84 + 100 k, a multiple of 4: 25 dwords, the second
// triangle's vertex floats shifted by two bytes.  The faces of a block are one
// contiguous range of the image; the emit pass stages them in LDS and writes
// the range with consecutive lanes on consecutive dwords.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "t3d_internal.hpp"

namespace t3d {

constexpr int STL_B = 256;                // threads of the count and emit workgroups
constexpr int STL_VT = 8;                 // consecutive voxels per thread
constexpr int STL_BV = STL_B * STL_VT;    // voxels per block
constexpr int STL_SB = 384;               // faces staged in LDS per round of the emit pass
constexpr int STL_SCAN = 1024;            // threads of the scan workgroup

struct StlVol {
    int ny, nx, nz;
    int64_t n;
};

__global__ void k_stl_fill(int64_t n, uint8_t val, uint8_t* __restrict__ ins)
{
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += int64_t(gridDim.x) * blockDim.x)
        ins[v] = val;
}

// ins[v] for the voxel of element e: the element's cell (z, y, x) in the
// solver's structured order (x fastest) moved to the volume's order (y fastest)
__global__ void k_stl_mark(int ne, const int32_t* __restrict__ compactToCell, int resX, int resY,
                           const float* __restrict__ xPhys, float iso, uint8_t* __restrict__ ins)
{
    const int e = int(blockIdx.x) * blockDim.x + int(threadIdx.x);
    if (e >= ne) return;
    const int c = compactToCell[e];
    const int ex = c % resX, ty = (c / resX) % resY, ez = c / (resX * resY);
    const int64_t v = int64_t(ty) + int64_t(resY) * (int64_t(ex) + int64_t(resX) * ez);
    ins[v] = xPhys[e] >= iso ? 1 : 0;
}

// face bits of the STL_VT voxels from v0 (6 bits each, bit k of a voxel for
// direction k of -X, +X, -Y, +Y, -Z, +Z); *cnt = their number of faces.  v0 is
// a multiple of STL_VT, so the voxels' own inside bytes are one aligned 8-byte
// load (the inside array is padded to whole blocks).  Cell counts fit in 32 bits.
__device__ __forceinline__ uint64_t stl_masks(const uint8_t* __restrict__ ins, const StlVol& V, int64_t v0,
                                              unsigned* cnt)
{
    *cnt = 0u;
    if (v0 >= V.n) return 0ull;
    const uint64_t own = *reinterpret_cast<const uint64_t*>(ins + v0);
    if (!own) return 0ull;
    int v = int(v0);
    const int pl = V.ny * V.nx;
    int y = v % V.ny, x = (v / V.ny) % V.nx, z = v / pl;
    uint64_t all = 0ull;
    unsigned c = 0u;
#pragma unroll
    for (int i = 0; i < STL_VT; ++i) {
        if (int64_t(v) < V.n && ((own >> (8 * i)) & 0xffull)) {
            unsigned m = 0u;
            if (x == 0 || !ins[v - V.ny]) m |= 1u;
            if (x == V.nx - 1 || !ins[v + V.ny]) m |= 2u;
            if (y == 0 || !ins[v - 1]) m |= 4u;
            if (y == V.ny - 1 || !ins[v + 1]) m |= 8u;
            if (z == 0 || !ins[v - pl]) m |= 16u;
            if (z == V.nz - 1 || !ins[v + pl]) m |= 32u;
            all |= uint64_t(m) << (6 * i);
            c += unsigned(__popc(m));
        }
        ++v;
        if (++y == V.ny) {
            y = 0;
            if (++x == V.nx) {
                x = 0;
                ++z;
            }
        }
    }
    *cnt = c;
    return all;
}

// exclusive prefix of c over the workgroup (STL_B threads); *total = the sum
__device__ __forceinline__ unsigned block_excl_scan(unsigned c, unsigned* total)
{
    __shared__ unsigned wsum[STL_B / 64];
    const int lane = int(threadIdx.x) & 63, w = int(threadIdx.x) >> 6;
    unsigned inc = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned before = 0u, all = 0u;
#pragma unroll
    for (int k = 0; k < STL_B / 64; ++k) {
        before += (k < w) ? wsum[k] : 0u;
        all += wsum[k];
    }
    *total = all;
    return before + inc - c;
}

__global__ void __launch_bounds__(STL_B) k_stl_count(StlVol V, const uint8_t* __restrict__ ins,
                                                     uint64_t* __restrict__ blk)
{
    const int64_t v0 = int64_t(blockIdx.x) * STL_BV + int64_t(threadIdx.x) * STL_VT;
    unsigned c;
    (void)stl_masks(ins, V, v0, &c);
    unsigned total;
    (void)block_excl_scan(c, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// exclusive scan of blk[0 .. nb) in place, blk[nb] = the sum; one workgroup
__global__ void __launch_bounds__(STL_SCAN) k_stl_scan(int64_t nb, uint64_t* __restrict__ blk)
{
    __shared__ uint64_t part[STL_SCAN];
    const int t = int(threadIdx.x);
    const int64_t per = (nb + STL_SCAN - 1) / STL_SCAN;
    const int64_t b0 = t * per, b1 = (b0 + per < nb) ? b0 + per : nb;
    uint64_t s = 0;
    for (int64_t b = b0; b < b1; ++b) s += blk[b];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < STL_SCAN; o <<= 1) {   // inclusive Hillis-Steele scan of the chunk sums
        const uint64_t add = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint64_t run = part[t] - s;
    for (int64_t b = b0; b < b1; ++b) {
        const uint64_t c = blk[b];
        blk[b] = run;
        run += c;
    }
    if (t == STL_SCAN - 1) blk[nb] = part[t];
}

// the 100 bytes of one face as 25 dwords: triangles (v0, v1, v2) and (v0, v2, v3)
__device__ __forceinline__ void stl_put_face(uint32_t* __restrict__ w, int kind, int y, int x, int z)
{
    // add_face(axis, y0, x0, z0) of voxel_surface.hpp:19-45 for the six directions
    const int axis = kind >> 1;
    int X = x, Y = y, Z = z;
    if (kind == 1) X += 1;
    if (kind == 3) Y += 1;
    if (kind == 5) Z += 1;
    // v1 = v0 + a, v2 = v0 + a + b, v3 = v0 + b
    const int ax = axis == 0 ? 0 : 1, ay = axis == 0 ? 1 : 0;
    const int bx = 0, by = axis == 2 ? 1 : 0, bz = axis == 2 ? 0 : 1;
    const uint32_t x0 = __float_as_uint(float(X)), y0 = __float_as_uint(float(Y)), z0 = __float_as_uint(float(Z));
    const uint32_t x1 = __float_as_uint(float(X + ax)), y1 = __float_as_uint(float(Y + ay));
    const uint32_t x2 = __float_as_uint(float(X + ax + bx)), y2 = __float_as_uint(float(Y + ay + by)),
                   z2 = __float_as_uint(float(Z + bz));
    const uint32_t x3 = __float_as_uint(float(X + bx)), y3 = __float_as_uint(float(Y + by)),
                   z3 = __float_as_uint(float(Z + bz));
    // first triangle: normal (dwords 0-2), v0, v1, v2, attribute (low half of dword 12)
    w[0] = 0u; w[1] = 0u; w[2] = 0u;
    w[3] = x0; w[4] = y0; w[5] = z0;
    w[6] = x1; w[7] = y1; w[8] = z0;
    w[9] = x2; w[10] = y2; w[11] = z2;
    // second triangle from byte 50: normal to byte 61, then v0, v2, v3 two bytes off the dword grid
    const uint32_t s[9] = {x0, y0, z0, x2, y2, z2, x3, y3, z3};
    w[12] = 0u; w[13] = 0u; w[14] = 0u;
    w[15] = s[0] << 16;
#pragma unroll
    for (int i = 1; i < 9; ++i) w[15 + i] = (s[i - 1] >> 16) | (s[i] << 16);
    w[24] = s[8] >> 16;
}

// each thread's faces go to LDS in rounds of STL_SB faces of the block (a face
// to its slot of 25 dwords), and the whole workgroup copies each round to the
// image with consecutive lanes on consecutive dwords
__global__ void __launch_bounds__(STL_B) k_stl_emit(StlVol V, const uint8_t* __restrict__ ins,
                                                    const uint64_t* __restrict__ blk, uint32_t* __restrict__ faces)
{
    __shared__ uint32_t stage[STL_SB * 25];
    const int64_t v0 = int64_t(blockIdx.x) * STL_BV + int64_t(threadIdx.x) * STL_VT;
    unsigned cnt;
    const uint64_t masks = stl_masks(ins, V, v0, &cnt);
    unsigned F;
    const unsigned f0 = block_excl_scan(cnt, &F);
    if (F == 0u) return;   // uniform over the workgroup
    const uint64_t gbase = blk[blockIdx.x];
    int y0 = 0, x0 = 0, z0 = 0;
    if (cnt) {
        const int v = int(v0), pl = V.ny * V.nx;
        y0 = v % V.ny; x0 = (v / V.ny) % V.nx; z0 = v / pl;
    }
    for (unsigned r0 = 0; r0 < F; r0 += STL_SB) {
        if (cnt && f0 < r0 + STL_SB && f0 + cnt > r0) {
            unsigned f = f0;
            int y = y0, x = x0, z = z0;
            for (int i = 0; i < STL_VT; ++i) {
                const unsigned m = unsigned(masks >> (6 * i)) & 63u;
                for (int kind = 0; kind < 6; ++kind) {
                    if (!((m >> kind) & 1u)) continue;
                    if (f >= r0 && f < r0 + STL_SB) stl_put_face(stage + 25 * (f - r0), kind, y, x, z);
                    ++f;
                }
                if (++y == V.ny) {
                    y = 0;
                    if (++x == V.nx) {
                        x = 0;
                        ++z;
                    }
                }
            }
        }
        __syncthreads();
        const unsigned nf = (F - r0 < unsigned(STL_SB)) ? F - r0 : unsigned(STL_SB);
        uint32_t* dst = faces + 25 * (gbase + r0);
        for (unsigned d = threadIdx.x; d < 25u * nf; d += STL_B) dst[d] = stage[d];
        __syncthreads();
    }
}

namespace {
// inside bytes, block counts and their scan for xPhys at iso; *nfaces = faces of the surface
int stl_prepare(t3d_problem* p, const float* xPhys, float iso, StlVol& V, int64_t& nb, uint64_t* nfaces)
{
    T3D_CHECK(p && xPhys && nfaces, T3D_ERR_ARG, "t3d_stl: null argument");
    hipStream_t st = p->stream;
    V.ny = p->g.resY; V.nx = p->g.resX; V.nz = p->g.resZ;
    V.n = p->g.ncell;
    nb = (V.n + STL_BV - 1) / STL_BV;
    const int64_t padded = nb * STL_BV;   // whole blocks: the 8-byte loads of the last one stay inside
    if (!p->stlIns) T3D_HIP(hipMalloc(&p->stlIns, size_t(padded)));
    if (!p->stlOff) T3D_HIP(hipMalloc(&p->stlOff, sizeof(uint64_t) * size_t(nb + 1)));
    const uint8_t outsideVal = (0.0f >= iso) ? 1 : 0;   // cells without an element hold density 0
    hipLaunchKernelGGL(k_stl_fill, dim3(4096), dim3(256), 0, st, padded, outsideVal, p->stlIns);
    if (p->ne > 0)
        hipLaunchKernelGGL(k_stl_mark, dim3(unsigned((p->ne + 255) / 256)), dim3(256), 0, st, p->ne,
                           (const int32_t*)p->compactToCell, p->g.resX, p->g.resY, xPhys, iso, p->stlIns);
    hipLaunchKernelGGL(k_stl_count, dim3(unsigned(nb)), dim3(STL_B), 0, st, V, (const uint8_t*)p->stlIns,
                       p->stlOff);
    hipLaunchKernelGGL(k_stl_scan, dim3(1), dim3(STL_SCAN), 0, st, nb, p->stlOff);
    T3D_HIP(hipGetLastError());
    uint64_t f = 0;
    T3D_HIP(hipMemcpyAsync(&f, p->stlOff + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    T3D_HIP(hipStreamSynchronize(st));
    T3D_CHECK(2 * f <= 0xffffffffull, T3D_ERR_ARG, "t3d_stl: more triangles than a binary STL can count");
    *nfaces = f;
    return T3D_OK;
}

void stl_header(uint8_t* h, uint32_t ntri)
{
    std::memset(h, 0, 84);
    const char tag[] = "Generated by TOP3D_XL";
    std::memcpy(h, tag, sizeof(tag) - 1);
    h[80] = uint8_t(ntri & 0xff);
    h[81] = uint8_t((ntri >> 8) & 0xff);
    h[82] = uint8_t((ntri >> 16) & 0xff);
    h[83] = uint8_t(ntri >> 24);
}

int stl_emit(t3d_problem* p, const StlVol& V, int64_t nb, uint64_t nfaces, uint8_t* image)
{
    hipStream_t st = p->stream;
    uint8_t h[84];
    stl_header(h, uint32_t(2 * nfaces));
    T3D_HIP(hipMemcpyAsync(image, h, 84, hipMemcpyHostToDevice, st));
    if (nfaces > 0)
        hipLaunchKernelGGL(k_stl_emit, dim3(unsigned(nb)), dim3(STL_B), 0, st, V, (const uint8_t*)p->stlIns,
                           (const uint64_t*)p->stlOff, reinterpret_cast<uint32_t*>(image + 84));
    T3D_HIP(hipGetLastError());
    T3D_HIP(hipStreamSynchronize(st));   // h is a host buffer of this scope
    return T3D_OK;
}
}  // namespace

}  // namespace t3d

using namespace t3d;

extern "C" {

int t3d_stl_size(t3d_problem* p, const float* xPhys, float iso, uint64_t* nbytes, uint64_t* ntri)
{
    StlVol V;
    int64_t nb = 0;
    uint64_t f = 0;
    int rc = stl_prepare(p, xPhys, iso, V, nb, &f);
    if (rc) return rc;
    if (nbytes) *nbytes = 84 + 100 * f;
    if (ntri) *ntri = 2 * f;
    return T3D_OK;
}

int t3d_stl_image(t3d_problem* p, const float* xPhys, float iso, void* image, uint64_t nbytes)
{
    T3D_CHECK(image, T3D_ERR_ARG, "t3d_stl_image: null image");
    T3D_CHECK((reinterpret_cast<uintptr_t>(image) & 3u) == 0, T3D_ERR_ARG,
              "t3d_stl_image: the image must be 4-byte aligned");
    StlVol V;
    int64_t nb = 0;
    uint64_t f = 0;
    int rc = stl_prepare(p, xPhys, iso, V, nb, &f);
    if (rc) return rc;
    T3D_CHECK(nbytes == 84 + 100 * f, T3D_ERR_ARG, "t3d_stl_image: nbytes differs from the size t3d_stl_size returns");
    return stl_emit(p, V, nb, f, static_cast<uint8_t*>(image));
}

int t3d_export_surface_stl(t3d_problem* p, const float* xPhys, const char* path, float iso)
{
    T3D_CHECK(path, T3D_ERR_ARG, "t3d_export_surface_stl: null path");
    StlVol V;
    int64_t nb = 0;
    uint64_t f = 0;
    int rc = stl_prepare(p, xPhys, iso, V, nb, &f);
    if (rc) return rc;
    const uint64_t bytes = 84 + 100 * f;
    uint8_t* img = nullptr;
    T3D_HIP(hipMalloc(&img, bytes));
    std::vector<uint8_t> host(bytes);
    rc = stl_emit(p, V, nb, f, img);
    if (!rc && hipMemcpy(host.data(), img, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("t3d_export_surface_stl: copy of the image failed");
        rc = T3D_ERR_HIP;
    }
    (void)hipFree(img);
    if (rc) return rc;
    FILE* fp = std::fopen(path, "wb");
    T3D_CHECK(fp, T3D_ERR_ARG, std::string("t3d_export_surface_stl: cannot open ") + path);
    const size_t wr = std::fwrite(host.data(), 1, bytes, fp);
    const int cl = std::fclose(fp);
    T3D_CHECK(wr == bytes && cl == 0, T3D_ERR_ARG, std::string("t3d_export_surface_stl: write failed: ") + path);
    return T3D_OK;
}

}  // extern "C"
