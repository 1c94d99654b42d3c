// model.cpp -- host voxel model builder (see model.hpp for the reference
// functions it restates).  All integer maps are bit-exact with the reference;
// Ke follows the reference's quadrature operation order exactly.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace t3d {

void padding_dims(int nelx, int nely, int nelz, uint64_t numSolid, int& rx, int& ry, int& rz, int& L)
{
    // padding.hpp:23-53: L = number of (n+4)/8 reductions of the solid voxel
    // count until n < 50000, at least 3
    uint64_t n = numSolid;
    int l = 0;
    while (n >= 50000u) { ++l; n = (n + 4u) / 8u; }
    L = std::max(3, l);
    const int64_t step = int64_t(1) << L;
    auto pad = [&](int v) { return int((int64_t(v) + step - 1) / step * step); };
    rx = pad(nelx); ry = pad(nely); rz = pad(nelz);
}

std::array<double, 576> voxel_ke(double nu, double h)
{
    // fea.cpp:117-179 -- 2x2x2 Gauss quadrature of B^T D B, Voigt [xx,yy,zz,yz,xz,xy]
    const double lam = 1.0 * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    double D[6][6] = {};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) D[i][j] = lam;
        D[i][i] = lam + 2 * mu;
        D[3 + i][3 + i] = mu;
    }
    std::array<double, 576> K{};
    const double q = 1.0 / std::sqrt(3.0);
    const double pts[2] = {-q, q};
    const double detJ = (h * h * h) / 8.0;
    const double sc = 2.0 / h;
    // corner signs of the 8 local nodes in (s,t,p)
    static const int SS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const int ST[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const int SP[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const double s = pts[a], t = pts[b], p = pts[c];
                double gx[8], gy[8], gz[8];
                for (int i = 0; i < 8; ++i) {
                    // same factor grouping as the reference: (+-0.125*(1+-u))*(1+-v)
                    const double ft = ST[i] < 0 ? (1 - t) : (1 + t);
                    const double fp = SP[i] < 0 ? (1 - p) : (1 + p);
                    const double fs = SS[i] < 0 ? (1 - s) : (1 + s);
                    gx[i] = ((SS[i] < 0 ? -0.125 : 0.125) * ft * fp) * sc;
                    gy[i] = ((ST[i] < 0 ? -0.125 : 0.125) * fs * fp) * sc;
                    gz[i] = ((SP[i] < 0 ? -0.125 : 0.125) * fs * ft) * sc;
                }
                double B[6][24] = {};
                for (int i = 0; i < 8; ++i) {
                    const int k = 3 * i;
                    B[0][k] = gx[i]; B[1][k + 1] = gy[i]; B[2][k + 2] = gz[i];
                    B[3][k + 1] = gz[i]; B[3][k + 2] = gy[i];
                    B[4][k] = gz[i]; B[4][k + 2] = gx[i];
                    B[5][k] = gy[i]; B[5][k + 1] = gx[i];
                }
                double DB[6][24];
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 24; ++j) {
                        double acc = 0.0;
                        for (int k = 0; k < 6; ++k) acc += D[i][k] * B[k][j];
                        DB[i][j] = acc;
                    }
                const double wgt = 1.0 * 1.0 * 1.0 * detJ;
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j) {
                        double acc = 0.0;
                        for (int k = 0; k < 6; ++k) acc += B[k][i] * DB[k][j];
                        K[i * 24 + j] += acc * wgt;
                    }
            }
    return K;
}

static inline uint64_t interleave3(uint32_t v)
{
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= uint64_t((v >> i) & 1u) << (3 * i);
    return r;
}

void build_cuboid(HostModel& m, int nely, int nelx, int nelz)
{
    build_voxels(m, nely, nelx, nelz, nullptr);
}

void build_voxels(HostModel& m, int nely, int nelx, int nelz, const uint8_t* solid)
{
    auto isSolid = [&](int ex, int ey, int ez) {
        return !solid || solid[(size_t(ez) * nelx + ex) * nely + ey] != 0;
    };
    uint64_t numSolid = 0;
    for (int ez = 0; ez < nelz; ++ez)
        for (int ex = 0; ex < nelx; ++ex)
            for (int ey = 0; ey < nely; ++ey) numSolid += isSolid(ex, ey, ez) ? 1u : 0u;
    m.nelx = nelx; m.nely = nely; m.nelz = nelz;
    padding_dims(nelx, nely, nelz, numSolid, m.resX, m.resY, m.resZ, m.baseL);
    const int nx = m.resX, ny = m.resY, nz = m.resZ;
    const int nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
    m.numNodes = nnx * nny * nnz;
    m.numDOFs = 3 * m.numNodes;
    const int ne = int(numSolid);
    m.numElements = ne;
    // Morton-sorted solid voxels (fea.cpp:289-317); key = x<<2 | y<<1 | z bits
    struct KE { uint64_t key; int32_t full; };
    std::vector<KE> ord;
    ord.reserve(ne);
    for (int ez = 0; ez < nelz; ++ez)
        for (int ex = 0; ex < nelx; ++ex)
            for (int ey = 0; ey < nely; ++ey) {
                if (!isSolid(ex, ey, ez)) continue;   // void voxel: no element (fea.cpp:203-227)
                const uint64_t key = (interleave3(ex) << 2) | (interleave3(ey) << 1) | interleave3(ez);
                ord.push_back({key, ny * nx * ez + ny * ex + (ny - 1 - ey)});
            }
    std::sort(ord.begin(), ord.end(), [](const KE& a, const KE& b) { return a.key < b.key; });
    m.eleMapBack.resize(ne);
    m.eNodMat.resize(size_t(8) * ne);
    m.elemColor.resize(ne);
    m.colorBuckets.assign(8, {});
    // local node a of cell (x, ty, z): node offsets (dx, dy, dz)
    static const int DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int DY[8] = {1, 1, 0, 0, 1, 1, 0, 0};
    static const int DZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int e = 0; e < ne; ++e) {
        const int full = ord[e].full;
        m.eleMapBack[e] = full;
        const int ez = full / (nx * ny), rem = full % (nx * ny);
        const int ex = rem / ny, ty = rem % ny, ey = ny - 1 - ty;
        m.elemColor[e] = (ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2);
        m.colorBuckets[m.elemColor[e]].push_back(e);
        for (int a = 0; a < 8; ++a)
            m.eNodMat[size_t(8) * e + a] = nnx * nny * (ez + DZ[a]) + nnx * (ty + DY[a]) + (ex + DX[a]);
    }
    m.Ke = voxel_ke(0.3, 1.0);
    m.Fx.assign(m.numNodes, 0.0);
    m.Fy.assign(m.numNodes, 0.0);
    m.Fz.assign(m.numNodes, 0.0);
    m.isFreeDOF.assign(m.numDOFs, 1);
    m.freeDofIndex.clear();
}

void apply_bcs(HostModel& m)
{
    const int nx = m.resX, nz = m.resZ;
    const int nnx = nx + 1, nny = m.resY + 1, nnz = nz + 1;
    std::fill(m.isFreeDOF.begin(), m.isFreeDOF.end(), 1);
    std::fill(m.Fx.begin(), m.Fx.end(), 0.0);
    std::fill(m.Fy.begin(), m.Fy.end(), 0.0);
    std::fill(m.Fz.begin(), m.Fz.end(), 0.0);
    // clamp the x = 0 face (fea.cpp:374-385)
    for (int iz = 0; iz < nnz; ++iz)
        for (int iy = 0; iy < nny; ++iy) {
            const int n = nnx * nny * iz + nnx * iy;
            m.isFreeDOF[3 * n] = m.isFreeDOF[3 * n + 1] = m.isFreeDOF[3 * n + 2] = 0;
        }
    // unit -z load spread over the lower slice of the x = nelx face (fea.cpp:387-403)
    {
        const int ixLoad = std::min(m.nelx, nnx - 1);
        const int izTop = std::max(1, nz / 6);
        const int count = (izTop + 1) * nny;
        const double fz = -1.0 / double(count);
        for (int iz = 0; iz <= izTop; ++iz)
            for (int iy = 0; iy < nny; ++iy) m.Fz[nnx * nny * iz + nnx * iy + ixLoad] += fz;
    }
    finish_bcs(m);
}

// nodes touched by no element are fixed (fea.cpp:405-417); the free-DOF index
// and the element-to-free-DOF map (fea.cpp:419-447)
void finish_bcs(HostModel& m)
{
    {
        std::vector<uint8_t> used(m.numNodes, 0);
        for (int32_t n : m.eNodMat) used[n] = 1;
        for (int n = 0; n < m.numNodes; ++n)
            if (!used[n]) m.isFreeDOF[3 * n] = m.isFreeDOF[3 * n + 1] = m.isFreeDOF[3 * n + 2] = 0;
    }
    m.freeDofIndex.clear();
    std::vector<int32_t> g2f(m.numDOFs, -1);
    for (int i = 0; i < m.numDOFs; ++i)
        if (m.isFreeDOF[i]) { g2f[i] = int32_t(m.freeDofIndex.size()); m.freeDofIndex.push_back(i); }
    m.eDofFreeMat.resize(size_t(24) * m.numElements);
    for (int e = 0; e < m.numElements; ++e)
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c)
                m.eDofFreeMat[size_t(24) * e + 3 * a + c] = g2f[3 * m.eNodMat[size_t(8) * e + a] + c];
}

// Fixations and loads given as lists on the nodes of the original box, as a
// TopVoxel file carries them (matlab/TOP3D_XL.m:470-552, applied as
// FEA_ApplyBoundaryCondition, :1319-1360): a node listed twice keeps its first
// entry (MATLAB's unique), a loaded node is not fixed (its setdiff), the load
// is the node's force, the fixed DOFs are those whose flag is set, and nodes
// touched by no element are fixed as in the cuboid (fea.cpp:405-417).
int apply_bc_lists(HostModel& m, const BcLists& bc, std::string& err)
{
    const int nnx = m.resX + 1, nny = m.resY + 1;
    auto node_of = [&](const int32_t* c, int& n) {
        const int x = c[0], y = c[1], z = c[2];
        if (x < 0 || x > m.nelx || y < 0 || y > m.nely || z < 0 || z > m.nelz) return false;
        n = nnx * nny * z + nnx * (m.resY - y) + x;   // node y counts up from the box's bottom
        return true;
    };
    std::fill(m.isFreeDOF.begin(), m.isFreeDOF.end(), 1);
    std::fill(m.Fx.begin(), m.Fx.end(), 0.0);
    std::fill(m.Fy.begin(), m.Fy.end(), 0.0);
    std::fill(m.Fz.begin(), m.Fz.end(), 0.0);
    std::vector<uint8_t> loaded(m.numNodes, 0);
    for (int i = 0; i < bc.nLoad; ++i) {
        int n = 0;
        if (!node_of(bc.loadNode + 3 * i, n)) {
            err = "load " + std::to_string(i) + ": node outside the box";
            return -1;
        }
        if (loaded[n]) continue;
        loaded[n] = 1;
        m.Fx[n] = bc.loadF[3 * i];
        m.Fy[n] = bc.loadF[3 * i + 1];
        m.Fz[n] = bc.loadF[3 * i + 2];
    }
    std::vector<uint8_t> seen(m.numNodes, 0);
    for (int i = 0; i < bc.nFix; ++i) {
        int n = 0;
        if (!node_of(bc.fix + 4 * i, n)) {
            err = "fixation " + std::to_string(i) + ": node outside the box";
            return -1;
        }
        if (seen[n] || loaded[n]) continue;
        seen[n] = 1;
        for (int c = 0; c < 3; ++c)
            if ((bc.fix[4 * i + 3] >> c) & 1) m.isFreeDOF[3 * size_t(n) + c] = 0;
    }
    finish_bcs(m);
    return 0;
}

}  // namespace t3d
