// model.hpp -- host-side voxel model for the TOP3D_XL inner loop.
//
// Integer/bit-exact restatement of the reference model builder:
//   CreateVoxelFEAmodel_Cuboid (cpp/src/fea.cpp:181-358),
//   ApplyBoundaryConditions    (cpp/src/fea.cpp:360-447),
//   ComputeVoxelKe             (cpp/src/fea.cpp:117-179),
//   compute_adjusted_dims      (cpp/include/multigrid/padding.hpp:42-53).
// The device context is built from this structure (or from the same arrays
// handed over by a caller that built its own reference Problem).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace t3d {

struct HostModel {
    int nelx = 0, nely = 0, nelz = 0;   // original element counts
    int resX = 0, resY = 0, resZ = 0;   // padded resolution
    int baseL = 0;                      // padding level
    int numElements = 0, numNodes = 0, numDOFs = 0;
    // compact element e (Morton order) -> reference full index ny*nx*ez + ny*ex + (ny-1-ey)
    std::vector<int32_t> eleMapBack;
    std::vector<int32_t> eNodMat;       // [8*ne]
    std::vector<int32_t> elemColor;     // [ne]
    std::vector<std::vector<int32_t>> colorBuckets;  // 8 buckets, ascending ids
    std::array<double, 576> Ke{};
    std::vector<double> Fx, Fy, Fz;     // [nN] loads (SoA, as DOFData)
    std::vector<uint8_t> isFreeDOF;     // [3*nN]
    std::vector<int32_t> freeDofIndex;  // ascending free global DOFs
    std::vector<int32_t> eDofFreeMat;   // [24*ne]
};

void padding_dims(int nelx, int nely, int nelz, uint64_t numSolid, int& rx, int& ry, int& rz, int& L);
std::array<double, 576> voxel_ke(double nu, double cellSize);
void build_cuboid(HostModel& m, int nely, int nelx, int nelz);
// general voxel domain: solid[(ez * nelx + ex) * nely + ey] != 0 marks a solid
// voxel of the nely x nelx x nelz region (ey as in the reference's element
// numbering, stored at full index ny*nx*ez + ny*ex + (ny-1-ey)); void voxels
// carry no element, as the padding (fea.cpp:203-227), and the padding level
// follows the solid voxel count (padding.hpp:42-53, TOP3D_XL.m:1763-1784).
// solid == nullptr: the full cuboid.
void build_voxels(HostModel& m, int nely, int nelx, int nelz, const uint8_t* solid);
void apply_bcs(HostModel& m);
// the fixed-node, free-DOF and element-to-free-DOF maps after the BCs are set
void finish_bcs(HostModel& m);

// fixations and loads as lists on the nodes (x, y, z) of the original
// nelx x nely x nelz box, y counted from the box's bottom (the element y of
// the solid mask): fix[4 i] = x y z mask (bit c set: DOF c fixed),
// loadNode[3 i] = x y z, loadF[3 i] = the node's force
struct BcLists {
    int nFix = 0;
    const int32_t* fix = nullptr;
    int nLoad = 0;
    const int32_t* loadNode = nullptr;
    const double* loadF = nullptr;
};
// 0 on success, -1 (err set) for a node outside the box
int apply_bc_lists(HostModel& m, const BcLists& bc, std::string& err);

// structured cell index used on the device: (cz*resY + ty)*resX + cx with
// ty = resY-1-ey (node-y of the cell's lower face)
inline int cell_of_full(int full, int resX, int resY)
{
    const int cpl = resX * resY;
    const int ez = full / cpl, rem = full % cpl;
    const int ex = rem / resY, ty = rem % resY;
    return (ez * resY + ty) * resX + ex;
}

}  // namespace t3d
