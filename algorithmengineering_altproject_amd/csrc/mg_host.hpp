// mg_host.hpp -- host side of the geometric multigrid preconditioner.
//
// Restates (bit-exact integer/dyadic work, fixed-order FP64 for the coarsest
// matrix):
//   ComputeAdaptiveMaxLevels / BuildMGHierarchy (cpp/src/multigrid/hierarchy.cpp:84-138)
//   make_trilinear_weights_table               (cpp/src/multigrid/hierarchy.cpp:26-42)
//   MG_BuildFixedMasks + MG_Restrict_nodes     (cpp/src/multigrid/masks.cpp:11-37,
//                                               cpp/src/multigrid/transfers.cpp:95-191)
//   EleMod_CompactToFull_Finest + MG_AssembleCoarsestDenseK_Galerkin
//                                              (cpp/src/multigrid/coarsest.cpp:126-323)
//   chol_spd_inplace                           (cpp/src/multigrid/coarsest.cpp:13-90)
// The coarsest matrix is assembled in the canonical element order (cez, cex,
// cey) -- the reference uses OpenMP atomics there, whose order is unspecified.
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"

namespace t3d {

struct MGLevelDims {
    int resX = 0, resY = 0, resZ = 0;
    int span = 1;        // coarsening span from the finer level (1 for the finest)
    int64_t nN = 0;      // nodes
};

struct HostMG {
    std::vector<MGLevelDims> lv;                  // lv[0] = finest
    std::vector<std::vector<uint8_t>> masks;      // per level, per DOF (1 = fixed)
    std::vector<std::vector<double>> W;           // per level (coarse), (span+1)^3 x 8 weights
    double weight = 0.6;                          // (double)cfg.weight
    double Emin = 1e-6;                           // (double)youngsModulusMin
    int NlimitDofs = 1000;
};

void mg_build_static(const HostModel& m, bool nonDyadic, int maxLevels, float weight, float Emin,
                     int NlimitDofs, HostMG& H);
// Returns N = 3 * coarsest nodes; Kc (N x N, lower triangle + BC rows) is filled
// in the canonical order.  Ecompact: element moduli in compact order.
int mg_assemble_coarsest(const HostModel& m, const HostMG& H, const double* Ecompact, std::vector<double>& Kc);
// Dense Cholesky in place (row form if N < 256, column form otherwise),
// sequential inner sums; returns false if a pivot is not positive.
bool chol_factor(std::vector<double>& A, int N);

}  // namespace t3d
