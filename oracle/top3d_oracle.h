/*
 * top3d_oracle.h -- CPU oracle for the TOP3D_XL finite-element inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (libt3d.so) never links or calls it.
 *
 * It restates, in deterministic C, the arithmetic of the reference C++ port
 * (cpp/src/*.cpp under the reference tree), operation by operation:
 *   - every floating-point operation is the reference's, in the reference's
 *     order, with no contraction (built with -ffp-contract=off, SSE2 only);
 *   - every reduction that the reference performs with an OpenMP
 *     reduction/atomic (nondeterministic order) is replaced by ONE canonical
 *     reduction tree (o_canon_*), which the HIP kernels implement identically;
 *   - powf(rho,3)/powf(rho,2) are replaced by correctly rounded cube/square.
 * Every function cites the reference file:line it follows.
 */
#ifndef TOP3D_ORACLE_H
#define TOP3D_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- canonical reduction (shared definition with the HIP kernels) ----
 * Items are grouped in tiles of O_CANON_TILE = 256 lanes x 4 items.  In tile t
 * lane l sums items t*1024 + k*256 + l, k = 0..3, sequentially from +0.0;
 * the 256 lane sums are then combined by the halving tree s[l] += s[l+w],
 * w = 128,64,...,1.  Tile partials are reduced again by the same rule until a
 * single value remains.  The order depends only on the item count. */
#define O_CANON_LANES 256
#define O_CANON_PER_LANE 4
#define O_CANON_TILE (O_CANON_LANES * O_CANON_PER_LANE)

double o_canon_sum(const double* items, size_t n);
double o_canon_dot(const double* a, const double* b, size_t n);

/* ---- model (fea.cpp:181-447, padding.hpp:42-53) ---- */
typedef struct o_model {
    int nelx, nely, nelz;           /* original (solid) element counts */
    int resX, resY, resZ, baseL;    /* padded resolution, padding level L */
    int nnx, nny, nnz;
    int ne, nN, nDOF, nF;
    int32_t* eleMapBack;    /* [ne] reference full index ny*nx*ez+ny*ex+(ny-1-ey), Morton order */
    int32_t* eNodMat;       /* [8*ne] */
    int8_t*  color;         /* [ne] */
    int32_t* bucket;        /* [ne] compact ids by colour, ascending within a colour */
    int32_t  bucketStart[9];
    double   Ke[576];
    double*  F;             /* [3*nN] AoS loads (node-major x,y,z) */
    uint8_t* isFree;        /* [3*nN] */
    int32_t* freeDofIndex;  /* [nF] ascending global DOF ids */
    int32_t* eDofFree;      /* [24*ne] free index or -1 */
    int32_t* cellToCompact; /* [resX*resY*resZ] structured cell (x fastest) -> compact id or -1 */
    int32_t* compactToCell; /* [ne] */
} o_model;

/* Build the cuboid model of CreateVoxelFEAmodel_Cuboid + ApplyBoundaryConditions. */
o_model* o_model_new(int nely, int nelx, int nelz);
/* general voxel domain: solid[(ez*nelx + ex)*nely + ey] != 0 (NULL: cuboid) */
o_model* o_model_new_voxels(int nely, int nelx, int nelz, const uint8_t* solid);
/* the voxel domain with fixations and loads as lists on the original box's
 * nodes (x, y, z), y from the box's bottom: fix[4i] = x y z mask, loadNode[3i],
 * loadF[3i] (matlab/TOP3D_XL.m:470-552, :1319-1360); NULL for a node outside */
o_model* o_model_new_bc(int nely, int nelx, int nelz, const uint8_t* solid, int nFix, const int32_t* fix, int nLoad,
                        const int32_t* loadNode, const double* loadF);
void o_model_free(o_model* m);
void o_voxel_ke(double nu, double cellSize, double Ke[576]);
void o_padding(int nelx, int nely, int nelz, int* rx, int* ry, int* rz, int* L);

/* restrict_to_free / scatter_from_free (solver.cpp:12-49) between the
 * compact free vector [nF] and the full AoS vector [3*nN] (zeros at fixed). */
void o_free_to_full(const o_model* m, const double* vfree, double* vfull);
void o_full_to_free(const o_model* m, const double* vfull, double* vfree);

/* K*u on free DOFs, reference layout (fea.cpp:692-837). */
void o_kxu_free(const o_model* m, const double* E, const double* uFree, double* yFree);
/* same operator on the full AoS layout; y = 0 at fixed DOFs */
void o_kxu_full(const o_model* m, const double* E, const double* u, double* y);
/* single precision operator used as the mixed-precision reference:
 * same colour order, fp32 arithmetic, y = 0 at fixed DOFs */
void o_kxu_full_f32(const o_model* m, const float* E, const float* u, float* y);

/* ComputeCompliance (fea.cpp:587-677, 851-864): U full AoS; returns canonical sum of E*ce */
double o_compliance(const o_model* m, const double* E, const double* U, double* ce);

/* ---- multigrid (multigrid/*.cpp) ---- */
#define O_MAXLEV 8
typedef struct o_mg {
    int nlev;
    int resX[O_MAXLEV], resY[O_MAXLEV], resZ[O_MAXLEV], span[O_MAXLEV], nN[O_MAXLEV];
    double* W[O_MAXLEV];         /* trilinear weights table per coarse level (hierarchy.cpp:26-42) */
    uint8_t* mask[O_MAXLEV];     /* [3*nN_l] 1 = fixed */
    double* diag[O_MAXLEV];      /* [3*nN_l] */
    double* Lc; int Nc;          /* Cholesky factor, Nc = 0 -> diagonal coarsest */
    double weight;               /* (double)cfg.weight (float 0.6f) */
    double Emin;                 /* (double)youngsModulusMin */
    int NlimitDofs;
    int cholOk;                  /* 1 if factorisation succeeded (or not attempted) */
    double* Kc;                  /* assembled coarsest matrix before factorisation (kept for tests) */
    /* workspaces */
    double* r[O_MAXLEV];
    double* x[O_MAXLEV];
    double* tmp;
} o_mg;

o_mg* o_mg_new(const o_model* m, int nonDyadic, int maxLevels, float weight, float Emin, int NlimitDofs);
void o_mg_free(o_mg* g);
void o_mg_setup(const o_model* m, o_mg* g, const double* E);
/* V-cycle on the full AoS layout: z = M r (preconditioner.cpp:88-265) */
void o_mg_apply_full(const o_model* m, o_mg* g, const double* r, double* z);
/* transfers, exposed for tests */
void o_restrict_vec3(const o_mg* g, int lc, const double* rf, double* rc);
void o_prolong_vec3(const o_mg* g, int lc, const double* xc, double* xf);
void o_restrict_scalar_div(const o_mg* g, int lc, const double* rf, double* rc);
int  o_chol_inplace(double* A, int N);
void o_chol_solve(const double* L, const double* b, double* x, int N);

/* ---- PCG (solver.cpp:87-217) on the full layout, canonical dots ---- */
typedef void (*o_precond_fn)(void* ctx, const double* r, double* z);
int o_pcg_full(const o_model* m, const double* E, const double* b, double* x,
               double tol, int maxIt, o_precond_fn M, void* Mctx);
/* convenience: MG-preconditioned PCG */
int o_pcg_mg(const o_model* m, o_mg* g, const double* E, const double* b, double* x,
             double tol, int maxIt);

/* ---- PDE filter (filter.cpp:9-275) ---- */
typedef struct o_filter {
    double kernel[64];
    double* diagPrecond;   /* [nN] */
    double* lastX;         /* [nN] */
    double* lastRhs;       /* [nN] */
    int lastIters;         /* CG iterations of the last apply (-1 = early exit) */
    int lastWarm;
} o_filter;
o_filter* o_filter_new(const o_model* m, float rMin);
void o_filter_free(o_filter* f);
void o_filter_apply(const o_model* m, o_filter* f, const float* src, float* dst);
void o_pde_matvec(const o_model* m, const o_filter* f, const double* x, double* y);

/* ---- SIMP / sensitivities / OC (driver.cpp:117-189) ---- */
float o_cr_cube(float x);
float o_cr_square(float x);
float o_seq_sum(const float* v, int n);          /* std::accumulate<float> */
float o_seq_sum_x1mx(const float* v, int n);     /* accumulate of v*(1-v) */
void  o_simp(int ne, const float* xPhys, float E0, float Emin, double* E);
void  o_sensitivity(int ne, const float* xPhys, const double* ce, double Csolid,
                    float E0, float Emin, float penal, float* dc);
/* OC bisection; returns number of bisection steps; writes xnew, change */
int   o_oc_update(int ne, const float* x, const float* dc, float V0, float move,
                  float* xnew, float* change);
/* the same with passive elements (passive[e] != 0) held at 1 (matlab/TOP3D_XL.m:154) */
int   o_oc_update_passive(int ne, const float* x, const float* dc, const uint8_t* passive, float V0, float move,
                          float* xnew, float* change);

/* ---- full optimisation loop (driver.cpp:23-218) ---- */
typedef struct o_params {
    float E0, Emin, nu, penal, cgTol;
    int cgMaxIt;
    float V0; int nLoop; float rMin;
    int mgMaxLevels; float mgWeight; int nonDyadic; int NlimitDofs;
    int rebuildEvery;
} o_params;
void o_params_default(o_params* p);

typedef struct o_loop_rec {
    int loop, pcgIters, filterIters, bisectSteps;
    double C, Cdisp;
    float volFrac, sharpness, change, fval;
    uint64_t xCheck;   /* o_x_check of the design after the loop */
} o_loop_rec;

/* Position-keyed checksum of a float design: sum over i of
 * splitmix64(((uint64)i << 32) | bits(x[i])) modulo 2^64 (order-free, so a
 * parallel device reduction gives the same value).  Not in the reference: a
 * compact per-loop fingerprint of the density for the parity traces. */
uint64_t o_x_check(const float* x, int64_t n);

/* Runs the solid solve and up to p->nLoop OC loops.  rec has room for nLoop
 * records; returns the number of loops run; x_out [ne] receives the final
 * design; Csolid/solidIters receive the solid solve results. */
int o_top3d_run(int nely, int nelx, int nelz, const o_params* p,
                o_loop_rec* rec, float* x_out, double* Csolid, int* solidIters);
/* the loop on a voxel domain (solid mask as o_model_new_voxels, NULL = cuboid)
 * with passive elements (voxel-order mask, NULL = none) */
int o_top3d_run_ex(int nely, int nelx, int nelz, const uint8_t* solid, const uint8_t* passive, const o_params* p,
                   o_loop_rec* rec, float* x_out, double* Csolid, int* solidIters);
/* the same with the boundary conditions of o_model_new_bc (-1 for a bad node) */
int o_top3d_run_bc(int nely, int nelx, int nelz, const uint8_t* solid, const uint8_t* passive, int nFix,
                   const int32_t* fix, int nLoad, const int32_t* loadNode, const double* loadF, const o_params* p,
                   o_loop_rec* rec, float* x_out, double* Csolid, int* solidIters);

/* ---- voxel-surface STL (io.cpp:80-87, geometry_out/voxel_surface.hpp:12-93,
 * geometry_out/export_stl.hpp:14-33) ----
 * vol[y + ny*x + ny*nx*z]; returns the image size 84 + 100 * faces; the image
 * is written only when out != NULL and cap >= that size. */
uint64_t o_stl_image(int ny, int nx, int nz, const float* vol, float iso, uint8_t* out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
