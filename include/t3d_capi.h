/*
 * t3d_capi.h -- C ABI of the MI355X (gfx950) TOP3D_XL finite-element inner loop.
 *
 * Plain C types only (pointers, sizes, status codes).  Device pointers are
 * HIP device addresses, e.g. from t3d_malloc().  Every entry point returns
 * T3D_OK (0) or a negative status; t3d_last_error() describes the last failure
 * on the calling thread.  Results are complete when a call returns (each call
 * synchronises the problem's stream unless it says otherwise).
 *
 * Vector layouts:
 *   free   -- the reference's compact free-DOF vector: entry i is global DOF
 *             freeDofIndex[i] (cpp/src/solver.cpp:12-49).
 *   full   -- the node-major AoS vector v[3*node + c] over every grid node,
 *             node = (iz*nny + iy)*nnx + ix; fixed DOFs hold 0.  The device
 *             kernels work on this layout; t3d_free_to_full/t3d_full_to_free
 *             convert.
 *   compact element arrays -- indexed by the reference's compact (Morton)
 *             element id (cpp/src/fea.cpp:289-317).
 *
 * Each declaration names the reference function it replaces (file:line in the
 * reference tree).
 */
#ifndef T3D_CAPI_H
#define T3D_CAPI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T3D_OK 0
#define T3D_ERR_HIP (-1)        /* a HIP runtime call failed */
#define T3D_ERR_ARG (-2)        /* bad argument / size */
#define T3D_ERR_STATE (-3)      /* object not set up for this call */
#define T3D_ERR_DIVERGED (-5)   /* t3d_pcg_mixed: the residual became non-finite or grew 1e6-fold */
#define T3D_ERR_NOT_SPD (-4)    /* the Galerkin form's coarsest matrix is not positive definite
                                   (t3d_mg_set_mode 1 has no fallback); the reference form falls
                                   back to the diagonal silently, as preconditioner.cpp:76-77,
                                   and t3d_mg_info reports cholOk = 0 */

typedef struct t3d_problem t3d_problem;
typedef struct t3d_mg t3d_mg;
typedef struct t3d_filter t3d_filter;
typedef struct t3d_mixed t3d_mixed;

/* ---- runtime / memory helpers (no reference counterpart: device plumbing) ---- */
const char* t3d_last_error(void);
int t3d_device_info(char* name, size_t name_len, int* num_cus, size_t* total_mem);
int t3d_malloc(void** ptr, size_t bytes);
int t3d_free(void* ptr);
int t3d_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int t3d_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int t3d_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes);
int t3d_memset(void* dst_dev, int value, size_t bytes);
int t3d_synchronize(void);

/* ---- problem (model + boundary conditions + device mirrors) ----
 * Replaces CreateVoxelFEAmodel_Cuboid (cpp/include/fea.hpp:38, cpp/src/fea.cpp:181-358)
 * followed by ApplyBoundaryConditions (cpp/include/fea.hpp:39, cpp/src/fea.cpp:360-447):
 * builds the padded voxel model on the host and uploads its device mirror. */
int t3d_problem_create_cuboid(int nely, int nelx, int nelz, t3d_problem** out);
/* ComputeVoxelKe (cpp/include/fea.hpp:37, cpp/src/fea.cpp:117-179): the 24x24
 * unit-modulus element matrix of a cubic voxel by 2x2x2 Gauss quadrature,
 * row-major into Ke[576] (host memory), bit-identical to the reference's. */
int t3d_voxel_ke(double nu, double cellSize, double* Ke);
/* General voxel domain (the voxel-model generalisation of
 * CreateVoxelFEAmodel_Cuboid, cpp/src/fea.cpp:181-358, and its padding rule,
 * cpp/include/multigrid/padding.hpp:42-53 with the solid voxel count, as
 * matlab/TOP3D_XL.m:1763-1784): solid[(ez*nelx + ex)*nely + ey] != 0 marks a
 * solid voxel of the nely x nelx x nelz region (ey as in the reference's
 * element numbering); void voxels carry no element, nodes touched by no
 * element are fixed, and the reference's boundary conditions and load
 * (fea.cpp:360-447) apply unchanged.  solid == NULL: the full cuboid.
 * T3D_ERR_ARG if no voxel is solid. */
int t3d_problem_create_voxels(int nely, int nelx, int nelz, const uint8_t* solid, t3d_problem** out);
/* The voxel domain with boundary conditions and load given as lists, as the
 * MATLAB original takes them from a TopVoxel file (matlab/TOP3D_XL.m:458-552,
 * applied as FEA_ApplyBoundaryCondition, :1319-1360; the C++ port has only the
 * cantilever of fea.cpp:360-447).  Nodes are (x, y, z) on the original
 * nelx x nely x nelz box, y counted from its bottom like the solid mask's ey:
 * fix[4 i] = x y z mask (bit c set: DOF c fixed), loadNode[3 i] = x y z,
 * loadF[3 i] = the node's force.  A node listed twice keeps its first entry,
 * a loaded node is not fixed, nodes touched by no element are fixed.
 * T3D_ERR_ARG for a node outside the box. */
int t3d_problem_create_bc(int nely, int nelx, int nelz, const uint8_t* solid, int nFix, const int32_t* fix,
                          int nLoad, const int32_t* loadNode, const double* loadF, t3d_problem** out);

/* A TopVoxel file (matlab/TOP3D_XL.m:458-586, version 1.0; inputs/README.md),
 * parsed.  The file's voxel and node indices (1-based, MATLAB order: y rows
 * from the top, then x, then z) become the masks and node lists above; the
 * forces and fixation flags are taken as written. */
typedef struct t3d_voxel_model {
    int nelx, nely, nelz;
    uint8_t* solid;        /* [nely*nelx*nelz] solid[(ez*nelx + ex)*nely + ey] != 0 */
    uint8_t* passive;      /* same indexing: passive elements of the file */
    float* density;        /* same indexing: the file's densities (NULL when it has none) */
    int nFix;
    int32_t* fix;          /* [4*nFix] x y z mask */
    int nLoad;
    int32_t* loadNode;     /* [3*nLoad] x y z */
    double* loadF;         /* [3*nLoad] */
    int extraLoadCases;    /* further load cases in the file; the loop solves the first */
} t3d_voxel_model;
/* T3D_ERR_ARG (with the reason in t3d_last_error) for a file that is not TopVoxel 1.0 */
int t3d_topvoxel_read(const char* path, t3d_voxel_model** out);
int t3d_voxel_model_destroy(t3d_voxel_model* m);
/* t3d_problem_create_bc with the model's domain and lists */
int t3d_problem_create_model(const t3d_voxel_model* m, t3d_problem** out);
int t3d_problem_destroy(t3d_problem* p);
/* sizes[0..9] = resX resY resZ numElements numNodes numDOFs numFree nnx nny nnz */
int t3d_problem_sizes(const t3d_problem* p, int64_t* sizes);
/* Copies the host model (integer maps of fea.cpp:181-447 and Ke of fea.cpp:117-179).
 * Any pointer may be NULL.  Sizes: eleMapBack[ne], eNodMat[8ne], elemColor[ne],
 * Ke[576], F[3nN] (AoS), isFree[3nN], freeDofIndex[nF], eDofFreeMat[24ne]. */
int t3d_problem_model(const t3d_problem* p, int32_t* eleMapBack, int32_t* eNodMat, int32_t* elemColor,
                      double* Ke, double* F, uint8_t* isFree, int32_t* freeDofIndex, int32_t* eDofFreeMat);

/* restrict_to_free / scatter_from_free (cpp/include/solver.hpp:8-9, cpp/src/solver.cpp:12-49)
 * between the compact free vector and the full device layout (device pointers). */
int t3d_free_to_full(t3d_problem* p, const double* free_dev, double* full_dev);
int t3d_full_to_free(t3d_problem* p, const double* full_dev, double* free_dev);
/* restrict_to_free(pb.F) of driver.cpp:75 in the full layout (device pointer) */
int t3d_load_vector(t3d_problem* p, double* b_full_dev);

/* Element moduli eleMod[e] (compact order, device pointer) used by every
 * operator until the next call; scattered to the structured cell grid. */
int t3d_set_modulus(t3d_problem* p, const double* E_dev);

/* K_times_u_finest (cpp/include/fea.hpp:42,45; cpp/src/fea.cpp:763-849):
 * y = K_ff u on the full layout in FP64, bit-identical to the reference's
 * operation order (colour-ordered accumulation, sequential 24-term rows, no FMA). */
int t3d_kxu_f64(t3d_problem* p, const double* u_full_dev, double* y_full_dev);
/* Legacy host-vector form of K_times_u_finest: E[ne], uFree[nF] -> yFree[nF] (host pointers). */
int t3d_kxu_free_host(t3d_problem* p, const double* E_host, const double* uFree_host, double* yFree_host);

/* parallel_dot (cpp/src/solver.cpp:77-84) with the canonical reduction tree:
 * result to *out_host. */
int t3d_dot_f64(t3d_problem* p, const double* a_dev, const double* b_dev, int64_t n, double* out_host);

/* ---- multigrid preconditioner ----
 * build_static_once (cpp/include/multigrid/multigrid.hpp:65, cpp/src/multigrid/preconditioner.cpp:19-34):
 * level hierarchy (hierarchy.cpp:84-138) and fixed-DOF masks (masks.cpp:11-37).
 * weight/Emin are the reference's float parameters (MGPrecondConfig.weight,
 * GlobalParams.youngsModulusMin); NlimitDofs <= 1024. */
int t3d_mg_create(t3d_problem* p, int nonDyadic, int maxLevels, float weight, float Emin, int NlimitDofs,
                  t3d_mg** out);
int t3d_mg_destroy(t3d_mg* g);
/* make_diagonal_preconditioner_from_static (cpp/include/multigrid/multigrid.hpp:66-68,
 * cpp/src/multigrid/preconditioner.cpp:38-79): per-level Jacobi diagonals on the
 * device (diagonal.cpp:11-160) and, when the coarsest level has <= NlimitDofs
 * DOFs, the Galerkin coarsest matrix (coarsest.cpp:138-323) assembled on the
 * host in canonical element order and Cholesky-factored (coarsest.cpp:13-90).
 * E_dev: element moduli, compact order; also becomes the problem's modulus. */
int t3d_mg_setup(t3d_mg* g, const double* E_dev);
/* The V-cycle closure M(r, z) (cpp/src/multigrid/preconditioner.cpp:88-265), full layout. */
int t3d_mg_apply(t3d_mg* g, const double* r_full_dev, double* z_full_dev);
/* inspection: dims[4*l..] = resX resY resZ span; Nc = coarsest Cholesky size (0 = diagonal) */
int t3d_mg_info(const t3d_mg* g, int* nlev, int* dims, int* Nc, int* cholOk);
int t3d_mg_mask(const t3d_mg* g, int level, uint8_t* out_host);
int t3d_mg_diag(const t3d_mg* g, int level, double* out_host);
int t3d_mg_coarsest(const t3d_mg* g, double* Kc_host, double* L_host);
/* Multigrid form, chosen before t3d_mg_setup.  0: the reference's V-cycle
 * above (default).  1: the Galerkin hierarchy of the MATLAB original
 * (matlab/TOP3D_XL.m:945-1090, Solving_AssembleFEAstencil) on the same levels
 * and transfers: level operators A_{l+1} = P^T A_l P from per-cell element
 * matrices (the finest A_0 = M K M + (I - M)), Jacobi diagonals of those
 * operators, restriction by P^T (TOP3D_XL.m:809-823) and the coarsest level
 * solved by Cholesky at any size, in the 'Adapted' V-cycle (TOP3D_XL.m:772-807).
 * Not bit-comparable with the reference (no counterpart in the C++ port);
 * t3d_pcg and t3d_pcg_mixed use it through the same object. */
int t3d_mg_set_mode(t3d_mg* g, int mode);
/* V-cycle of the Galerkin form (matlab/TOP3D_XL.m:772-807): 0 'Adapted' (default;
 * restriction of the residual itself, one Jacobi correction), 1 'Standard'
 * (TOP3D_XL.m:783-803: the residual recomputed with the level operator before
 * restriction and after prolongation, A_l applied from the level's element
 * matrices).  'Standard' runs in the FP64 path (t3d_pcg, t3d_mg_apply) and in
 * the mixed solver (FP32 levels; set it before t3d_mixed_setup); the batched
 * solver refuses it (T3D_ERR_STATE).  Needs mode 1. */
int t3d_mg_set_vcycle(t3d_mg* g, int type);

/* ---- PCG ----
 * PCG_free (cpp/include/solver.hpp:22-25, cpp/src/solver.cpp:87-217) on the full
 * layout with the problem's current modulus (t3d_set_modulus / t3d_mg_setup);
 * M = NULL runs unpreconditioned CG.  x is the initial guess and the result;
 * *iters = iterations (maxIt when not converged), as the reference returns. */
int t3d_pcg(t3d_problem* p, t3d_mg* M, const double* b_full_dev, double* x_full_dev, double tol, int maxIt,
            int* iters);
/* ||r_{k+1}||^2 of the last solve for k < n (n <= maxIt of that solve) */
int t3d_pcg_history(t3d_problem* p, int maxIt, int n, double* rn2_host);

/* ---- mixed precision (the chosen "precision" form; no single reference
 * counterpart: it replaces PCG_free + the V-cycle closure, cpp/src/solver.cpp:87-217,
 * cpp/src/multigrid/preconditioner.cpp:88-265, with FP32 K*u and V-cycle inside a
 * flexible FP64 outer loop; FP64 dots, alpha/beta, x, r, residual replacement
 * and coarsest Cholesky, the last applied as x = Li^T (Li b) with the FP64
 * inverse factor Li = L^-1 instead of the substitution of
 * cpp/src/multigrid/coarsest.cpp:92-122) ---- */
int t3d_mixed_create(t3d_problem* p, t3d_mg* M, t3d_mixed** out);
int t3d_mixed_destroy(t3d_mixed* s);
/* refresh the FP32 copies of the multigrid diagonals and the inverse coarsest
 * Cholesky factor after t3d_mg_setup */
int t3d_mixed_setup(t3d_mixed* s);
/* x = the mixed solver's coarsest-level solve of b (FP64 device arrays of the
 * coarsest size 3 nN; masked DOFs of x are 0); the counterpart of
 * chol_solve_lower (cpp/src/multigrid/coarsest.cpp:92-122) through Li.
 * T3D_ERR_STATE when the coarsest level has no Cholesky factor. */
int t3d_mixed_coarse_solve(t3d_mixed* s, const double* b_dev, double* x_dev);
/* z = M32(r): one application of the mixed solver's FP32 V-cycle (the
 * multigrid object's cycle: 'Adapted' or 'Standard') to an FP32 full-layout
 * residual (device arrays of 3 nN floats); the counterpart of the
 * preconditioner closure (cpp/src/multigrid/preconditioner.cpp:88-265) in the
 * mixed solver's arithmetic */
int t3d_mixed_vcycle(t3d_mixed* s, const float* r_dev, float* z_dev);
/* r = b - K x with the FP64 operator of the mixed solver's residual
 * replacement (Walsh form in FP64; fixed DOFs of K x are 0, so r = b there),
 * stored in FP32 into r_dev, and *rn2 = the FP64 r.r of the unrounded
 * residual (device arrays of 3 nN; b and x FP64).  The counterpart of the
 * reference's residual r = b - K x (cpp/src/solver.cpp:110-122) in the
 * mixed solver's arithmetic; overwrites the solver's residual, so it is
 * called between solves. */
int t3d_mixed_true_residual(t3d_mixed* s, const double* b_dev, const double* x_dev, float* r_dev, double* rn2);
/* b, x: FP64 full layout (device); trueEvery > 0: residual replacement period;
 * flexible != 0: Polak-Ribiere beta, else the reference's beta = rz_new/rz_old.
 * Convergence is reported only when the true residual ||b - K x|| (FP64
 * operator) is below tol * max(1e-30, ||b||).  T3D_ERR_DIVERGED (x holds the
 * last iterate, *iters the iterations run) when the residual norm becomes
 * non-finite or exceeds 1e6 times the initial one, checked every 32 iterations. */
int t3d_pcg_mixed(t3d_mixed* s, const double* b_full_dev, double* x_full_dev, double tol, int maxIt, int trueEvery,
                  int flexible, int* iters);
/* y = K u in FP32 on the full layout (K_times_u_finest, cpp/src/fea.cpp:763-837,
 * in single precision; fixed DOFs of y are 0).  form 0: Walsh form (45-nonzero
 * element matrix in the voxel's Walsh-Hadamard basis, FMA), form 1: node gather
 * with the full 24x24 element matrix (FMA), form 2: the 24x24 x 24xN voxel
 * contraction on MFMA as split-bf16 x3 products (2-piece operands) accumulated
 * in FP32, form 3: the same with 3-piece operands and 6 products (bf16 x6) */
int t3d_kxu_f32(t3d_mixed* s, int form, const float* u_full_dev, float* y_full_dev);
/* operator of t3d_pcg_mixed: 0 Walsh form (default), 2 MFMA bf16 x3, 3 MFMA bf16 x6.
 * The MFMA forms split each voxel's corner displacements less their mean
 * (a rigid translation, which Ke maps to zero), so the rounding of form 2's
 * 16-bit operands scales with how the displacement varies across a voxel,
 * not with the displacement itself; the mixed solve with form 2 converges at
 * 256x512x256 in the iterations of form 0. */
int t3d_mixed_set_operator(t3d_mixed* s, int form);
/* average milliseconds per launch over reps launches, HIP events on the
 * problem's stream: op 0..3 the FP32 operator of form 0..3; op 4 the solver's
 * fused kernel of the selected operator (p = z + beta*u, y = K p, p.y partials);
 * op 5 / op 6: the Walsh-form fused kernel / operator in the other workgroup
 * tile order (hardware order vs XCD-grouped); op 7: the fused kernel's memory
 * streams as a plain float4 streaming pass (reads u, z, the solver's dx and
 * the cell moduli, writes two node vectors and dx; 4 (18 nN + ncell) bytes),
 * the bandwidth the same access mix reaches without the stencil */
int t3d_mixed_time_op(t3d_mixed* s, int op, const float* u_full_dev, float* y_full_dev, int reps, float* ms);
/* HIP-event timing of the solver's fused operator kernel inside t3d_pcg_mixed
 * (no counterpart in the reference; for the benchmark's roofline): on != 0
 * creates the events and zeroes the totals; get returns the summed kernel
 * milliseconds and the number of launches of iterations that ran. */
int t3d_mixed_set_timing(t3d_mixed* s, int on);
int t3d_mixed_get_timing(const t3d_mixed* s, double* ms_total, long long* launches);

/* ---- batching (the chosen "batching" form; no reference counterpart: the
 * reference runs its simulation_count problems one after another,
 * cpp/src/driver.cpp:23-218).  Many problems of uneven sizes solved together
 * by the mixed-precision MG-PCG: each kernel of an iteration is launched once
 * for the whole batch (device block map -> problem, block), the iteration is
 * replayed from hipGraphs, the Walsh-form tile depth and the row kernels'
 * rows per block are autotuned on the batch.  All problems share one stream. */
typedef struct t3d_batch t3d_batch;
/* n cuboids (nely[i], nelx[i], nelz[i]); multigrid with the reference's
 * defaults (non-dyadic, 5 levels, weight 0.6, NlimitDofs 1000) and Emin */
int t3d_batch_create(int n, const int* nely, const int* nelx, const int* nelz, float Emin, t3d_batch** out);
/* the same with voxel domains: solid[i] is problem i's mask as in
 * t3d_problem_create_voxels (solid == NULL or solid[i] == NULL: full cuboid) */
int t3d_batch_create_voxels(int n, const int* nely, const int* nelx, const int* nelz, const uint8_t* const* solid,
                            float Emin, t3d_batch** out);
int t3d_batch_destroy(t3d_batch* b);
/* the batch's problem / multigrid object i (owned by the batch); a problem's
 * multigrid may be switched to the Galerkin form (t3d_mg_set_mode) before
 * t3d_batch_setup */
int t3d_batch_problem(t3d_batch* b, int i, t3d_problem** p);
int t3d_batch_mg(t3d_batch* b, int i, t3d_mg** g);
/* E_dev[i]: compact element moduli of problem i (device); multigrid set-up */
int t3d_batch_setup(t3d_batch* b, const double* const* E_dev);
/* tiles: kwNZ in {8,16,32} node planes per Walsh-form tile, rowsPerBlock in
 * {2,4,8}; (0, 0) = autotune on the next solve (the default) */
int t3d_batch_set_tiles(t3d_batch* b, int kwNZ, int rowsPerBlock);
int t3d_batch_info(const t3d_batch* b, int* kwNZ, int* rowsPerBlock);
/* kernel timing of the next t3d_batch_solve calls (no reference counterpart):
 * on != 0 zeroes the totals.  Solves with useGraph = 0 record HIP events
 * around every kernel of each lockstep iteration (one stream synchronisation
 * per iteration); solves with useGraph = 1 record events around every graph
 * replay.  t3d_batch_get_timing: k >= 0 the k-th kernel name's total
 * milliseconds and launches, k < 0 the graph replays ("graph"); returns the
 * number of kernel names, or a negative status. */
int t3d_batch_set_timing(t3d_batch* b, int on);
int t3d_batch_get_timing(const t3d_batch* b, int k, char* name, int nameLen, double* ms_total, long long* launches);
/* b_dev[i], x_dev[i]: FP64 full layout of problem i (x = initial guess and
 * result); iters[i] as t3d_pcg_mixed; useGraph != 0 replays hipGraphs */
int t3d_batch_solve(t3d_batch* b, const double* const* b_dev, double* const* x_dev, double tol, int maxIt,
                    int trueEvery, int flexible, int useGraph, int* iters);

/* ---- compliance ----
 * ComputeCompliance (cpp/include/fea.hpp:47, cpp/src/fea.cpp:587-677,851-864):
 * ce[e] = ue^T Ke ue (0 when E <= 1e-16), returns sum E*ce (canonical tree).
 * E: compact element moduli; U: full layout displacements; ce: compact [ne]. */
int t3d_compliance(t3d_problem* p, const double* E_dev, const double* U_full_dev, double* ce_dev, double* C_host);
/* t3d_compliance without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results, bit for bit; the reduction's partial sums live in a static device
 * array of the library (about 1 MiB per device, held until the process ends)
 * instead of the problem's lazily allocated buffer, and the call makes no
 * hipMalloc/hipFree.  T3D_ERR_ARG above T3D_NOALLOC_MAX_ELEMENTS elements.
 * All calls on one device share that array: calls that overlap in time on
 * one device (other threads, other problems' streams) race. */
#define T3D_NOALLOC_MAX_ELEMENTS 67108864
int t3d_compliance_noalloc(t3d_problem* p, const double* E_dev, const double* U_full_dev, double* ce_dev,
                           double* C_host);

/* ---- PDE filter ----
 * SetupPDEFilter (cpp/include/filter.hpp:17, cpp/src/filter.cpp:9-75) */
int t3d_filter_create(t3d_problem* p, float rMin, t3d_filter** out);
int t3d_filter_destroy(t3d_filter* f);
/* kernel64: the 8x8 Helmholtz kernel; diag_host: [nN] Jacobi preconditioner */
int t3d_filter_info(const t3d_filter* f, double* kernel64, double* diag_host);
/* ApplyPDEFilter (cpp/include/filter.hpp:28, cpp/src/filter.cpp:175-275) on
 * float element arrays (compact order, device); keeps the warm-start state.
 * *iters = CG iterations, -1 for the early exit before the loop. */
int t3d_filter_apply(t3d_filter* f, const float* src_dev, float* dst_dev, int* iters);

/* ---- SIMP / OC (inline in TOP3D_XL_GLOBAL, cpp/src/driver.cpp:117-189) ---- */
/* E[e] = Emin + rho^3 (E0 - Emin) in float, rho = clamp(x,0,1) (driver.cpp:117-121) */
int t3d_simp(t3d_problem* p, const float* x_dev, float E0, float Emin, double* E_dev);
/* OC bisection + update (driver.cpp:169-189): dc = post-filter sensitivities;
 * out_host[0] = change, [1] = sharpness, [2] = volume fraction, [3] = bisection steps */
int t3d_oc_update(t3d_problem* p, const float* x_dev, const float* dc_dev, float V0, float move, float* xnew_dev,
                  float* out_host);
/* The same with passive elements (passive_dev[ne], compact order, nonzero =
 * passive) held at density 1 in every trial of the bisection and in xnew, as
 * matlab/TOP3D_XL.m:154,169 (the C++ port declares passive layers,
 * core.hpp:39-41, but never applies them).  passive_dev == NULL: t3d_oc_update. */
int t3d_oc_update_passive(t3d_problem* p, const float* x_dev, const float* dc_dev, const uint8_t* passive_dev,
                          float V0, float move, float* xnew_dev, float* out_host);
/* std::accumulate<float> of v[0..n) reproduced exactly (driver.cpp:181,188,201) */
int t3d_seq_sum_f32(t3d_problem* p, const float* v_dev, int64_t n, float* out_host);

/* ---- the optimisation loop ----
 * TOP3D_XL_GLOBAL (cpp/include/driver.hpp:4, cpp/src/driver.cpp:23-218) for one
 * simulation, parameters explicit (defaults = the reference's GlobalParams,
 * MGPrecondConfig and CLI values). */
typedef struct t3d_params {
    float E0, Emin, nu, penal, cgTol;
    int cgMaxIt;
    float V0;
    int nLoop;
    float rMin;
    int mgMaxLevels;
    float mgWeight;
    int nonDyadic;
    int NlimitDofs;
    int rebuildEvery;
    int precision;   /* 0 = FP64 plain path, 1 = mixed-precision solves (t3d_pcg_mixed) */
    int mgMode;      /* 0 = the C++ port's V-cycle, 1 = Galerkin multigrid (t3d_mg_set_mode;
                        no counterpart in the C++ port), reassembled every loop as in
                        matlab/TOP3D_XL.m:122 (rebuildEvery applies to mode 0 only) */
    const char* stlPath; /* when set, the final design's voxel surface at stlIso is written
                            there (t3d_export_surface_stl), as driver.cpp:272-275 does */
    float stlIso;    /* default 0.3 (driver.cpp:274) */
    const char* summaryPath; /* when set, the run's summary (t3d_summary_format) is written
                                there, as driver.cpp:264-270 writes it to out/comments/ */
    const uint8_t* solid;    /* optional voxel domain, solid[(ez*nelx + ex)*nely + ey] != 0
                                (t3d_problem_create_voxels); NULL = the cuboid */
    const uint8_t* passive;  /* optional passive voxels, same indexing: held at density 1 by
                                the OC update (t3d_oc_update_passive); NULL = none */
    const t3d_voxel_model* model; /* optional: the domain, boundary conditions and load of a
                                voxel model (t3d_topvoxel_read); its passive elements apply
                                unless passive is set; nely/nelx/nelz must be the model's */
} t3d_params;
typedef struct t3d_loop_record {
    int loop, pcgIters, filterIters, bisectSteps;
    double C, Cdisp;
    float volFrac, sharpness, change, fval;
    double tSolve, tOpti, tFilter, tTotal;   /* host wall-clock seconds */
    uint64_t xCheck;   /* checksum of the design after the loop: sum over elements i
                          (compact order) of splitmix64((i << 32) | bits(x[i])) mod 2^64
                          (no reference counterpart; a per-loop density fingerprint) */
} t3d_loop_record;
void t3d_params_default(t3d_params* p);
/* Runs the solid reference solve and up to P->nLoop OC loops; rec[recCap]
 * receives one record per loop, x_out_host[ne] (optional) the final design. */
int t3d_top3d_run(int nely, int nelx, int nelz, const t3d_params* P, t3d_loop_record* rec, int recCap,
                  float* x_out_host, double* Csolid, int* solidIters, int* loopsRun);

/* ---- run summary (driver.cpp:225-258): the text the reference prints after a
 * run and stores as markdown.  Values as the reference computes them: objFirst
 * and objLast are the first and last loop's Cdisp as float, the sums are the
 * per-loop solve, optimisation, filtering and total times in seconds, threads
 * the host threads of the run. */
typedef struct t3d_summary {
    int origX, origY, origZ;     /* nelx, nely, nelz */
    int resX, resY, resZ;        /* padded dims */
    float V0;
    int iterations;              /* loops run */
    int cgMaxIt;
    double Csolid;
    float objFirst, objLast;
    float sumCG, sumOpt, sumFilter, sumIter;
    int threads;
} t3d_summary;
/* the text into buf[cap] (NUL-terminated), *len = its length; T3D_ERR_ARG if it does not fit */
int t3d_summary_format(const t3d_summary* s, char* buf, size_t cap, size_t* len);

/* ---- voxel-surface STL of a design (io.cpp:80-87 export_surface_stl with
 * geometry_out/voxel_surface.hpp:12-93 extract_faces and
 * geometry_out/export_stl.hpp:14-33 write_stl_binary), built on the device.
 * xPhys_dev[ne] is the density in compact element order; a voxel of the padded
 * resY x resX x resZ volume is inside when its density is >= iso (cells without
 * an element hold 0), and each inside voxel's face towards an outside voxel or
 * the volume's boundary becomes two triangles, in the reference's order.  The
 * image is the reference's file byte for byte: 80-byte header, uint32 triangle
 * count, 50 bytes per triangle.  T3D_ERR_ARG beyond 2^32 - 1 triangles. */
/* image size: *nbytes = 84 + 50 * (*ntri) (either pointer may be null) */
int t3d_stl_size(t3d_problem* p, const float* xPhys_dev, float iso, uint64_t* nbytes, uint64_t* ntri);
/* the image into image_dev (4-byte aligned device memory of nbytes, which must
 * equal the size t3d_stl_size returns for the same arguments) */
int t3d_stl_image(t3d_problem* p, const float* xPhys_dev, float iso, void* image_dev, uint64_t nbytes);
/* export_surface_stl (io.cpp:80-87): the image built on the device, copied to
 * the host and written to path */
int t3d_export_surface_stl(t3d_problem* p, const float* xPhys_dev, const char* path, float iso);

/* ---- diagnostics (no reference counterpart): device floating-point semantics
 * the plain path relies on.  op 0: out = sqrtf(a) (float), 1: out = a / b
 * (float), 2: out = sqrt(a) (double); all must be correctly rounded. */
int t3d_fp_probe(int op, int64_t n, const void* a_dev, const void* b_dev, void* out_dev);

/* ---- timing (no reference counterpart): average device time of `reps`
 * back-to-back launches of an operator, measured with HIP events on the
 * problem's stream.  op: 0 = t3d_kxu_f64 (in = u, out = y, full layout);
 * 1 = streaming copy of 3*numNodes doubles (bandwidth / counter calibration). */
int t3d_time_op(t3d_problem* p, int op, const void* in_dev, void* out_dev, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
