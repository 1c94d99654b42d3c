#!/usr/bin/env python3
"""Benchmark of the TOP3D_XL finite-element inner loop on one MI355X.

A step is one multigrid-preconditioned PCG solve (PCG_free, cpp/src/solver.cpp:87-217)
from x0 = 0 to ||b - K x|| < tol ||b|| on the 192x384x192 cuboid (14.2M voxels,
42.9M free DOFs), with the reference's boundary conditions and load, element
moduli from SIMP of the uniform initial design rho = 0.12 (Emin = 1e-9), and the
multigrid preconditioner built for those moduli (outside the timed region).

`value` is solves/s of the mixed-precision solver (FP32 Walsh-form K*u and FP32
V-cycle inside a flexible FP64 outer loop, true-residual stopping) with the
Galerkin multigrid of the MATLAB original (matlab/TOP3D_XL.m:772-1090) on the
C++ port's levels, to tol = 1e-6, inputs resident in HBM.  The same run solves
the problem with that multigrid in FP64, with the FP64 plain path (the C++
port's V-cycle, bit-identical to the reference's arithmetic) and with the
mixed solver on that V-cycle at the reference's cgTol = 1e-3, and reports every
compliance against the converged compliance (FP64 to 1e-10).  `--mg reference`
times the mixed solver on the C++ port's V-cycle instead (iteration cap raised
from the reference's 800 to 4000: neither solver reaches 1e-3 within 800
iterations at this size).  `roofline` is the dominant kernel of the mixed solve
(the fused FP32 K*u), timed with HIP events around each of its launches inside
the timed solves, on its stream; `cpu_baseline` is the reference's own OpenMP
code timed on this host.  `--batch 64` runs the 64-problem uneven batch;
`--loop N` times N loops of the device-resident optimisation loop (mixed
Galerkin solves against the C++ port's algorithm on the device); `--stl`
times the device-built voxel-surface STL image of a lattice design on the
same grid against the reference's writer on one host core.  Each mode prints
its own single result line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "MG-PCG solves/s at 43M DOF, mixed vs FP64 GPU at equal compliance; K·u GB/s"
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nely", type=int, default=192)
    ap.add_argument("--nelx", type=int, default=384)
    ap.add_argument("--nelz", type=int, default=192)
    ap.add_argument("--mg", choices=["galerkin", "reference"], default="galerkin",
                    help="multigrid of the timed mixed solves: the Galerkin hierarchy of the MATLAB original "
                         "(default) or the C++ port's additive V-cycle")
    ap.add_argument("--vcycle", choices=["adapted", "standard"], default="adapted",
                    help="V-cycle of the Galerkin multigrid: the 'Adapted' cycle (default) or MATLAB's 'Standard' "
                         "cycle (residuals recomputed with the level operators; FP32 levels in the mixed solver)")
    ap.add_argument("--non-dyadic", type=int, choices=[0, 1], default=1,
                    help="hierarchy of the Galerkin multigrid: first coarsening by 4 (1, the reference driver's and "
                         "the MATLAB original's default) or by 2 at every level (0)")
    ap.add_argument("--levels", type=int, default=5, help="multigrid levels of the Galerkin hierarchy (maxLevels)")
    ap.add_argument("--no-dyadic-check", action="store_true",
                    help="skip the comparison solve on the hierarchy coarsened by 2 at every level (6 levels)")
    ap.add_argument("--tol", type=float, default=None,
                    help="tolerance of the timed solves (default 1e-6 with --mg galerkin, else --ref-tol)")
    ap.add_argument("--ref-tol", type=float, default=float(np.float32(1e-3)),
                    help="the reference's cgTol, used for the solves with the C++ port's V-cycle")
    ap.add_argument("--tight-tol", type=float, default=1e-5)
    ap.add_argument("--maxit", type=int, default=4000)
    ap.add_argument("--true-every", type=int, default=None,
                    help="residual replacement period of the mixed solver (default 25 with the Galerkin multigrid, "
                         "whose residual falls 1e3 in about 25 iterations, else 50)")
    ap.add_argument("--design", choices=["uniform", "random-filtered"], default="uniform",
                    help="uniform rho = 0.12, or rho ~ U[1e-3, 1] (seeded) PDE-filtered with rmin = 2")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--batch", type=int, default=0,
                    help="batched workload instead: this many uneven cantilevers (seeded sizes, V0, voids)")
    ap.add_argument("--operator", choices=["walsh", "mfma3", "mfma6"], default="walsh",
                    help="FP32 operator of the mixed solve: Walsh form (FMA), MFMA bf16 x3 or MFMA bf16 x6")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch mode: launch each iteration's kernels directly instead of replaying hipGraphs "
                         "(same kernels; for kernel-trace runs)")
    ap.add_argument("--no-seq", action="store_true", help="batch mode: skip the sequential comparison")
    ap.add_argument("--voids", choices=["removed", "density0"], default="removed",
                    help="batch mode: spherical voids as absent voxels (default) or as elements of density 0")
    ap.add_argument("--loop", type=int, default=0,
                    help="time N loops of the device-resident optimisation loop instead of the solves")
    ap.add_argument("--stl", action="store_true",
                    help="the voxel-surface STL image of a lattice design on the same grid instead of the solves")
    ap.add_argument("--kxu-reps", type=int, default=50)
    ap.add_argument("--tile-order-compare", action="store_true",
                    help="with --kxu-only: also time the Walsh-form kernels in the other workgroup tile order")
    ap.add_argument("--kxu-only", action="store_true", help="only the K*u timing blocks (for counter runs)")
    ap.add_argument("--copy-only", action="store_true", help="only the streaming-copy calibration kernel")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain solves")
    ap.add_argument("--no-tight", action="store_true", help="skip the tight-tolerance comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=20, help="reference PCG iterations timed on the CPU")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class stdout_to_stderr:
    """Sends what native code writes to file descriptor 1 to stderr for the
    duration (the reference's code logs to std::cout; stdout carries only the
    result line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, E, iters):
    """The reference's own OpenMP code (oracle/_ref: its sources built as its job
    builds them, -O3 -ffast-math -fopenmp) on this host's cores, threads bound
    close to cores as its job script sets them (OMP_PROC_BIND=close,
    OMP_PLACES=cores).  A bounded sample of the headline workload (192x384x192):
    the median of 20 K_times_u_finest applies and 20 MG-PCG iterations, scaled to
    solves/s with the iteration count of one FP64 solve (the reference runs the
    FP64 plain path's iterations); and one whole reference solve of the
    96x192x96 cantilever (uniform rho = 0.12, its tol 1e-3, its cap of 800)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    if not O.have_ref(fast=True):
        return None
    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ["OMP_PROC_BIND"] = "close"   # read by libgomp when the reference library loads
    os.environ["OMP_PLACES"] = "cores"
    R = O.ref(fast=True)
    import ctypes as C
    t0 = time.time()
    sizes = np.zeros(7, np.int32)
    R.ref_sizes(args.nely, args.nelx, args.nelz, sizes)   # builds the reference model + MG statics
    t_setup = time.time() - t0
    u = np.random.default_rng(1).uniform(-1, 1, int(sizes[5]))
    Ec = np.ascontiguousarray(E)
    kxu_ms = R.ref_time_kxu(args.nely, args.nelx, args.nelz, Ec, u, 20)
    nF, ne = int(sizes[5]), int(sizes[3])
    it = C.c_int()
    ms = R.ref_time_pcg(args.nely, args.nelx, args.nelz, Ec, float(np.float32(1e-9)), 0.0, args.cpu_iters,
                        C.byref(it))
    per_it = ms / 1e3 / it.value
    # one whole solve of the 96x192x96 cantilever with the reference's tolerance and cap
    d1 = (96, 192, 96)
    s1 = np.zeros(7, np.int32)
    R.ref_sizes(*d1, s1)
    rho3 = np.float32(0.12) * np.float32(0.12) * np.float32(0.12)
    emin = np.float32(1e-9)
    E1 = np.full(int(s1[3]), float(emin + rho3 * (np.float32(1.0) - emin)))
    it1 = C.c_int()
    ms1 = R.ref_time_pcg(*d1, E1, float(emin), float(np.float32(1e-3)), 800, C.byref(it1))
    return {
        "value": 1.0 / (per_it * iters),
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "cpu": cpu_model(),
        "omp": "OMP_NUM_THREADS=%d OMP_PROC_BIND=close OMP_PLACES=cores" % cores,
        "kxu_ms_median_of_20": kxu_ms,
        "kxu_gbps": 8.0 * (2 * nF + ne) / kxu_ms / 1e6,
        "pcg_s_per_iteration": per_it,
        "solve_96x192x96": {"iterations": it1.value, "seconds": ms1 / 1e3, "tol": 1e-3, "max_iters": 800},
        "sample": "reference PCG_free + MG V-cycle (oracle/_ref, OpenMP -O3 -ffast-math), %d iterations on the "
                  "same 192x384x192 problem = %.3f s/iteration, scaled by the %d iterations of one FP64 solve; "
                  "reference K_times_u_finest median of 20 applies %.1f ms; one whole reference solve of "
                  "96x192x96 (tol 1e-3, cap 800): %d iterations, %.1f s; model setup %.1f s not timed"
                  % (it.value, per_it, iters, kxu_ms, it1.value, ms1 / 1e3, t_setup),
    }


def mfma_flops(p, products):
    """bf16 MFMA flops the split-bf16 operator issues per apply: per 32 voxels of
    a tile row, 2 K steps x `products` products of a 32x32x16 MFMA (2*32*32*16 flop)."""
    nnx, nny, nnz = p.nnx, p.nny, p.nnz
    tiles = -(-nnx // 31) * -(-nny // 7) * -(-nnz // 16)
    steps_per_tile = 17                      # 16 node planes + the voxel layer below
    return tiles * steps_per_tile * 8 * 2 * products * 2 * 32 * 32 * 16


def true_residual(t3d, p, db, dx):
    dy = t3d.DeviceArray(p.nDOF)
    p.kxu_f64(dx, dy)
    b = db.download()
    return float(np.linalg.norm(b - dy.download()) / np.linalg.norm(b))


def batch_workload(rng, n):
    """n cantilevers: nelx ~ U{48, 56, ..., 160}, nely = nelz = nelx/2 rounded to a
    multiple of 8, V0 ~ U[0.08, 0.30]; every second one with 1-6 spherical voids
    of radius 3-15 voxels at seeded centres inside the domain."""
    probs = []
    for i in range(n):
        nelx = int(rng.choice(np.arange(48, 161, 8)))
        ny = max(8, int(round(nelx / 2 / 8)) * 8)
        v0 = float(rng.uniform(0.08, 0.30))
        sph = []
        if i % 2 == 1:
            for _ in range(int(rng.integers(1, 7))):
                r = float(rng.uniform(3, 15))
                sph.append((float(rng.uniform(0, nelx)), float(rng.uniform(0, ny)), float(rng.uniform(0, ny)), r))
        probs.append(((ny, nelx, ny), v0, sph))
    return probs


def batch_mask(dims, spheres):
    """solid[ez, ex, ey] of a cuboid with spherical voids (voxel centres inside a sphere are void)."""
    nely, nelx, nelz = dims
    if not spheres:
        return None
    ez, ex, ey = np.meshgrid(np.arange(nelz), np.arange(nelx), np.arange(nely), indexing="ij")
    solid = np.ones((nelz, nelx, nely), np.uint8)
    for cx, cy, cz, r in spheres:
        solid[(ex + 0.5 - cx) ** 2 + (ey + 0.5 - cy) ** 2 + (ez + 0.5 - cz) ** 2 <= r * r] = 0
    return solid


def batch_design(t3d, p, v0, spheres):
    x = np.full(p.ne, np.float32(v0), np.float32)
    if spheres:
        f = p.model()["eleMapBack"].astype(np.int64)
        ny, nx = p.resY, p.resX
        ez = f // (ny * nx)
        rem = f % (ny * nx)
        ex, ey = rem // ny, ny - 1 - rem % ny
        for cx, cy, cz, r in spheres:
            x[(ex + 0.5 - cx) ** 2 + (ey + 0.5 - cy) ** 2 + (ez + 0.5 - cz) ** 2 <= r * r] = 0.0   # passive void
    dx = t3d.DeviceArray.from_host(x)
    dE = t3d.DeviceArray(p.ne)
    t3d.simp(p, dx, 1.0, float(np.float32(1e-9)), dE)
    return dE


def main_stl(args, t3d):
    """export_surface_stl (io.cpp:80-87) of a gyroid lattice design (period 32
    voxels, density 0.5 + 0.35 g) at the driver's threshold 0.3 on the
    nely x nelx x nelz cuboid: device images per second with the density resident
    in HBM; the reference's extract_faces + write_stl_binary on the host as the
    CPU baseline (writing to /tmp)."""
    t0 = time.time()
    p = t3d.Problem(args.nely, args.nelx, args.nelz)
    emb = p.ele_map_back().astype(np.int64)
    ny, nx, nz = p.resY, p.resX, p.resZ
    ez, rem = emb // (ny * nx), emb % (ny * nx)
    ex, ty = rem // ny, rem % ny
    k = 2 * np.pi / 32
    g = np.sin(k * ex) * np.cos(k * ty) + np.sin(k * ty) * np.cos(k * ez) + np.sin(k * ez) * np.cos(k * ex)
    x = np.clip(0.5 + 0.35 * g, 0.0, 1.0).astype(np.float32)
    del g, ez, rem, ex, ty
    iso = float(np.float32(0.3))
    dx = t3d.DeviceArray.from_host(x)
    nbytes, ntri = t3d.stl_size(p, dx, iso)
    img = t3d.DeviceArray((nbytes + 3) // 4, np.uint32)
    log("STL workload %dx%dx%d: %d voxels, %d triangles, %.1f MB image (set-up %.1fs)"
        % (args.nely, args.nelx, args.nelz, nx * ny * nz, ntri, nbytes / 1e6, time.time() - t0))
    L = t3d.lib()
    for _ in range(args.warmup):
        t3d.check(L.t3d_stl_image(p.h, dx.ptr, iso, img.ptr, nbytes), "t3d_stl_image")
    t3d.check(L.t3d_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t3d.check(L.t3d_stl_image(p.h, dx.ptr, iso, img.ptr, nbytes), "t3d_stl_image")
    t3d.check(L.t3d_synchronize(), "sync")
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ncell = nx * ny * nz
    # algorithmic bytes of one image: density and element-to-cell map (8 per element),
    # the inside bytes written twice (fill, mark) and read twice (count, emit), the image
    alg = 8 * p.ne + 4 * ncell + nbytes
    log("STL image on the device: %.3f ms/image, %.0f GB/s algorithmic" % (ms, alg / ms / 1e6))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        if O.have_ref(fast=True):
            vol = np.zeros(ncell, np.float32)
            vol[emb] = x
            path = "/tmp/t3d_bench_ref.stl"
            t1 = time.perf_counter()
            with stdout_to_stderr():
                rc = O.ref(fast=True).ref_stl_volume(ny, nx, nz, vol, iso, path.encode())
            tc = time.perf_counter() - t1
            same = rc == 0 and os.path.getsize(path) == nbytes
            os.remove(path)
            cpu = {"value": 1.0 / tc, "unit": "images/s", "cores": 1, "kind": "reference",
                   "sample": "one image of the same volume: the reference's extract_faces + write_stl_binary "
                             "(voxel_surface.hpp, export_stl.hpp; -O3 -ffast-math build), file written to /tmp; "
                             "image size %s the device's" % ("equal to" if same else "DIFFERENT from")}
            log("reference on the host: %.2f s/image" % tc)
    line = {"metric": "voxel-surface STL images/s (export_surface_stl) at 14.2M voxels",
            "value": 1e3 / ms, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic (gyroid lattice density, period 32 voxels, threshold 0.3)",
            "config": {"workload": "binary STL image of the voxel surface, cuboid %dx%dx%d (nely x nelx x nelz), "
                                   "%d triangles" % (args.nely, args.nelx, args.nelz, ntri),
                       "image_bytes": nbytes, "triangles": ntri, "voxels": ncell},
            "roofline": {"bound": "hbm", "achieved": alg / ms / 1e6, "peak": 8000.0, "unit": "GB/s",
                         "frac": alg / ms / 1e6 / 8000.0, "traffic": None, "algorithmic_bytes": alg,
                         "timing": "wall clock around whole t3d_stl_image calls (both passes, the count's "
                                   "device-to-host copy and the synchronisations included)"},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def main_loop(args, t3d):
    """TOP3D_XL_GLOBAL (driver.cpp:23-218) on the device for args.loop loops of the
    nely x nelx x nelz cantilever (V0 0.12, rMin sqrt(3), Emin 1e-9): the mixed
    solves with the Galerkin multigrid (tol 1e-6, reassembled every loop), and the
    C++ port's algorithm (FP64 plain path, its V-cycle, tol 1e-3, 800-iteration
    cap, rebuilt every 5 loops) for comparison.  Per-loop wall time from the
    loop's own records (host clock around each loop, every device step inside)."""
    common = dict(V0=0.12, nLoop=args.loop, Emin=float(np.float32(1e-9)))
    runs = {}
    for name, kw in (("mixed_galerkin", dict(precision=1, mgMode=1, cgTol=1e-6, cgMaxIt=args.maxit)),
                     ("fp64_reference_algorithm", dict(precision=0, mgMode=0))):
        t0 = time.time()
        recs, x, Cs, si = t3d.top3d_run(args.nely, args.nelx, args.nelz, **dict(common, **kw))
        wall = time.time() - t0
        tot = sum(r["tTotal"] for r in recs)
        runs[name] = {"loops": len(recs), "s_per_loop": tot / len(recs), "wall_s_incl_setup_and_solid": wall,
                      "solid_iters": si, "pcg_iters": [r["pcgIters"] for r in recs],
                      "s_solve_per_loop": sum(r["tSolve"] for r in recs) / len(recs),
                      "s_filter_per_loop": sum(r["tFilter"] for r in recs) / len(recs),
                      "s_opti_per_loop": sum(r["tOpti"] for r in recs) / len(recs),
                      "objective": [r["Cdisp"] for r in recs]}
        log("%s: %d loops, %.3f s/loop (solve %.3f, filter %.3f, OC %.3f), PCG iterations %s, run %.1f s"
            % (name, len(recs), runs[name]["s_per_loop"], runs[name]["s_solve_per_loop"],
               runs[name]["s_filter_per_loop"], runs[name]["s_opti_per_loop"], runs[name]["pcg_iters"], wall))
    a, b = runs["mixed_galerkin"], runs["fp64_reference_algorithm"]
    line = {"metric": "optimisation loops/s (TOP3D_XL_GLOBAL), cuboid %dx%dx%d" % (args.nely, args.nelx, args.nelz),
            "value": 1.0 / a["s_per_loop"],
            "unit": "loops/s", "n_gpus": 1, "steps": a["loops"], "warmup": 0, "ms_per_step": 1e3 * a["s_per_loop"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic (the reference's cantilever, V0 = 0.12)",
            "config": {"workload": "device-resident optimisation loop, cuboid %dx%dx%d, mixed solves with the "
                                   "Galerkin multigrid to tol 1e-6" % (args.nely, args.nelx, args.nelz),
                       "runs": runs, "speedup_vs_reference_algorithm_on_gpu": b["s_per_loop"] / a["s_per_loop"]},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def main_batch(args, t3d):
    galerkin = args.mg == "galerkin"
    btol = args.tol if args.tol is not None else (1e-6 if galerkin else args.ref_tol)
    bte = args.true_every if args.true_every is not None else (25 if galerkin else 50)
    rng = np.random.default_rng(args.seed)
    probs = batch_workload(rng, args.batch)
    dims = [d for d, _, _ in probs]
    t0 = time.time()
    if args.voids == "removed":   # void voxels carry no element (voxel-domain model)
        solids = [batch_mask(d, s) for d, _, s in probs]
        B = t3d.Batch(dims, Emin=float(np.float32(1e-9)), solids=solids, galerkin=galerkin)
        dEs = [batch_design(t3d, p, v, ()) for p, (_, v, s) in zip(B.problems, probs)]
    else:                         # void voxels kept as elements of density 0 (E = Emin)
        solids = [None] * len(probs)
        B = t3d.Batch(dims, Emin=float(np.float32(1e-9)), galerkin=galerkin)
        dEs = [batch_design(t3d, p, v, s) for p, (_, v, s) in zip(B.problems, probs)]
    B.setup(dEs)
    dbs, xs = [], []
    for p in B.problems:
        db = t3d.DeviceArray(p.nDOF)
        p.load_vector(db)
        dbs.append(db)
        xs.append(t3d.DeviceArray(p.nDOF))
    ne = sum(p.ne for p in B.problems)
    nF = sum(p.nF for p in B.problems)
    log("batch of %d problems, %d elements, %d free DOFs (set-up %.1fs)" % (len(probs), ne, nF, time.time() - t0))
    for _ in range(max(1, args.warmup)):          # the first solve also autotunes the tiles
        for x in xs:
            x.zero()
        B.solve(dbs, xs, btol, args.maxit, trueEvery=bte, useGraph=not args.no_graph)
    tiles = B.info()
    t3d.lib().t3d_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for x in xs:
            x.zero()
        its = B.solve(dbs, xs, btol, args.maxit, trueEvery=bte, useGraph=not args.no_graph)
    t3d.lib().t3d_synchronize()
    ms_batch = (time.perf_counter() - t0) / args.steps * 1e3
    res = [true_residual(t3d, p, db, x) for p, db, x in zip(B.problems, dbs, xs)]
    log("batched: %.1f ms for all %d solves (%.1f solves/s), tiles %s, iterations %d..%d, max true residual %.2e"
        % (ms_batch, len(probs), len(probs) * 1e3 / ms_batch, tiles, min(its), max(its), max(res)))
    # kernel-level evidence outside the timed steps: HIP events around every graph replay of one solve,
    # then around every kernel of the same lockstep iterations launched directly (one sync per iteration)
    ktiming = None
    if not args.no_graph:
        B.set_timing(True)
        for x in xs:
            x.zero()
        B.solve(dbs, xs, btol, args.maxit, trueEvery=bte, useGraph=True)
        g_ms, g_n = B.timing()["graph"]
        B.set_timing(True)
        for x in xs:
            x.zero()
        B.solve(dbs, xs, btol, args.maxit, trueEvery=bte, useGraph=False)
        kt = B.timing()
        B.set_timing(False)
        per = {k: {"us_per_launch": 1e3 * v[0] / v[1], "launches": v[1]} for k, v in kt.items() if k != "graph" and v[1]}
        n_it = kt["kb_kxw32"][1] if "kb_kxw32" in kt else 0
        direct = sum(v[0] for k, v in kt.items() if k != "graph") / max(1, n_it)
        ktiming = {"graph_replay_ms_per_iteration": g_ms / max(1, g_n), "graph_replays": g_n,
                   "direct_launch_ms_per_iteration": direct, "direct_iterations": n_it,
                   "per_kernel_direct": per,
                   "timing": "HIP events on the batch stream: around each graph replay (one lockstep iteration of "
                             "every active problem); around each kernel of the directly launched iterations "
                             "(launch gap included)"}
        log("graph replay: %.3f ms per lockstep iteration (%d replays); direct launches: %.3f ms per iteration"
            % (ktiming["graph_replay_ms_per_iteration"], g_n, direct))
        for k, v in sorted(per.items(), key=lambda kv: -kv[1]["us_per_launch"] * kv[1]["launches"]):
            log("  %-18s %9.1f us per launch, %6d launches" % (k, v["us_per_launch"], v["launches"]))
    # the same problems one after another with the single-problem solver
    ms_seq, its_seq = None, []
    if not args.no_seq:
        Ms = []
        for i, p in enumerate(B.problems):   # the batch's problems, each with its own multigrid and solver state
            M = t3d.MG(p, Emin=float(np.float32(1e-9)), galerkin=galerkin)
            M.setup(dEs[i])
            mx = t3d.Mixed(p, M)
            mx.setup()
            Ms.append((M, mx))
        for (M, mx), db, x in zip(Ms, dbs, xs):   # warm-up
            x.zero()
            mx.pcg(db, x, btol, args.maxit, trueEvery=bte)
        t3d.lib().t3d_synchronize()
        t0 = time.perf_counter()
        its_seq = []
        for (M, mx), db, x in zip(Ms, dbs, xs):
            x.zero()
            its_seq.append(mx.pcg(db, x, btol, args.maxit, trueEvery=bte))
        t3d.lib().t3d_synchronize()
        ms_seq = (time.perf_counter() - t0) * 1e3
        log("sequential: %.1f ms for all solves (%.1f solves/s), speed-up %.2fx" % (ms_seq, len(probs) * 1e3 / ms_seq,
                                                                                 ms_seq / ms_batch))
    line = {
        "metric": METRIC,
        "value": len(probs) * 1e3 / ms_batch,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (%d cantilevers, seed %d: nelx~U{48..160}, nely=nelz=nelx/2, V0~U[0.08,0.30], half with "
                "1-6 spherical voids r~U[3,15], %s)" % (len(probs), args.seed,
                                                        "voids removed from the voxel domain" if args.voids == "removed"
                                                        else "voids as density-0 elements"),
        "config": {
            "workload": "batched mixed-precision MG-PCG (%s multigrid): %d uneven problems solved together to true "
                        "residual < %.0e (one launch per kernel for the batch, hipGraph-replayed iterations, "
                        "autotuned tiles); a step solves all %d"
                        % ("Galerkin" if galerkin else "reference", len(probs), btol, len(probs)),
            "multigrid": args.mg,
            "problems": len(probs),
            "elements": ne,
            "free_dofs": nF,
            "tiles_kw_nz_rows_per_block": list(tiles),
            "iterations": [int(v) for v in its],
            "sequential_iterations": [int(v) for v in its_seq],
            "max_true_residual": max(res),
            "sequential_ms": ms_seq,
            "sequential_solves_per_s": len(probs) * 1e3 / ms_seq if ms_seq else None,
            "speedup_vs_sequential": ms_seq / ms_batch if ms_seq else None,
            "kernel_timing": ktiming,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def timed_solves(t3d, mx, db, dsol, tol, maxit, true_every, warmup, steps):
    """warmup + steps mixed solves from x0 = 0; (iterations, ms per solve)"""
    it = 0
    for _ in range(warmup):
        dsol.zero()
        it = mx.pcg(db, dsol, tol, maxit, trueEvery=true_every)
    t3d.lib().t3d_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dsol.zero()
        it = mx.pcg(db, dsol, tol, maxit, trueEvery=true_every)
    t3d.lib().t3d_synchronize()
    return it, (time.perf_counter() - t0) / steps * 1e3


def fp64_solve(t3d, p, M, db, dsol, tol, maxit):
    dsol.zero()
    t3d.lib().t3d_synchronize()
    t0 = time.perf_counter()
    it = t3d.pcg(p, M, db, dsol, tol, maxit)
    t3d.lib().t3d_synchronize()
    return it, (time.perf_counter() - t0) * 1e3


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import algorithmengineering_altproject_amd as t3d
    name, cus, mem = t3d.device_info()
    log("device %s, %d CUs" % (name, cus))
    if args.batch:
        return main_batch(args, t3d)
    if args.stl:
        return main_stl(args, t3d)
    if args.loop:
        return main_loop(args, t3d)
    galerkin = args.mg == "galerkin"
    tol = args.tol if args.tol is not None else (1e-6 if galerkin else args.ref_tol)
    te = args.true_every if args.true_every is not None else (25 if galerkin else 50)
    t0 = time.time()
    p = t3d.Problem(args.nely, args.nelx, args.nelz)
    log("problem %dx%dx%d: %d elements, %d free DOFs (setup %.1fs)" % (args.nely, args.nelx, args.nelz, p.ne, p.nF,
                                                                         time.time() - t0))
    # SIMP moduli (driver.cpp:117-121), Emin = 1e-9, of the uniform design rho = V0
    # or of seeded random densities smoothed by the PDE filter (rmin = 2)
    if args.design == "uniform":
        dx = t3d.DeviceArray.from_host(np.full(p.ne, np.float32(0.12), np.float32))
    else:
        raw = t3d.DeviceArray.from_host(np.random.default_rng(args.seed).uniform(1e-3, 1.0, p.ne).astype(np.float32))
        dx = t3d.DeviceArray(p.ne, np.float32)
        t3d.Filter(p, 2.0).apply(raw, dx)
    dE = t3d.DeviceArray(p.ne)
    t3d.simp(p, dx, 1.0, float(np.float32(1e-9)), dE)
    p.set_modulus(dE)
    du = t3d.DeviceArray(p.nDOF)
    dy = t3d.DeviceArray(p.nDOF)
    if args.copy_only:
        du.zero()
        ms = p.time_op(1, du, dy, args.kxu_reps)
        log("stream copy: %.4f ms, %.0f GB/s" % (ms, 2 * 8 * p.nDOF / ms / 1e6))
        return

    t0 = time.time()
    M = t3d.MG(p, Emin=float(np.float32(1e-9)))   # the C++ port's V-cycle (bit-exact FP64 path)
    M.setup(dE)
    mx = t3d.Mixed(p, M)
    mx.setup()
    form = {"walsh": 0, "mfma3": 2, "mfma6": 3}[args.operator]
    mx.set_operator(form)
    log("MG: levels %s, coarsest Cholesky N=%d (setup %.2fs)" % (M.dims, M.info()[0], time.time() - t0))
    Mg = mxg = None
    gal_setup_s = gal_resetup_s = None
    if galerkin:
        t0 = time.time()
        Mg = t3d.MG(p, Emin=float(np.float32(1e-9)), galerkin=True, standard=args.vcycle == "standard",
                    nonDyadic=args.non_dyadic, maxLevels=args.levels)
        Mg.setup(dE)
        mxg = t3d.Mixed(p, Mg)
        mxg.setup()
        mxg.set_operator(form)
        gal_setup_s = time.time() - t0
        # a repeated set-up (buffers allocated): what an optimisation loop pays per design
        t0 = time.time()
        Mg.setup(dE)
        mxg.setup()
        gal_resetup_s = time.time() - t0
        log("Galerkin MG: coarsest Cholesky N=%d (first set-up %.2fs, repeated %.3fs)"
            % (Mg.info()[0], gal_setup_s, gal_resetup_s))

    # ---- K*u blocks: HIP events around back-to-back launches on the problem stream
    uh = np.random.default_rng(0).uniform(-1, 1, p.nF)
    p.free_to_full(t3d.DeviceArray.from_host(uh), du)
    p.time_op(0, du, dy, 3)
    kxu64_ms = p.time_op(0, du, dy, args.kxu_reps)
    u32 = t3d.DeviceArray.from_host(du.download().astype(np.float32))
    y32 = t3d.DeviceArray(p.nDOF, np.float32)
    mx.time_kxu(u32, y32, 3, 0)
    kxu32_ms = mx.time_kxu(u32, y32, args.kxu_reps, 0)
    mx.time_kxu(u32, y32, 3, 2)
    kxum_ms = mx.time_kxu(u32, y32, args.kxu_reps, 2)
    mx.time_kxu(u32, y32, 3, 3)
    kxum6_ms = mx.time_kxu(u32, y32, args.kxu_reps, 3)
    mx.time_kxu(u32, y32, 3, 4)
    fused_ms = mx.time_kxu(u32, y32, args.kxu_reps, 4)
    # streaming-copy rate of this GPU (FP64 vector copy, 2 x 8 B per DOF): boxes differ by up to ~20 %
    p.time_op(1, du, dy, 3)
    copy_ms = p.time_op(1, du, dy, args.kxu_reps)
    copy_gbps = 2 * 8 * p.nDOF / copy_ms / 1e6
    log("stream copy on this GPU: %.0f GB/s" % copy_gbps)
    # the fused kernel's own streams (read u, z, dx, moduli; write two vectors and dx) as a plain float4 copy
    mx.time_kxu(u32, y32, 3, 7)
    copyf_ms = mx.time_kxu(u32, y32, args.kxu_reps, 7)
    copyf_gbps = (16 * 6 * (p.nDOF // 4) + 16 * (p.ne // 4)) / copyf_ms / 1e6
    log("copy with the fused kernel's streams (3 in + 3 out + moduli): %.0f GB/s" % copyf_gbps)
    b64 = 8 * (2 * p.nF + p.ne)       # read u, write y (free DOFs), read E
    b32 = 4 * (2 * p.nF + p.ne)
    bfused = 4 * (6 * p.nF + p.ne)    # read p_prev, z, dx; write p, y, dx; read E
    log("K*u FP64 (plain, bit-exact): %.4f ms/apply, %.0f GB/s algorithmic, %.2f TFLOP/s"
        % (kxu64_ms, b64 / kxu64_ms / 1e6, 1152 * p.ne / kxu64_ms / 1e9))
    log("K*u FP32 Walsh form: %.4f ms/apply, %.0f GB/s algorithmic" % (kxu32_ms, b32 / kxu32_ms / 1e6))
    log("K*u FP32 MFMA bf16 x3: %.4f ms/apply, %.0f GB/s algorithmic, %.1f TFLOP/s bf16 MFMA issued"
        % (kxum_ms, b32 / kxum_ms / 1e6, mfma_flops(p, 3) / kxum_ms / 1e9))
    log("K*u FP32 MFMA bf16 x6: %.4f ms/apply, %.0f GB/s algorithmic, %.1f TFLOP/s bf16 MFMA issued"
        % (kxum6_ms, b32 / kxum6_ms / 1e6, mfma_flops(p, 6) / kxum6_ms / 1e9))
    log("K*u FP32 fused (solver kernel, %s): %.4f ms/launch, %.0f GB/s algorithmic"
        % (args.operator, fused_ms, bfused / fused_ms / 1e6))
    if args.kxu_only:
        if args.operator == "walsh" and args.tile_order_compare:   # the other tile order
            mx.time_kxu(u32, y32, 3, 5)
            alt_f = mx.time_kxu(u32, y32, args.kxu_reps, 5)
            alt_p = mx.time_kxu(u32, y32, args.kxu_reps, 6)
            log("Walsh form, other tile order: fused %.4f ms (%.0f GB/s), operator %.4f ms (%.0f GB/s)"
                % (alt_f, bfused / alt_f / 1e6, alt_p, b32 / alt_p / 1e6))
        return

    db = t3d.DeviceArray(p.nDOF)
    p.load_vector(db)
    dsol = t3d.DeviceArray(p.nDOF)
    dce = t3d.DeviceArray(p.ne)
    comp = lambda: t3d.compliance(p, dE, dsol, dce)
    # ---- the timed steps: mixed-precision solves with the selected multigrid
    head_mx = mxg if galerkin else mx
    for _ in range(args.warmup):
        dsol.zero()
        head_mx.pcg(db, dsol, tol, args.maxit, trueEvery=te)
    head_mx.set_timing(True)   # HIP events around every fused operator launch of the timed solves
    it_m, ms_step = timed_solves(t3d, head_mx, db, dsol, tol, args.maxit, te, 0, args.steps)
    k_ms, k_n = head_mx.timing()
    head_mx.set_timing(False)
    fused_solve_ms = k_ms / max(1, k_n)
    Cm = comp()
    res_m = true_residual(t3d, p, db, dsol)
    log("mixed PCG (%s MG): %d iterations/solve%s, %.1f ms/solve, %.3f ms/iteration, compliance %.9e, "
        "true residual %.3e" % (args.mg, it_m, " (capped)" if it_m >= args.maxit else "", ms_step, ms_step / it_m,
                                Cm, res_m))
    log("fused K*u inside the timed solves: %d launches, %.4f ms average, %.0f GB/s algorithmic"
        % (k_n, fused_solve_ms, bfused / fused_solve_ms / 1e6))
    # converged compliance of this load: FP64 PCG with the Galerkin V-cycle to 1e-10
    C_star = None
    if galerkin and not args.no_fp64:
        it_s, _ = fp64_solve(t3d, p, Mg, db, dsol, 1e-10, 2000)
        C_star = comp()
        log("converged compliance (FP64, Galerkin MG, tol 1e-10, %d iterations): %.12e" % (it_s, C_star))
    rel = (lambda C: abs(C - C_star) / abs(C_star)) if C_star else (lambda C: None)

    fp64_same = None
    if galerkin and not args.no_fp64:   # the same preconditioner in FP64: the precision comparison
        it64g, ms64g = fp64_solve(t3d, p, Mg, db, dsol, tol, args.maxit)
        C64g = comp()
        res64g = true_residual(t3d, p, db, dsol)
        fp64_same = {"solves_per_s": 1e3 / ms64g, "ms_per_solve": ms64g, "pcg_iters": it64g,
                     "ms_per_iteration": ms64g / it64g, "compliance": C64g, "true_residual": res64g,
                     "compliance_rel_err": rel(C64g), "speedup_mixed_vs_fp64": ms64g / ms_step,
                     "compliance_rel_diff_mixed": abs(Cm - C64g) / abs(C64g)}
        log("FP64 PCG (Galerkin MG): %d iterations, %.1f ms/solve, compliance %.9e, true residual %.3e; "
            "mixed speed-up %.2fx" % (it64g, ms64g, C64g, res64g, ms64g / ms_step))
        # the same hierarchy with the other V-cycle (FP64)
        other = "adapted" if args.vcycle == "standard" else "standard"
        t3d.check(t3d.lib().t3d_mg_set_vcycle(Mg.h, 1 if other == "standard" else 0), "t3d_mg_set_vcycle")
        it64s, ms64s = fp64_solve(t3d, p, Mg, db, dsol, tol, args.maxit)
        C64s = comp()
        res64s = true_residual(t3d, p, db, dsol)
        t3d.check(t3d.lib().t3d_mg_set_vcycle(Mg.h, 1 if other == "adapted" else 0), "t3d_mg_set_vcycle")
        fp64_same[other + "_vcycle"] = {"ms_per_solve": ms64s, "pcg_iters": it64s, "ms_per_iteration": ms64s / it64s,
                                        "compliance_rel_err": rel(C64s), "true_residual": res64s}
        log("FP64 PCG (Galerkin MG, '%s' V-cycle): %d iterations, %.1f ms/solve, %.3f ms/iteration, "
            "true residual %.3e" % (other, it64s, ms64s, ms64s / it64s, res64s))

    dyadic = None
    if galerkin and not args.no_fp64 and not args.no_dyadic_check and args.non_dyadic == 1:
        # the same solve on the hierarchy coarsened by 2 at every level (the MATLAB original's nonDyadic_ = 0),
        # 6 levels down to the same coarsest grid: the mixed solver and the FP64 solver with that multigrid
        t0 = time.time()
        Md = t3d.MG(p, Emin=float(np.float32(1e-9)), galerkin=True, standard=args.vcycle == "standard",
                    nonDyadic=0, maxLevels=args.levels + 1)
        Md.setup(dE)
        mxd = t3d.Mixed(p, Md)
        mxd.setup()
        mxd.set_operator(form)
        setup_d = time.time() - t0
        itd, msd = timed_solves(t3d, mxd, db, dsol, tol, args.maxit, te, 1, args.steps)
        Cd = comp()
        resd = true_residual(t3d, p, db, dsol)
        it64d, ms64d = fp64_solve(t3d, p, Md, db, dsol, tol, args.maxit)
        C64d = comp()
        dyadic = {"levels": [list(d) for d in Md.dims], "coarsest_dofs": Md.info()[0], "setup_s": setup_d,
                  "mixed": {"solves_per_s": 1e3 / msd, "ms_per_solve": msd, "pcg_iters": itd,
                            "ms_per_iteration": msd / itd, "compliance_rel_err": rel(Cd), "true_residual": resd},
                  "fp64": {"ms_per_solve": ms64d, "pcg_iters": it64d, "compliance_rel_err": rel(C64d)},
                  "speedup_mixed_vs_fp64": ms64d / msd}
        log("Galerkin MG coarsened by 2 at every level (%d levels): mixed %d iterations %.1f ms/solve "
            "(%.3f ms/iteration), FP64 %d iterations %.1f ms/solve" % (Md.nlev, itd, msd, msd / itd, it64d, ms64d))
        del mxd, Md

    fp64 = None
    if not args.no_fp64:   # the FP64 plain path: the reference's algorithm, bit-exact, at the reference's tol
        it64, ms64 = fp64_solve(t3d, p, M, db, dsol, args.ref_tol, args.maxit)
        C64 = comp()
        res64 = true_residual(t3d, p, db, dsol)
        log("FP64 plain PCG (reference V-cycle, tol %.0e): %d iterations%s, %.1f ms/solve, %.3f ms/iteration, "
            "compliance %.9e, true residual %.3e" % (args.ref_tol, it64, " (capped)" if it64 >= args.maxit else "",
                                                     ms64, ms64 / it64, C64, res64))
        fp64 = {"tol": args.ref_tol, "solves_per_s": 1e3 / ms64, "ms_per_solve": ms64, "pcg_iters": it64,
                "capped": it64 >= args.maxit, "ms_per_iteration": ms64 / it64, "compliance": C64,
                "true_residual": res64, "compliance_rel_err": rel(C64), "speedup_mixed_vs_fp64": ms64 / ms_step,
                "compliance_rel_diff_mixed": abs(Cm - C64) / abs(C64)}
    mixed_ref = None
    if galerkin:   # the mixed solver with the reference's V-cycle at the reference's tol
        itr, msr = timed_solves(t3d, mx, db, dsol, args.ref_tol, args.maxit, 50, 0, 1)
        Cr = comp()
        mixed_ref = {"tol": args.ref_tol, "solves_per_s": 1e3 / msr, "ms_per_solve": msr, "pcg_iters": itr,
                     "ms_per_iteration": msr / itr, "compliance": Cr, "compliance_rel_err": rel(Cr),
                     "true_residual": true_residual(t3d, p, db, dsol)}
        log("mixed PCG (reference V-cycle, tol %.0e): %d iterations, %.1f ms/solve, %.3f ms/iteration"
            % (args.ref_tol, itr, msr, msr / itr))
    tight = None
    if not galerkin and not args.no_tight and not args.no_fp64:
        dsol.zero()
        itm_t = mx.pcg(db, dsol, args.tight_tol, args.maxit, trueEvery=te)
        Cm_t = comp()
        dsol.zero()
        it64_t = t3d.pcg(p, M, db, dsol, args.tight_tol, args.maxit)
        C64_t = comp()
        res64_t = true_residual(t3d, p, db, dsol)
        tight = {"tol": args.tight_tol, "mixed_iters": itm_t, "fp64_iters": it64_t, "mixed_compliance": Cm_t,
                 "fp64_compliance": C64_t, "fp64_true_residual": res64_t,
                 "compliance_rel_diff": abs(Cm_t - C64_t) / abs(C64_t)}
        log("tol %.0e: mixed %d it C=%.9e, FP64 %d it C=%.9e (true residual %.2e), rel diff %.2e"
            % (args.tight_tol, itm_t, Cm_t, it64_t, C64_t, res64_t, tight["compliance_rel_diff"]))

    small = None
    if not args.no_fp64:   # the FP64 plain path on the 96x192x96 cantilever, as the CPU baseline's whole solve
        p1 = t3d.Problem(96, 192, 96)
        x1 = t3d.DeviceArray.from_host(np.full(p1.ne, np.float32(0.12), np.float32))
        E1 = t3d.DeviceArray(p1.ne)
        t3d.simp(p1, x1, 1.0, float(np.float32(1e-9)), E1)
        M1 = t3d.MG(p1, Emin=float(np.float32(1e-9)))
        M1.setup(E1)
        b1 = t3d.DeviceArray(p1.nDOF)
        p1.load_vector(b1)
        s1 = t3d.DeviceArray(p1.nDOF)
        fp64_solve(t3d, p1, M1, b1, s1, args.ref_tol, 800)   # warm-up
        it1, ms1 = fp64_solve(t3d, p1, M1, b1, s1, args.ref_tol, 800)
        small = {"iterations": it1, "ms_per_solve": ms1, "tol": args.ref_tol, "max_iters": 800}
        log("FP64 plain PCG at 96x192x96 (tol %.0e, cap 800): %d iterations, %.1f ms/solve" % (args.ref_tol, it1, ms1))
        del p1, x1, E1, M1, b1, s1
    cpu = None
    if not args.no_cpu:
        try:
            with stdout_to_stderr():
                cpu = cpu_baseline(args, dE.download(), fp64["pcg_iters"] if fp64 else it_m)
            if cpu:
                log("CPU reference: %s" % cpu["sample"])
        except Exception as e:  # the baseline is reported, never required
            log("CPU baseline skipped: %r" % (e,))
    traffic = None
    tpath = os.path.join(REPO, "profiles", "kxu_f32_%s_%dx%dx%d_traffic.json"
                         % ("fused" if form == 0 else "mfma%d_fused" % (3 if form == 2 else 6), args.nely,
                            args.nelx, args.nelz))
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    opname = {0: "Walsh-form", 2: "MFMA bf16 x3", 3: "MFMA bf16 x6"}[form]
    levels = ("the C++ port's levels (%d)" % Mg.nlev if args.non_dyadic == 1
              else "%d levels coarsened by 2" % Mg.nlev) if galerkin else ""
    mgname = ("Galerkin multigrid of the MATLAB original ('%s' V-cycle, P^T restriction, coarsest Cholesky) "
              "on %s" % (args.vcycle.capitalize(), levels) if galerkin else "the C++ port's additive V-cycle")
    line = {
        "metric": METRIC,
        "value": 1e3 / ms_step,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": ("synthetic (uniform rho=0.12 SIMP moduli, reference cantilever BCs/load)" if args.design == "uniform"
                 else "synthetic (rho ~ U[1e-3,1], seed %d, PDE-filtered rmin=2; SIMP moduli; reference "
                      "cantilever BCs/load)" % args.seed),
        "config": {
            "workload": "MG-PCG solve from x0=0 to true residual < %.0e, cuboid %dx%dx%d (nely x nelx x nelz), "
                        "%d free DOFs; mixed precision: FP32 %s K*u and FP32 V-cycle of the %s, FP64 dots/alpha/"
                        "beta/true residual/coarsest solve, flexible beta, residual replacement every %d iterations"
                        % (tol, args.nely, args.nelx, args.nelz, p.nF, opname, mgname, te),
            "multigrid": args.mg,
            "tol": tol,
            "max_iters": args.maxit,
            "pcg_iters": it_m,
            "capped": it_m >= args.maxit,
            "ms_per_iteration": ms_step / it_m,
            "dof_per_s": p.nF * it_m / (ms_step / 1e3),
            "compliance": Cm,
            "compliance_converged": C_star,
            "compliance_rel_err": rel(Cm),
            "true_residual": res_m,
            "galerkin_setup_s": gal_setup_s,
            "galerkin_resetup_s": gal_resetup_s,
            "fp64_same_multigrid": fp64_same,
            "hierarchy": {"nonDyadic": args.non_dyadic, "levels": [list(d) for d in Mg.dims]} if galerkin else None,
            "galerkin_dyadic_hierarchy": dyadic,
            "fp64_plain": fp64,
            "fp64_plain_96x192x96": small,
            "mixed_reference_vcycle": mixed_ref,
            "tight_tolerance_check": tight,
            "kxu_fp64_ms": kxu64_ms,
            "kxu_fp64_gbps": b64 / kxu64_ms / 1e6,
            "kxu_fp32_walsh_ms": kxu32_ms,
            "kxu_fp32_walsh_gbps": b32 / kxu32_ms / 1e6,
            "kxu_fp32_mfma_bf16x3_ms": kxum_ms,
            "kxu_fp32_mfma_bf16x3_bf16_tflops_issued": mfma_flops(p, 3) / kxum_ms / 1e9,
            "kxu_fp32_mfma_bf16x6_ms": kxum6_ms,
            "kxu_fp32_mfma_bf16x6_bf16_tflops_issued": mfma_flops(p, 6) / kxum6_ms / 1e9,
        },
        "roofline": {
            "kernel": ("fused FP32 Walsh-form K*u of the mixed solve (k_kxw32<16,true>)" if form == 0
                       else "fused FP32 MFMA bf16 x%d K*u of the mixed solve (k_kxw32<16,true,16,%d>)"
                       % (3 if form == 2 else 6, form)),
            "bound": "hbm",
            "achieved": bfused / fused_solve_ms / 1e6,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": bfused / fused_solve_ms / 1e6 / HBM_PEAK_GBPS,
            "traffic": traffic,
            "ms_per_launch": fused_solve_ms,
            "launches_timed": k_n,
            "timing": "HIP events around every launch of the kernel in the timed solves, on the solver's stream",
            "ms_per_launch_back_to_back": fused_ms,
            "stream_copy_gbps_this_gpu": copy_gbps,
            "frac_of_stream_copy": bfused / fused_solve_ms / 1e6 / copy_gbps,
            "copy_of_fused_streams_gbps": copyf_gbps,
            "frac_of_copy_of_fused_streams": bfused / fused_solve_ms / 1e6 / copyf_gbps,
            "algorithmic_bytes": bfused,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
