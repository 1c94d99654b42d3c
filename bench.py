#!/usr/bin/env python3
"""Benchmark of the TOP3D_XL finite-element inner loop on one MI355X.

A step is one multigrid-preconditioned PCG solve (PCG_free, cpp/src/solver.cpp:87-217)
from x0 = 0 on the 192x384x192 cuboid (14.2M voxels, 42.9M free DOFs), with the
reference's boundary conditions and load, element moduli from SIMP of the uniform
initial design rho = 0.12, and the MG preconditioner built for those moduli
(built once, outside the timed region).  Tolerance and iteration cap are the
reference's defaults (cgTol = 1e-3f, cgMaxIt = 800).

Prints one JSON line.  `value` is solves/s of the FP64 plain path (inputs already
resident in HBM).  The line also carries the K*u roofline (HIP-event timing of the
operator on its stream) and the reference CPU code timed on this host.
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
    ap.add_argument("--tol", type=float, default=float(np.float32(1e-3)))
    ap.add_argument("--maxit", type=int, default=800)
    ap.add_argument("--kxu-reps", type=int, default=50)
    ap.add_argument("--kxu-only", action="store_true", help="only the K*u timing block (for counter runs)")
    ap.add_argument("--copy-only", action="store_true", help="only the streaming-copy calibration kernel")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=4, help="reference PCG iterations timed on the CPU")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(args, E, gpu_iters):
    """The reference's own OpenMP code (oracle/_ref, -O3 -ffast-math -fopenmp) on this
    host: a bounded sample of `cpu_iters` MG-PCG iterations on the same problem,
    scaled to solves/s with the GPU's iteration count for one solve (the two paths
    run the same iterations)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    if not O.have_ref(fast=True):
        return None
    R = O.ref(fast=True)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    import ctypes as C
    t0 = time.time()
    sizes = np.zeros(7, np.int32)
    R.ref_sizes(args.nely, args.nelx, args.nelz, sizes)   # builds the reference model + MG statics
    t_setup = time.time() - t0
    it = C.c_int()
    ms = R.ref_time_pcg(args.nely, args.nelx, args.nelz, np.ascontiguousarray(E), 1e-6, 0.0, args.cpu_iters,
                        C.byref(it))
    per_it = ms / 1e3 / it.value
    u = np.random.default_rng(1).uniform(-1, 1, int(sizes[5]))
    kxu_ms = R.ref_time_kxu(args.nely, args.nelx, args.nelz, np.ascontiguousarray(E), u, 3)
    return {
        "value": 1.0 / (per_it * gpu_iters),
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": "reference PCG_free + MG V-cycle (oracle/_ref, OpenMP -O3 -ffast-math), %d iterations on the "
                  "same 192x384x192 problem = %.3f s/iteration, scaled by the %d iterations of one solve; "
                  "reference K_times_u_finest %.1f ms/apply; model setup %.1f s not timed"
                  % (it.value, per_it, gpu_iters, kxu_ms, t_setup),
        "kxu_ms": kxu_ms,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import algorithmengineering_altproject_amd as t3d
    name, cus, mem = t3d.device_info()
    log("device %s, %d CUs" % (name, cus))
    t0 = time.time()
    p = t3d.Problem(args.nely, args.nelx, args.nelz)
    log("problem %dx%dx%d: %d elements, %d free DOFs (setup %.1fs)" % (args.nely, args.nelx, args.nelz, p.ne, p.nF,
                                                                         time.time() - t0))
    # SIMP moduli of the uniform design rho = V0 (driver.cpp:117-121), Emin = 1e-9
    dx = t3d.DeviceArray.from_host(np.full(p.ne, np.float32(0.12), np.float32))
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
    # ---- K*u roofline block: HIP events around back-to-back launches on the problem stream
    p.free_to_full(t3d.DeviceArray.from_host(np.random.default_rng(0).uniform(-1, 1, p.nF)), du)
    p.time_op(0, du, dy, 3)
    kxu_ms = p.time_op(0, du, dy, args.kxu_reps)
    kxu_bytes = 8 * (2 * p.nF + p.ne)          # algorithmic bytes: read u, write y (free DOFs), read E
    kxu_flops = 1152 * p.ne
    kxu_gbps = kxu_bytes / kxu_ms / 1e6
    log("K*u FP64: %.4f ms/apply, %.1f GB/s algorithmic, %.2f TFLOP/s" % (kxu_ms, kxu_gbps,
                                                                         kxu_flops / kxu_ms / 1e9))
    if args.kxu_only:
        return

    t0 = time.time()
    M = t3d.MG(p, Emin=float(np.float32(1e-9)))
    M.setup(dE)
    log("MG: levels %s, coarsest Cholesky N=%d (setup %.2fs)" % (M.dims, M.info()[0], time.time() - t0))
    db = t3d.DeviceArray(p.nDOF)
    p.load_vector(db)
    dsol = t3d.DeviceArray(p.nDOF)
    iters = []
    for _ in range(args.warmup):
        dsol.zero()
        iters.append(t3d.pcg(p, M, db, dsol, args.tol, args.maxit))
    t3d.lib().t3d_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dsol.zero()
        iters.append(t3d.pcg(p, M, db, dsol, args.tol, args.maxit))
    t3d.lib().t3d_synchronize()
    dt = time.perf_counter() - t0
    it_solve = iters[-1]
    ms_step = dt / args.steps * 1e3
    dce = t3d.DeviceArray(p.ne)
    C = t3d.compliance(p, dE, dsol, dce)
    log("PCG: %d iterations/solve%s, %.1f ms/solve, %.3f ms/iteration, compliance %.9e"
        % (it_solve, " (capped)" if it_solve >= args.maxit else "", ms_step, ms_step / it_solve, C))

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, dE.download(), it_solve)
            if cpu:
                log("CPU reference: %s" % cpu["sample"])
        except Exception as e:  # the baseline is reported, never required
            log("CPU baseline skipped: %r" % (e,))
    traffic = None
    tpath = os.path.join(REPO, "profiles", "kxu_f64_192x384x192_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
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
        "dtype": "f64",
        "data": "synthetic (uniform rho=0.12 SIMP moduli, reference cantilever BCs/load)",
        "config": {
            "workload": "MG-PCG solve from x0=0, cuboid %dx%dx%d (nely x nelx x nelz), %d free DOFs, "
                        "FP64 plain path (bit-identical to the reference's operation order)"
                        % (args.nely, args.nelx, args.nelz, p.nF),
            "tol": args.tol,
            "max_iters": args.maxit,
            "pcg_iters": it_solve,
            "capped": it_solve >= args.maxit,
            "ms_per_iteration": ms_step / it_solve,
            "dof_per_s": p.nF * it_solve / (ms_step / 1e3),
            "compliance": C,
        },
        "roofline": {
            "kernel": "K*u FP64 (k_kxu_f64_lds)",
            "bound": "hbm",
            "achieved": kxu_gbps,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": kxu_gbps / HBM_PEAK_GBPS,
            "traffic": traffic,
            "ms_per_launch": kxu_ms,
            "algorithmic_bytes": kxu_bytes,
            "fp64_tflops": kxu_flops / kxu_ms / 1e9,
        },
        "cpu_baseline": ({k: v for k, v in cpu.items() if k != "kxu_ms"} if cpu else None),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
