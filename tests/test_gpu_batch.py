"""Batched mixed-precision MG-PCG (t3d_batch_*): many uneven problems solved
together (one launch per kernel for the whole batch, hipGraph-replayed
iterations, autotuned tiles) against the same problems solved one by one with
t3d_pcg_mixed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def element_coords(p):
    """(ex, ey, ez) of every compact element (eleMapBack = resY*resX*ez + resY*ex + (resY-1-ey))."""
    f = p.model()["eleMapBack"].astype(np.int64)
    ny, nx = p.resY, p.resX
    ez = f // (ny * nx)
    rem = f % (ny * nx)
    return rem // ny, ny - 1 - rem % ny, ez


def design(t3d, p, V0, spheres=()):
    """SIMP moduli of the uniform design V0 with spherical voids (density 0,
    passive) of the given (cx, cy, cz, r) in element units."""
    x = np.full(p.ne, V0, np.float32)
    if spheres:
        ex, ey, ez = element_coords(p)
        for cx, cy, cz, r in spheres:
            x[(ex + 0.5 - cx) ** 2 + (ey + 0.5 - cy) ** 2 + (ez + 0.5 - cz) ** 2 <= r * r] = 0.0
    dx = t3d.DeviceArray.from_host(x)
    dE = t3d.DeviceArray(p.ne)
    t3d.simp(p, dx, 1.0, 1e-9, dE)
    return dE


def true_rel_residual(t3d, p, db, dx):
    dy = t3d.DeviceArray(p.nDOF)
    p.kxu_f64(dx, dy)
    b = db.download()
    return np.linalg.norm(b - dy.download()) / np.linalg.norm(b)


DIMS = [(8, 16, 8), (16, 32, 16), (24, 48, 24), (12, 24, 16), (16, 48, 16), (8, 24, 8)]
V0S = [0.12, 0.2, 0.3, 0.08, 0.15, 0.25]
VOIDS = [(), ((24, 8, 8, 5),), (), ((12, 6, 8, 3), (6, 6, 8, 3)), (), ()]


def build(t3d, dims, v0s, voids, galerkin=False):
    B = t3d.Batch(dims, galerkin=galerkin)
    dEs = [design(t3d, p, v, s) for p, v, s in zip(B.problems, v0s, voids)]
    B.setup(dEs)
    dbs = []
    for p in B.problems:
        db = t3d.DeviceArray(p.nDOF)
        p.load_vector(db)
        dbs.append(db)
    return B, dEs, dbs


def sequential(t3d, dims, v0s, voids, tol, maxIt, galerkin=False, trueEvery=50):
    out = []
    for d, v, s in zip(dims, v0s, voids):
        p = t3d.Problem(*d)
        dE = design(t3d, p, v, s)
        M = t3d.MG(p, Emin=1e-9, galerkin=galerkin)
        M.setup(dE)
        mx = t3d.Mixed(p, M)
        mx.setup()
        db = t3d.DeviceArray(p.nDOF)
        p.load_vector(db)
        x = t3d.DeviceArray(p.nDOF)
        x.zero()
        it = mx.pcg(db, x, tol, maxIt, trueEvery=trueEvery)
        out.append((it, x.download()))
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_sequential(t3d, graph):
    """With the single solver's tiles (16 planes, 4 rows per block) every
    problem of the batch runs the single solver's arithmetic: the same
    iteration count and a bit-identical solution (problems whose first
    convergence is confirmed by the true residual)."""
    tol, maxIt = 1e-3, 3000
    B, dEs, dbs = build(t3d, DIMS, V0S, VOIDS)
    B.set_tiles(16, 4)
    xs = [t3d.DeviceArray(p.nDOF) for p in B.problems]
    for x in xs:
        x.zero()
    its = B.solve(dbs, xs, tol, maxIt, useGraph=graph)
    ref = sequential(t3d, DIMS, V0S, VOIDS, tol, maxIt)
    for i, (p, x, db) in enumerate(zip(B.problems, xs, dbs)):
        it_ref, x_ref = ref[i]
        print(DIMS[i], "batch", its[i], "sequential", it_ref)
        assert its[i] == it_ref
        assert np.array_equal(x.download(), x_ref)
        assert true_rel_residual(t3d, p, db, x) < tol


def test_batch_autotuned(t3d):
    """Autotuned tiles: every problem converges to the true-residual tolerance
    in about the sequential iteration count, the chosen tiles are among the
    candidates, and a second solve reuses them and reproduces the first."""
    tol, maxIt = 1e-3, 3000
    B, dEs, dbs = build(t3d, DIMS, V0S, VOIDS)
    xs = [t3d.DeviceArray(p.nDOF) for p in B.problems]
    for x in xs:
        x.zero()
    its = B.solve(dbs, xs, tol, maxIt)
    nz, rpb = B.info()
    assert nz in (8, 16, 32) and rpb in (2, 4, 8)
    first = [x.download() for x in xs]
    ref = sequential(t3d, DIMS, V0S, VOIDS, tol, maxIt)
    for i, (p, x, db) in enumerate(zip(B.problems, xs, dbs)):
        print(DIMS[i], "tiles", nz, rpb, "batch", its[i], "sequential", ref[i][0])
        assert its[i] < maxIt
        assert true_rel_residual(t3d, p, db, x) < tol
        assert abs(its[i] - ref[i][0]) <= max(5, 0.1 * ref[i][0])
    for x in xs:
        x.zero()
    its2 = B.solve(dbs, xs, tol, maxIt)
    assert list(its2) == list(its)
    for x, f in zip(xs, first):
        assert np.array_equal(x.download(), f)


def test_batch_galerkin_matches_sequential(t3d):
    """Every problem with the Galerkin multigrid: the batch runs the single
    solver's arithmetic (same iterations, bit-identical solutions) and reaches
    1e-6 in tens of iterations."""
    tol, maxIt, te = 1e-6, 500, 25
    B, dEs, dbs = build(t3d, DIMS, V0S, VOIDS, galerkin=True)
    B.set_tiles(16, 4)
    xs = [t3d.DeviceArray(p.nDOF) for p in B.problems]
    for x in xs:
        x.zero()
    its = B.solve(dbs, xs, tol, maxIt, trueEvery=te)
    ref = sequential(t3d, DIMS, V0S, VOIDS, tol, maxIt, galerkin=True, trueEvery=te)
    for i, (p, x, db) in enumerate(zip(B.problems, xs, dbs)):
        it_ref, x_ref = ref[i]
        print(DIMS[i], "batch", its[i], "sequential", it_ref)
        assert its[i] == it_ref and its[i] < 150
        assert np.array_equal(x.download(), x_ref)
        assert true_rel_residual(t3d, p, db, x) < tol


def test_batch_timing_leaves_results_unchanged(t3d):
    """t3d_batch_set_timing: the graph replays and, without graphs, every kernel
    of the lockstep iterations are timed with HIP events; timing changes no
    result (same iterations, bit-identical solutions as an untimed solve), the
    per-kernel launch counts agree with the iterations run, and the times are
    positive."""
    tol, maxIt, te = 1e-6, 500, 25
    B, dEs, dbs = build(t3d, DIMS, V0S, VOIDS, galerkin=True)
    B.set_tiles(16, 4)
    xs = [t3d.DeviceArray(p.nDOF) for p in B.problems]
    for x in xs:
        x.zero()
    its = B.solve(dbs, xs, tol, maxIt, trueEvery=te, useGraph=True)
    ref = [x.download() for x in xs]
    for graph in (True, False):
        B.set_timing(True)
        for x in xs:
            x.zero()
        its_t = B.solve(dbs, xs, tol, maxIt, trueEvery=te, useGraph=graph)
        t = B.timing()
        B.set_timing(False)
        assert list(its_t) == list(its)
        for x, r in zip(xs, ref):
            assert np.array_equal(x.download(), r)
        # lockstep iterations run: the slowest problem's, less the residual
        # replacements, up to the end of its chunk of 32 iterations
        n_plain = max(its) - max(its) // te
        n_max = -(-(max(its) + 1) // 32) * 32
        if graph:
            ms, n = t["graph"]
            assert n_plain - 2 <= n <= n_max and ms > 0.0
            assert len(t) == 1
        else:
            assert t["graph"][1] == 0
            for name in ("kb_kxw32", "kb_fin_alpha", "kb_xr_rows", "kb_fin_rn2", "kb_restrict_yz",
                         "kb_coarse_tiles", "kb_coarse_sum", "kb_pro_rows", "kb_fin_beta"):
                ms, n = t[name]
                assert ms > 0.0, name
            n_kw = t["kb_kxw32"][1]
            assert n_plain - 2 <= n_kw <= n_max
            assert t["kb_pro_rows"][1] == n_kw and t["kb_coarse_tiles"][1] == n_kw
