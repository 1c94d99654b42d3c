"""The oracle against the reference itself (oracle/_ref: the reference's own
cpp/src sources compiled by oracle/Makefile, strict build, one OpenMP thread).

Bit-identical where the reference's arithmetic order is deterministic (model
builder, K*u, masks, diagonals, transfers, filter); within round-off where the
reference reduces with OpenMP reductions / omp simd (compliance, Cholesky,
PCG, V-cycle through the coarsest factor).  Skipped where oracle/_ref is not
built (it needs the reference tree at build time); the committed golden
fixtures (tests/golden) pin the same quantities without it.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (reference tree absent)")

os.environ.setdefault("OMP_NUM_THREADS", "1")
SHAPES = [(2, 4, 2), (8, 16, 8), (30, 60, 30)]


@pytest.mark.parametrize("dims", SHAPES)
def test_model_kxu_mg_bitexact(dims):
    m = O.Model(*dims)
    r = O.RefModel(*dims)
    for f in ("eleMapBack", "eNodMat", "color", "Ke", "F", "isFree", "freeDofIndex", "eDofFree"):
        assert np.array_equal(getattr(m, f), getattr(r, f)), f
    assert np.array_equal(m.full_to_free(m.b), r.bFree)
    rng = np.random.default_rng(0)
    E = rng.uniform(1e-3, 1, m.ne)
    E[::7] = 1e-17
    u = rng.uniform(-1, 1, m.nF)
    assert np.array_equal(m.kxu_free(E, u), r.kxu(E, u))
    C1, ce1 = m.compliance(E, m.free_to_full(u))
    C2, ce2 = r.compliance(E, u)
    assert abs(C1 - C2) <= 1e-13 * abs(C2)
    assert np.abs(ce1 - ce2).max() <= 1e-13 * np.abs(ce2).max()
    _check_mg(m, r, O.MG(m), E, rng)


def _check_mg(m, r, g, E, rng):
    """levels, masks, diagonals, transfers, coarsest matrix and factor, one V-cycle"""
    assert [tuple(d) for d in g.dims] == r.levels()
    for l in range(g.nlev):
        assert np.array_equal(g.masks[l], r.mask(l, g.nN[l]))
    g.setup(E)
    for l in range(g.nlev):
        assert np.array_equal(g.diag[l], r.diag(E, l, g.nN[l])), l
    for lc in range(1, g.nlev):
        rf = rng.uniform(-1, 1, 3 * g.nN[lc - 1])
        assert np.array_equal(g.restrict_vec3(lc, rf), r.restrict_vec3(lc, rf, g.nN[lc]))
        xc = rng.uniform(-1, 1, 3 * g.nN[lc])
        assert np.array_equal(g.prolong_vec3(lc, xc), r.prolong_vec3(lc, xc, g.nN[lc - 1]))
    if g.Kc is not None:
        N = g.Kc.shape[0]
        Kc, Lc, ok = r.coarsest(E, 1e-6, N)
        assert ok == g.cholOk
        assert np.abs(g.Kc - Kc).max() <= 1e-14 * np.abs(Kc).max()
        if g.Lc is not None:
            assert np.abs(g.Lc - Lc).max() <= 1e-13 * np.abs(Lc).max()
    rr = rng.uniform(-1, 1, m.nF)
    z1 = m.full_to_free(g.apply(m.free_to_full(rr)))
    z2 = r.mg_apply(E, 1e-6, rr)
    assert np.abs(z1 - z2).max() <= 1e-12 * np.abs(z2).max()


# the other hierarchies of MGPrecondConfig (multigrid.hpp:9-13; hierarchy.cpp:84-138):
# every coarsening by 2 (nonDyadic = false), and more levels
HIERARCHIES = [(0, 5), (0, 6), (1, 6)]


@pytest.mark.parametrize("hier", HIERARCHIES)
@pytest.mark.parametrize("dims", [(8, 16, 8), (16, 32, 16), (12, 20, 8)])
def test_mg_hierarchies_bitexact(dims, hier):
    m = O.Model(*dims)
    r = O.RefModel(*dims, nonDyadic=hier[0], maxLevels=hier[1])
    rng = np.random.default_rng(7)
    E = rng.uniform(1e-3, 1, m.ne)
    E[::11] = 1e-17
    g = O.MG(m, nonDyadic=hier[0], maxLevels=hier[1])
    _check_mg(m, r, g, E, rng)


@pytest.mark.parametrize("dims,tol,maxIt,hier", [((8, 16, 8), 1e-3, 800, (1, 5)), ((8, 16, 8), 1e-8, 300, (1, 5)),
                                                 ((30, 60, 30), 1e-3, 60, (1, 5)),
                                                 ((16, 32, 16), 1e-8, 300, (0, 6))])
def test_pcg_iterations_and_solution(dims, tol, maxIt, hier):
    m = O.Model(*dims)
    r = O.RefModel(*dims, nonDyadic=hier[0], maxLevels=hier[1])
    rng = np.random.default_rng(1)
    E = rng.uniform(1e-3, 1, m.ne)
    g = O.MG(m, nonDyadic=hier[0], maxLevels=hier[1])
    g.setup(E)
    it1, x1 = g.pcg(E, m.b, np.zeros(m.nDOF), tol, maxIt)
    it2, x2 = r.pcg(E, 1e-6, r.bFree, np.zeros(m.nF), tol, maxIt)
    assert it1 == it2
    x1 = m.full_to_free(x1)
    assert np.abs(x1 - x2).max() <= 1e-10 * np.abs(x2).max()


@pytest.mark.parametrize("dims,rmin", [((8, 16, 8), 3 ** 0.5), ((30, 60, 30), 1.5)])
def test_filter_bitexact(dims, rmin):
    m = O.Model(*dims)
    r = O.RefModel(*dims)
    f = O.Filter(m, rmin)
    k, d = r.filter_setup(rmin)
    assert np.array_equal(f.kernel, k) and np.array_equal(f.diag, d)
    rng = np.random.default_rng(2)
    base = rng.uniform(-1, 0, m.ne).astype(np.float32)
    for call in range(3):
        src = (base * (1 + 0.1 * call)).astype(np.float32)
        a, _, _ = f.apply(src)
        b = r.filter_apply(src)
        assert np.abs(a - b).max() <= 2e-7 * np.abs(b).max()


def test_optimisation_loop_vs_reference():
    """driver.cpp:112-218 with the reference's functions vs the oracle loop.

    The first loop agrees to round-off (1e-12); after that the loop is chaotic
    in the last bits: the reference run against itself with 1, 2 and 8 OpenMP
    threads differs by ~6e-10 relative in loop 2 and ~3e-6 in loop 5, with
    density differences up to 2e-5 (measured on this 16x8x8 case).  The oracle
    stays inside that spread: PCG iteration counts identical, compliance within
    1e-9 in loop 2 and 1e-5 afterwards, densities within 1e-4."""
    dims = (8, 16, 8)
    kw = dict(V0=0.12, nLoop=5, rMin=1.5, Emin=1e-9, cgTol=1e-3, cgMaxIt=200)
    ro, xo, Cso, _ = O.run(*dims, **kw)
    rr, xr, Csr = O.ref_run(*dims, **kw)
    assert abs(Cso - Csr) <= 1e-12 * abs(Csr)
    assert len(ro) == len(rr)
    for k, (a, b) in enumerate(zip(ro, rr)):
        assert a["pcgIters"] == b["pcgIters"]
        bound = 1e-12 if k == 0 else (1e-9 if k == 1 else 1e-5)
        assert abs(a["Cdisp"] - b["Cdisp"]) <= bound * abs(b["Cdisp"]), (k, a["Cdisp"], b["Cdisp"])
    assert np.abs(xo - xr).max() <= 1e-4


def test_cli_known_answers(tmp_path):
    """The reference CLI (strict build) on a small cuboid, default parameters:
    the solid compliance and the loop-1 objective (printed to 9 significant
    digits) agree with the oracle to print precision; later objectives are
    chaotic in the reference itself, so they are compared against the spread of
    the reference's own runs with 1, 2 and 4 OpenMP threads, widened to the
    1e-5 relative spread that the reference shows across thread counts and
    repeated runs by loop 4 (three runs can happen to agree more closely)."""
    import re
    import subprocess
    cli = os.path.join(O.HERE, "_ref", "top3d_xl_cli")
    if not os.path.exists(cli):
        pytest.skip("reference CLI not built")
    runs = []
    for t in ("1", "2", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        out = subprocess.run([cli, "GLOBAL", "8", "16", "8", "0.12", "4", "1"], cwd=tmp_path, env=env,
                             capture_output=True, text=True, timeout=600).stdout
        objs = [float(v) for v in re.findall(r"It\.:\s*\d+ Obj\.:\s*([-+0-9.eE]+)", out)]
        solid = float(re.search(r"Compliance of Fully Solid Domain:\s*([-+0-9.eE]+)", out).group(1))
        runs.append((solid, objs))
    recs, _, Cs, _ = O.run(8, 16, 8, nLoop=4)
    solid, objs = runs[0]
    assert abs(Cs - solid) <= 1e-6 * abs(solid)
    assert len(objs) == len(recs)
    assert abs(recs[0]["Cdisp"] - objs[0]) <= 1e-8 * abs(objs[0])
    for k in range(1, len(objs)):
        vals = [r[1][k] for r in runs]
        spread = max(2 * (max(vals) - min(vals)), 1e-5 * abs(vals[0]))
        assert min(vals) - spread <= recs[k]["Cdisp"] <= max(vals) + spread, (k, recs[k]["Cdisp"], vals)
