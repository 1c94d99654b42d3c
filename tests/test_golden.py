"""The oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py runs the reference's own sources, strict build,
one thread).  Same tolerances as test_oracle_vs_reference.py; runs without the
reference tree."""
import glob
import os

import numpy as np
import pytest

import oracle as O

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_*.npz")))


@pytest.fixture(params=FIX, ids=[os.path.basename(f) for f in FIX])
def gold(request):
    return dict(np.load(request.param))


def test_model(gold):
    m = O.Model(*[int(v) for v in gold["dims"]])
    for f in ("eleMapBack", "eNodMat", "color", "Ke", "F", "isFree", "freeDofIndex", "eDofFree"):
        assert np.array_equal(getattr(m, f), gold[f]), f
    assert np.array_equal(m.full_to_free(m.b), gold["bFree"])


def test_kxu_compliance(gold):
    m = O.Model(*[int(v) for v in gold["dims"]])
    assert np.array_equal(m.kxu_free(gold["E"], gold["u"]), gold["kxu"])
    C, ce = m.compliance(gold["E"], m.free_to_full(gold["u"]))
    assert abs(C - gold["C"][0]) <= 1e-13 * abs(gold["C"][0])
    assert np.abs(ce - gold["ce"]).max() <= 1e-13 * np.abs(gold["ce"]).max()


def test_multigrid(gold):
    m = O.Model(*[int(v) for v in gold["dims"]])
    g = O.MG(m)
    assert np.array_equal(np.array([list(d) for d in g.dims]), gold["levels"])
    g.setup(gold["E_mg"])
    for l in range(g.nlev):
        assert np.array_equal(g.masks[l], gold["mask%d" % l])
        assert np.array_equal(g.diag[l], gold["diag%d" % l])
    for lc in range(1, g.nlev):
        assert np.array_equal(g.restrict_vec3(lc, gold["restrict_in%d" % lc]), gold["restrict_out%d" % lc])
        assert np.array_equal(g.prolong_vec3(lc, gold["prolong_in%d" % lc]), gold["prolong_out%d" % lc])
    z = m.full_to_free(g.apply(m.free_to_full(gold["vcycle_in"])))
    assert np.abs(z - gold["vcycle_out"]).max() <= 1e-12 * np.abs(gold["vcycle_out"]).max()
    it, x = g.pcg(gold["E_mg"], m.b, np.zeros(m.nDOF), 1e-3, 800)
    assert it == gold["pcg_iters"][0]
    # round-off of the OpenMP-order dots is amplified by the conditioning (1e-7 on the tiny case)
    assert np.abs(m.full_to_free(x) - gold["pcg_x"]).max() <= 1e-6 * np.abs(gold["pcg_x"]).max()


def test_filter(gold):
    m = O.Model(*[int(v) for v in gold["dims"]])
    f = O.Filter(m, 1.5)
    assert np.array_equal(f.kernel, gold["filter_kernel"])
    assert np.array_equal(f.diag, gold["filter_diag"])
    for src, dst in zip(gold["filter_src"], gold["filter_dst"]):
        out, _, _ = f.apply(src)
        assert np.abs(out - dst).max() <= 2e-7 * np.abs(dst).max()


def test_loop(gold):
    dims = [int(v) for v in gold["dims"]]
    recs, x, Cs, _ = O.run(*dims, V0=0.12, nLoop=4, rMin=1.5, Emin=1e-9, cgTol=1e-3, cgMaxIt=200)
    assert abs(Cs - gold["loop_Csolid"][0]) <= 1e-12 * abs(gold["loop_Csolid"][0])
    assert [r["pcgIters"] for r in recs] == list(gold["loop_iters"])
    for k, (r, c) in enumerate(zip(recs, gold["loop_Cdisp"])):
        assert abs(r["Cdisp"] - c) <= (1e-12 if k == 0 else 1e-5) * abs(c)
    assert np.abs(x - gold["loop_x"]).max() <= 1e-4
