"""Writes tests/golden/*.npz: input/output vectors of the reference itself.

Runs the reference's own C++ sources as compiled by oracle/Makefile
(oracle/_ref/libtop3d_ref.so, strict build, one OpenMP thread) on small
cuboids with seeded inputs and stores inputs and outputs.  The fixtures let
tests/test_golden.py pin the oracle where the reference tree is absent.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
os.environ.setdefault("OMP_NUM_THREADS", "1")
import oracle as O  # noqa: E402


def case(dims, seed):
    r = O.RefModel(*dims)
    rng = np.random.default_rng(seed)
    E = rng.uniform(1e-3, 1, r.ne)
    E[::7] = 1e-17
    u = rng.uniform(-1, 1, r.nF)
    out = dict(dims=np.array(dims), E=E, u=u,
               eleMapBack=r.eleMapBack, eNodMat=r.eNodMat, color=r.color, Ke=r.Ke, F=r.F, isFree=r.isFree,
               freeDofIndex=r.freeDofIndex, eDofFree=r.eDofFree, bFree=r.bFree)
    out["kxu"] = r.kxu(E, u)
    C, ce = r.compliance(E, u)
    out["C"] = np.array([C])
    out["ce"] = ce
    levels = r.levels()
    out["levels"] = np.array(levels)
    nN = [(a + 1) * (b + 1) * (c + 1) for a, b, c, _ in levels]
    Eg = rng.uniform(1e-3, 1, r.ne)
    out["E_mg"] = Eg
    for l in range(len(levels)):
        out["mask%d" % l] = r.mask(l, nN[l])
        out["diag%d" % l] = r.diag(Eg, l, nN[l])
    for lc in range(1, len(levels)):
        rf = rng.uniform(-1, 1, 3 * nN[lc - 1])
        xc = rng.uniform(-1, 1, 3 * nN[lc])
        out["restrict_in%d" % lc] = rf
        out["restrict_out%d" % lc] = r.restrict_vec3(lc, rf, nN[lc])
        out["prolong_in%d" % lc] = xc
        out["prolong_out%d" % lc] = r.prolong_vec3(lc, xc, nN[lc - 1])
    rr = rng.uniform(-1, 1, r.nF)
    out["vcycle_in"] = rr
    out["vcycle_out"] = r.mg_apply(Eg, 1e-6, rr)
    it, x = r.pcg(Eg, 1e-6, r.bFree, np.zeros(r.nF), 1e-3, 800)
    out["pcg_iters"] = np.array([it])
    out["pcg_x"] = x
    k, d = r.filter_setup(1.5)
    out["filter_kernel"] = k
    out["filter_diag"] = d
    srcs, dsts = [], []
    base = rng.uniform(-1, 0, r.ne).astype(np.float32)
    for call in range(3):
        src = (base * (1 + 0.1 * call)).astype(np.float32)
        srcs.append(src)
        dsts.append(r.filter_apply(src))
    out["filter_src"] = np.stack(srcs)
    out["filter_dst"] = np.stack(dsts)
    recs, x, Cs = O.ref_run(*dims, V0=0.12, nLoop=4, rMin=1.5, Emin=1e-9, cgTol=1e-3, cgMaxIt=200)
    out["loop_Cdisp"] = np.array([q["Cdisp"] for q in recs])
    out["loop_iters"] = np.array([q["pcgIters"] for q in recs])
    out["loop_Csolid"] = np.array([Cs])
    out["loop_x"] = x
    return out


if __name__ == "__main__":
    if not O.have_ref():
        sys.exit("oracle/_ref/libtop3d_ref.so is not built (make -C oracle ref)")
    for dims, seed in (((2, 4, 2), 10), ((8, 16, 8), 11)):
        d = case(dims, seed)
        path = os.path.join(HERE, "ref_%dx%dx%d.npz" % dims)
        np.savez_compressed(path, **d)
        print(path, os.path.getsize(path))
