"""GPU parity of the rest of the inner loop vs the oracle, bit-exact:
ComputeCompliance (cpp/src/fea.cpp:587-677), the PDE filter (cpp/src/filter.cpp:9-275),
SIMP / OC / sequential float sums (cpp/src/driver.cpp:117-201) and the whole
TOP3D_XL_GLOBAL loop (cpp/src/driver.cpp:23-218)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(2, 4, 2), (8, 16, 8), (30, 60, 30)])
def test_compliance_bitexact(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    rng = np.random.default_rng(1)
    E = rng.uniform(1e-3, 1, m.ne)
    E[::9] = 1e-17
    U = m.free_to_full(rng.uniform(-1, 1, m.nF))
    C_o, ce_o = m.compliance(E, U)
    dce = t3d.DeviceArray(m.ne)
    C_g = t3d.compliance(p, t3d.DeviceArray.from_host(E), t3d.DeviceArray.from_host(U), dce)
    assert C_g == C_o
    assert np.array_equal(dce.download(), ce_o)


@pytest.mark.parametrize("dims,rmin", [((2, 4, 2), 1.5), ((8, 16, 8), 3 ** 0.5), ((30, 60, 30), 2.0)])
def test_filter_bitexact_with_warm_start(t3d, dims, rmin):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    f_o = O.Filter(m, rmin)
    f_g = t3d.Filter(p, rmin)
    k, d = f_g.info()
    assert np.array_equal(k, f_o.kernel) and np.array_equal(d, f_o.diag)
    rng = np.random.default_rng(2)
    dst = t3d.DeviceArray(m.ne, np.float32)
    base = rng.uniform(-1, 0, m.ne).astype(np.float32)
    for call in range(4):
        # successive calls with slowly changing input exercise the warm start
        src = (base * (1 + 0.05 * call)).astype(np.float32)
        if call == 3:
            src = np.zeros(m.ne, np.float32)  # ||rhs|| = 0: cold start, early exit
        out_o, it_o, warm_o = f_o.apply(src)
        it_g = f_g.apply(t3d.DeviceArray.from_host(src), dst)
        assert it_g == it_o
        assert np.array_equal(dst.download(), out_o)


def _adversarial_sums(rng, n):
    yield rng.uniform(1e-3, 1, n).astype(np.float32)
    yield np.full(n, 0.12, np.float32)
    yield np.full(n, 1e-3, np.float32)
    yield (rng.integers(0, 1025, n) / 1024.0).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    yield (t * (np.float32(1) - t)).astype(np.float32)
    v = rng.uniform(1e-3, 1, n).astype(np.float32)
    v[rng.integers(0, 3, n) == 0] = 0
    yield v


@pytest.mark.parametrize("n", [1, 1000, 9000, 100003, 2_000_000])
def test_seq_sum_exact(t3d, n):
    p = t3d.Problem(2, 4, 2)
    rng = np.random.default_rng(n)
    for v in _adversarial_sums(rng, n):
        got = t3d.seq_sum(p, t3d.DeviceArray.from_host(v))
        assert got == O.lib().o_seq_sum(v, n)


def test_fp_semantics_correctly_rounded(t3d):
    rng = np.random.default_rng(9)
    a = np.concatenate([rng.uniform(0, 1, 200000), np.exp(rng.uniform(-80, 10, 200000)),
                        rng.uniform(1e-30, 1e-20, 1000)]).astype(np.float32)
    b = np.concatenate([rng.uniform(1e-6, 1e3, a.size - 1000), np.ones(1000)]).astype(np.float32)
    assert np.array_equal(t3d.fp_probe(0, a), np.sqrt(a))
    assert np.array_equal(t3d.fp_probe(1, a, b), (a / b).astype(np.float32))
    d = np.exp(rng.uniform(-700, 700, 400000))
    assert np.array_equal(t3d.fp_probe(2, d), np.sqrt(d))


@pytest.mark.parametrize("dims", [(30, 60, 30), (96, 192, 128)])
@pytest.mark.parametrize("xmode", ["random", "uniform0.12"])
def test_simp_and_oc_bitexact(t3d, xmode, dims):
    """54 000 elements (at most 27 chunks of 2048: one 64-lane batch of
    chunk aggregates) and 2.36M elements (1152 chunks: grid-stride aggregation,
    carries over more than 64 chunks, repeated window doublings)."""
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    rng = np.random.default_rng(4)
    x = rng.uniform(1e-3, 1, m.ne).astype(np.float32)
    x[::13] = 1.0
    if xmode == "uniform0.12":
        x[:] = np.float32(0.12)
    E_o = np.zeros(m.ne)
    O.lib().o_simp(m.ne, x, 1.0, 1e-9, E_o)
    dE = t3d.DeviceArray(m.ne)
    t3d.simp(p, t3d.DeviceArray.from_host(x), 1.0, 1e-9, dE)
    assert np.array_equal(dE.download(), E_o)
    dc = -rng.uniform(0, 5, m.ne).astype(np.float32)
    xnew_o = np.zeros(m.ne, np.float32)
    import ctypes as C
    ch = C.c_float()
    steps_o = O.lib().o_oc_update(m.ne, x, dc, 0.12, 0.2, xnew_o, C.byref(ch))
    dxn = t3d.DeviceArray(m.ne, np.float32)
    r = t3d.oc_update(p, t3d.DeviceArray.from_host(x), t3d.DeviceArray.from_host(dc), 0.12, 0.2, dxn)
    assert r["steps"] == steps_o
    assert np.array_equal(dxn.download(), xnew_o)
    assert r["change"].tobytes() == np.float32(ch.value).tobytes()
    assert r["sharpness"] == np.float32(4.0) * O.lib().o_seq_sum_x1mx(xnew_o, m.ne) / np.float32(m.ne)
    assert r["volFrac"] == O.lib().o_seq_sum(xnew_o, m.ne) / np.float32(m.ne)


@pytest.mark.parametrize("dims,kw", [
    ((8, 16, 8), dict(nLoop=6, cgMaxIt=60, Emin=1e-9, rMin=1.5)),
    ((30, 60, 30), dict(nLoop=3, cgMaxIt=40, cgTol=1e-8, Emin=1e-9, rMin=1.5)),
    ((12, 20, 8), dict(nLoop=5, cgMaxIt=800)),
])
def test_top3d_loop_bitexact(t3d, dims, kw):
    rec_o, x_o, Cs_o, si_o = O.run(*dims, **kw)
    rec_g, x_g, Cs_g, si_g = t3d.top3d_run(*dims, **kw)
    assert Cs_g == Cs_o and si_g == si_o
    assert len(rec_g) == len(rec_o)
    for a, b in zip(rec_o, rec_g):
        for key in ("pcgIters", "filterIters", "C", "Cdisp", "volFrac", "sharpness", "change", "fval"):
            assert np.float64(a[key]) == np.float64(b[key]), (a["loop"], key, a[key], b[key])
    assert np.array_equal(x_g, x_o)


@pytest.mark.parametrize("dims", [(30, 60, 30), (96, 192, 128)])
def test_oc_passive_bitexact(t3d, dims):
    """OC with passive elements held at 1 (matlab/TOP3D_XL.m:154) against the
    oracle; 54 000 and 2.36M elements, as test_simp_and_oc_bitexact."""
    import ctypes as C
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    rng = np.random.default_rng(12)
    x = rng.uniform(0.01, 0.6, m.ne).astype(np.float32)
    dc = -rng.uniform(0, 5, m.ne).astype(np.float32)
    passive = (rng.random(m.ne) < 0.05).astype(np.uint8)
    xnew_o = np.zeros(m.ne, np.float32)
    ch = C.c_float()
    steps_o = O.lib().o_oc_update_passive(m.ne, x, dc, passive.ctypes.data, 0.3, 0.2, xnew_o, C.byref(ch))
    dxn = t3d.DeviceArray(m.ne, np.float32)
    r = t3d.oc_update(p, t3d.DeviceArray.from_host(x), t3d.DeviceArray.from_host(dc), 0.3, 0.2, dxn,
                      passive=t3d.DeviceArray.from_host(passive))
    assert r["steps"] == steps_o
    assert np.array_equal(dxn.download(), xnew_o)
    assert r["change"].tobytes() == np.float32(ch.value).tobytes()


def test_loop_voxel_domain_passive_bitexact(t3d):
    """The loop on a voxel domain with voids and a passive layer at the loaded
    end (t3d_params.solid / passive) against the oracle's loop, record by record."""
    from test_voxels import spheres_mask
    dims = (16, 32, 16)
    solid = spheres_mask(*dims, [(10.0, 8.0, 8.0, 4.5), (22.0, 4.0, 10.0, 3.0)])
    passive = np.zeros_like(solid)
    passive[:, 31:, :] = 1
    passive &= solid
    kw = dict(nLoop=4, V0=0.2, cgMaxIt=200, Emin=1e-9, rMin=1.5)
    rec_o, x_o, Cs_o, si_o = O.run(*dims, solid=solid, passive=passive, **kw)
    rec_g, x_g, Cs_g, si_g = t3d.top3d_run(*dims, solid=solid, passive=passive, **kw)
    assert Cs_g == Cs_o and si_g == si_o
    assert len(rec_g) == len(rec_o) == 4
    for a, b in zip(rec_o, rec_g):
        for key in ("pcgIters", "filterIters", "C", "Cdisp", "volFrac", "sharpness", "change", "fval"):
            assert np.float64(a[key]) == np.float64(b[key]), (a["loop"], key, a[key], b[key])
    assert np.array_equal(x_g, x_o)
