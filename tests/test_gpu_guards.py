"""The reference's guard paths on the device, bit-exact against the oracle:

* the coarsest matrix that is not positive definite: with Emin = 0 and void
  elements of modulus 0 the coarsest Galerkin matrix (coarsest.cpp:138-323,
  moduli clamped below by Emin) has zero pivots, chol_spd_inplace
  (coarsest.cpp:25,56-58) fails and the V-cycle falls back to the diagonal
  (preconditioner.cpp:53,76-77) -- both Cholesky forms (N = 135 row form,
  N = 315 column form);
* a zero right-hand side: PCG_free stops after one iteration with x = 0
  (||r|| = 0 < max(1e-30, ||b||) tol, solver.cpp:147-149,184);
* a one-level hierarchy (maxLevels = 1): no coarse level, no coarsest solve
  (preconditioner.cpp:53)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _voids(m, seed):
    E = np.random.default_rng(seed).uniform(1e-3, 1.0, m.ne)
    E[m.ne // 4:] = 0.0
    return E


@pytest.mark.parametrize("dims", [(8, 16, 8), (12, 20, 8)])
def test_non_spd_coarsest_falls_back_to_diagonal(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m, Emin=0.0)
    G = t3d.MG(p, Emin=0.0)
    E = _voids(m, 1)
    g.setup(E)
    dE = t3d.DeviceArray.from_host(E)
    G.setup(dE)
    assert g.cholOk == 0 and g.Nc == 0
    assert G.info() == (0, 0)
    for l in range(g.nlev):
        assert np.array_equal(G.diag(l), g.diag[l]), l
    r = m.free_to_full(np.random.default_rng(5).uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-3, 40)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    assert t3d.pcg(p, G, db, dx, 1e-3, 40) == it_o
    assert np.array_equal(dx.download(), x_o)


@pytest.mark.parametrize("dims", [(8, 16, 8), (30, 60, 30)])
def test_zero_rhs(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m)
    G = t3d.MG(p)
    E = np.random.default_rng(2).uniform(1e-3, 1.0, m.ne)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    it_o, x_o = g.pcg(E, np.zeros(m.nDOF), np.zeros(m.nDOF), 1e-3, 50)
    db = t3d.DeviceArray(m.nDOF)
    db.zero()
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, 1e-3, 50)
    assert it_o == it_g == 1
    assert np.array_equal(dx.download(), x_o) and not np.any(x_o)
    # the mixed solver: its true residual ||b - K x|| = 0 stops it with x = 0
    mx = t3d.Mixed(p, G)
    mx.setup()
    dx.zero()
    mx.pcg(db, dx, 1e-6, 50)
    assert not np.any(dx.download())


@pytest.mark.parametrize("dims", [(8, 16, 8), (30, 60, 30)])
def test_one_level_hierarchy(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m, maxLevels=1)
    G = t3d.MG(p, maxLevels=1)
    assert g.nlev == G.nlev == 1
    E = np.random.default_rng(3).uniform(1e-3, 1.0, m.ne)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    assert G.info()[0] == g.Nc == 0
    assert np.array_equal(G.diag(0), g.diag[0])
    r = m.free_to_full(np.random.default_rng(6).uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-3, 300)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    assert t3d.pcg(p, G, db, dx, 1e-3, 300) == it_o
    assert np.array_equal(dx.download(), x_o)
