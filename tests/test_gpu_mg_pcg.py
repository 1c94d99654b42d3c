"""GPU parity: multigrid preconditioner (cpp/src/multigrid/*.cpp) and PCG_free
(cpp/src/solver.cpp:87-217) on the device vs the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 4, 2), (8, 16, 8), (30, 60, 30), (12, 20, 8)]


def _E(m, seed):
    rng = np.random.default_rng(seed)
    E = rng.uniform(1e-3, 1.0, m.ne)
    return E


@pytest.mark.parametrize("dims", SHAPES)
def test_mg_static_and_setup(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m)
    G = t3d.MG(p)
    assert G.nlev == g.nlev
    assert [d for d in G.dims] == [tuple(d) for d in g.dims]
    for l in range(g.nlev):
        assert np.array_equal(G.mask(l), g.masks[l])
    E = _E(m, 1)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    for l in range(g.nlev):
        assert np.array_equal(G.diag(l), g.diag[l]), l
    nc, ok = G.info()
    assert nc == g.Nc and ok == g.cholOk
    if g.Kc is not None:
        Kc, L = G.coarsest()
        assert np.array_equal(Kc, g.Kc)
        if g.Nc:
            assert np.array_equal(L, g.Lc)


@pytest.mark.parametrize("dims", SHAPES)
def test_vcycle_bitexact(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m)
    G = t3d.MG(p)
    E = _E(m, 2)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    rng = np.random.default_rng(5)
    r = m.free_to_full(rng.uniform(-1, 1, m.nF))
    z_o = g.apply(r)
    dr = t3d.DeviceArray.from_host(r)
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(dr, dz)
    z_g = dz.download()
    assert np.array_equal(z_o, z_g), np.abs(z_o - z_g).max()


@pytest.mark.parametrize("dims,tol,maxIt", [((8, 16, 8), 1e-3, 800), ((8, 16, 8), 0.0, 60),
                                            ((30, 60, 30), 1e-3, 200), ((12, 20, 8), 1e-8, 800)])
def test_pcg_bitexact(t3d, dims, tol, maxIt):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m)
    G = t3d.MG(p)
    E = _E(m, 3)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), tol, maxIt)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    assert np.array_equal(db.download(), m.b)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, tol, maxIt)
    x_g = dx.download()
    assert it_g == it_o
    assert np.array_equal(x_g, x_o), np.abs(x_g - x_o).max()


def test_pcg_unpreconditioned(t3d):
    dims = (8, 16, 8)
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m)
    E = _E(m, 4)
    p.set_modulus(t3d.DeviceArray.from_host(E))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-6, 300, precond=False)
    db = t3d.DeviceArray.from_host(m.b)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, None, db, dx, 1e-6, 300)
    assert it_g == it_o
    assert np.array_equal(dx.download(), x_o)


# the other hierarchies of MGPrecondConfig (multigrid.hpp:9-13): every
# coarsening by 2 (nonDyadic = false; the finest transfers then run with span 2)
# and more levels; the oracle's levels are pinned against the reference build
# for the same hierarchies (test_oracle_vs_reference.py)
@pytest.mark.parametrize("hier", [(0, 5), (0, 6), (1, 6)])
@pytest.mark.parametrize("dims", [(8, 16, 8), (16, 32, 16), (12, 20, 8)])
def test_hierarchies_bitexact(t3d, dims, hier):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    g = O.MG(m, nonDyadic=hier[0], maxLevels=hier[1])
    G = t3d.MG(p, nonDyadic=hier[0], maxLevels=hier[1])
    assert [d for d in G.dims] == [tuple(d) for d in g.dims]
    for l in range(g.nlev):
        assert np.array_equal(G.mask(l), g.masks[l])
    E = _E(m, 6)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    for l in range(g.nlev):
        assert np.array_equal(G.diag(l), g.diag[l]), l
    nc, ok = G.info()
    assert nc == g.Nc and ok == g.cholOk
    r = m.free_to_full(np.random.default_rng(8).uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-8, 300)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, 1e-8, 300)
    assert it_g == it_o
    assert np.array_equal(dx.download(), x_o)
