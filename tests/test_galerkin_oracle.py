"""Pins of the Galerkin-multigrid oracle (oracle/galerkin.py) on the CPU.

* The trilinear prolongation built from the formula equals the C oracle's
  MG_Prolongate_nodes_Vec3_Inner restatement (transfers.cpp:378-445) on every
  level, so the Galerkin levels use the C++ port's own transfer operator.
* A_0 is the masked stiffness matrix with unit Dirichlet rows: A_0 u equals
  the C oracle's K*u on free DOFs and u itself on fixed DOFs.
* Every A_l is symmetric positive definite, and the 'Adapted' V-cycle is a
  symmetric positive definite preconditioner (y^T M(r) = r^T M(y), r^T M(r) > 0).
* PCG with it converges to the direct solution, in far fewer iterations than
  with the C++ port's V-cycle (the reason this form exists).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from galerkin import GalerkinMG, finest_operator, prolongation

CASES = [((8, 16, 8), None), ((16, 32, 16), None), ((12, 20, 8), None), ((16, 24, 16), "sphere")]


def _model(dims, kind):
    if kind is None:
        return O.Model(*dims)
    nely, nelx, nelz = dims
    ez, ex, ey = np.meshgrid(np.arange(nelz), np.arange(nelx), np.arange(nely), indexing="ij")
    solid = ((ex + 0.5 - nelx / 2) ** 2 + (ey + 0.5 - nely / 2) ** 2 + (ez + 0.5 - nelz / 2) ** 2 > 16).astype(np.uint8)
    return O.Model(*dims, solid=solid)


def _E(m, seed):
    return np.random.default_rng(seed).uniform(1e-3, 1.0, m.ne)


@pytest.mark.parametrize("dims,kind", CASES)
def test_prolongation_matches_oracle(dims, kind):
    m = _model(dims, kind)
    g = O.MG(m, Emin=1e-9)
    rng = np.random.default_rng(1)
    for l in range(1, g.nlev):
        P = prolongation(g.dims[l - 1], g.dims[l])
        xc = rng.standard_normal(3 * g.nN[l])
        assert np.allclose(P @ xc, g.prolong_vec3(l, xc), rtol=0, atol=1e-14)


@pytest.mark.parametrize("dims,kind", CASES[:2] + CASES[3:])
def test_finest_operator_is_masked_stiffness(dims, kind):
    m = _model(dims, kind)
    E = _E(m, 2)
    E[::5] = 1e-17   # below the modulus threshold: no contribution
    A = finest_operator(m, E, untouched=True)
    u = np.random.default_rng(3).standard_normal(m.nDOF)
    free = m.isFree.astype(bool)
    Au = A @ u
    uf = np.where(free, u, 0.0)
    assert np.allclose(Au[free], m.kxu_full(E, uf)[free], rtol=0, atol=1e-12 * np.abs(Au).max())
    assert np.array_equal(Au[~free], u[~free])


@pytest.mark.parametrize("dims,kind", CASES)
def test_levels_spd_and_vcycle_symmetric(dims, kind):
    m = _model(dims, kind)
    g = O.MG(m, Emin=1e-9)
    G = GalerkinMG(m, g, _E(m, 4))
    for l, A in enumerate(G.A):
        assert abs(A - A.T).max() <= 1e-13 * abs(A).max(), l
    assert np.linalg.eigvalsh(G.Ac).min() > 0
    rng = np.random.default_rng(5)
    free = m.isFree.astype(bool)
    r = np.where(free, rng.standard_normal(m.nDOF), 0.0)
    y = np.where(free, rng.standard_normal(m.nDOF), 0.0)
    Mr, My = G.apply(r), G.apply(y)
    assert abs(y @ Mr - r @ My) <= 1e-12 * abs(y @ Mr)
    assert r @ Mr > 0


@pytest.mark.parametrize("dims,kind", [((16, 32, 16), None), ((16, 24, 16), "sphere")])
def test_pcg_converges_faster_than_reference_vcycle(dims, kind):
    m = _model(dims, kind)
    g = O.MG(m, Emin=1e-9)
    rho = 0.12
    E = np.full(m.ne, 1e-9 + rho ** 3 * (1 - 1e-9))
    g.setup(E)
    G = GalerkinMG(m, g, E)
    it_g, x = G.pcg(m.b, 1e-8, 400)
    x_direct = spla.spsolve(G.A0.tocsc(), m.b)
    assert np.linalg.norm(x - x_direct) <= 1e-6 * np.linalg.norm(x_direct)
    it_ref, _ = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-8, 2000)
    assert it_g < 150 and it_g * 4 < it_ref, (it_g, it_ref)


@pytest.mark.parametrize("dims,kind", [((16, 32, 16), None), ((16, 24, 16), "sphere")])
def test_standard_vcycle_symmetric_and_faster(dims, kind):
    """The 'Standard' V-cycle (TOP3D_XL.m:783-803): symmetric positive definite
    as a preconditioner (one Jacobi sweep before and one after the coarse
    correction, the same weight), and PCG with it converges to the direct
    solution in fewer iterations than with the 'Adapted' cycle."""
    m = _model(dims, kind)
    g = O.MG(m, Emin=1e-9)
    E = np.full(m.ne, 1e-9 + 0.12 ** 3 * (1 - 1e-9))
    g.setup(E)
    G = GalerkinMG(m, g, E)
    rng = np.random.default_rng(6)
    free = m.isFree.astype(bool)
    r = np.where(free, rng.standard_normal(m.nDOF), 0.0)
    y = np.where(free, rng.standard_normal(m.nDOF), 0.0)
    Mr, My = G.apply(r, True), G.apply(y, True)
    assert abs(y @ Mr - r @ My) <= 1e-11 * abs(y @ Mr)
    assert r @ Mr > 0
    it_s, x = G.pcg(m.b, 1e-8, 400, standard=True)
    it_a, _ = G.pcg(m.b, 1e-8, 400)
    x_direct = spla.spsolve(G.A0.tocsc(), m.b)
    assert np.linalg.norm(x - x_direct) <= 1e-6 * np.linalg.norm(x_direct)
    assert it_s < it_a, (it_s, it_a)
