"""Galerkin form of the multigrid preconditioner on the device
(t3d_mg_set_mode(g, 1), the hierarchy of matlab/TOP3D_XL.m:772-1090 on the C++
port's levels) against the scipy restatement in oracle/galerkin.py, which forms
every level operator as the global product P^T A P.

The device builds the same operators cell by cell in a different summation
order, so the comparisons are to FP64 round-off: diagonals and the coarsest
matrix within 1e-12 of their largest entry, the V-cycle within 1e-11.  PCG
with this preconditioner (FP64 plain loop and the mixed-precision loop) is
checked by its true residual and by the compliance of its solution."""
import numpy as np
import pytest

import oracle as O
from galerkin import GalerkinMG
from test_voxels import spheres_mask

pytestmark = pytest.mark.gpu

CASES = [((8, 16, 8), None), ((16, 32, 16), None), ((12, 20, 8), None),
         ((16, 32, 16), [(10.0, 8.0, 8.0, 4.5), (22.0, 4.0, 10.0, 3.0)])]


def _both(t3d, dims, spheres):
    solid = spheres_mask(*dims, spheres) if spheres else None
    return t3d.Problem(*dims, solid=solid), O.Model(*dims, solid=solid)


def _E(m, seed):
    E = np.random.default_rng(seed).uniform(1e-3, 1.0, m.ne)
    E[::13] = 1e-17
    return E


def _setup(t3d, dims, spheres, seed, hier=(1, 5)):
    p, m = _both(t3d, dims, spheres)
    E = _E(m, seed)
    g = O.MG(m, Emin=1e-9, nonDyadic=hier[0], maxLevels=hier[1])
    Go = GalerkinMG(m, g, E)
    G = t3d.MG(p, Emin=1e-9, galerkin=True, nonDyadic=hier[0], maxLevels=hier[1])
    G.setup(t3d.DeviceArray.from_host(E))
    return p, m, E, Go, G


@pytest.mark.parametrize("dims,spheres", CASES)
def test_diagonals_masks_coarsest(t3d, dims, spheres):
    p, m, E, Go, G = _setup(t3d, dims, spheres, 1)
    assert G.nlev == len(Go.A)
    for l in range(G.nlev):
        d = G.diag(l)
        assert np.abs(d - Go.D[l]).max() <= 1e-12 * np.abs(Go.D[l]).max(), l
        assert np.array_equal(G.mask(l).astype(bool), Go.mask[l]), l
    nc, ok = G.info()
    assert nc == Go.Ac.shape[0] and ok == 1
    Kc, L = G.coarsest()
    assert np.abs(Kc - Go.Ac).max() <= 1e-12 * np.abs(Go.Ac).max()
    assert np.abs(L @ L.T - Go.Ac).max() <= 1e-11 * np.abs(Go.Ac).max()


@pytest.mark.parametrize("dims,spheres", CASES)
def test_vcycle(t3d, dims, spheres):
    p, m, E, Go, G = _setup(t3d, dims, spheres, 2)
    rng = np.random.default_rng(5)
    r = m.free_to_full(rng.uniform(-1, 1, m.nF))
    z_o = Go.apply(r)
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    z_g = dz.download()
    assert np.abs(z_g - z_o).max() <= 1e-11 * np.abs(z_o).max()


def test_mode_switch_restores_reference(t3d):
    """Back in mode 0 the object reproduces the reference's V-cycle bit-exactly."""
    dims = (8, 16, 8)
    p, m = _both(t3d, dims, None)
    E = _E(m, 3)
    G = t3d.MG(p, Emin=1e-9, galerkin=True)
    dE = t3d.DeviceArray.from_host(E)
    G.setup(dE)
    t3d.check(t3d.lib().t3d_mg_set_mode(G.h, 0), "t3d_mg_set_mode")
    G.setup(dE)
    g = O.MG(m, Emin=1e-9)
    g.setup(E)
    for l in range(g.nlev):
        assert np.array_equal(G.mask(l), g.masks[l]) and np.array_equal(G.diag(l), g.diag[l]), l
    r = m.free_to_full(np.random.default_rng(6).uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))


@pytest.mark.parametrize("dims,spheres", [((16, 32, 16), None), ((30, 60, 30), None),
                                          ((16, 32, 16), [(10.0, 8.0, 8.0, 4.5), (22.0, 4.0, 10.0, 3.0)])])
def test_pcg_fp64_and_mixed(t3d, dims, spheres):
    """FP64 PCG_free with the Galerkin V-cycle follows the oracle's textbook PCG
    (same iteration count within one, solutions within 1e-9); the mixed solver
    with the same object converges on its FP64 true residual and its compliance
    agrees with the FP64 one within 1e-6 at tol 1e-8."""
    p, m, E, Go, G = _setup(t3d, dims, spheres, 4)
    tol = 1e-8
    it_o, x_o = Go.pcg(m.b, tol, 500)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, tol, 500)
    x_g = dx.download()
    assert abs(it_g - it_o) <= 1 and it_g < 500, (it_g, it_o)
    assert np.linalg.norm(x_g - x_o) <= 1e-9 * np.linalg.norm(x_o) + np.linalg.norm(x_o) * tol * 10
    dE = t3d.DeviceArray.from_host(E)
    dce = t3d.DeviceArray(p.ne)
    C64 = t3d.compliance(p, dE, dx, dce)
    mx = t3d.Mixed(p, G)
    mx.setup()
    dx.zero()
    it_m = mx.pcg(db, dx, tol, 500)
    assert it_m < 500 and it_m <= 1.5 * it_g + 5, (it_m, it_g)
    dy = t3d.DeviceArray(m.nDOF)
    p.kxu_f64(dx, dy)
    res = np.linalg.norm(m.b - dy.download()) / np.linalg.norm(m.b)
    assert res < tol
    Cm = t3d.compliance(p, dE, dx, dce)
    assert abs(Cm - C64) <= 1e-6 * abs(C64)


def test_loop_galerkin_mixed_vs_fp64(t3d):
    """The optimisation loop (TOP3D_XL_GLOBAL) with the Galerkin multigrid
    (t3d_params.mgMode = 1) and converged solves (tol 1e-6): the mixed-precision
    loop follows the FP64 loop (per-loop compliance within 1e-5, the final
    density within 1e-3), and every solve takes far fewer iterations than with
    the C++ port's V-cycle.  The hierarchy is reassembled every loop (its
    operators carry the moduli); reused from an earlier design it would not
    converge within the cap."""
    kw = dict(V0=0.12, nLoop=6, rMin=1.5, Emin=1e-9, cgTol=1e-6, cgMaxIt=800, mgMode=1)
    r64, x64, Cs64, si64 = t3d.top3d_run(30, 60, 30, precision=0, **kw)
    rmx, xmx, Csmx, simx = t3d.top3d_run(30, 60, 30, precision=1, **kw)
    assert len(r64) == len(rmx) == 6
    assert abs(Csmx - Cs64) <= 1e-6 * abs(Cs64)
    rel = [abs(a["Cdisp"] - b["Cdisp"]) / abs(a["Cdisp"]) for a, b in zip(r64, rmx)]
    dx = float(np.abs(x64 - xmx).max())
    its = [a["pcgIters"] for a in r64] + [b["pcgIters"] for b in rmx]
    print("solid", Cs64, Csmx, si64, simx, "rel", ["%.2e" % v for v in rel], "density Linf", dx, "iters", its)
    assert max(rel) <= 1e-5
    assert dx <= 1e-3
    assert max(its) < 200 and si64 < 150
    ref, _, _, siref = t3d.top3d_run(30, 60, 30, precision=0, **dict(kw, mgMode=0, nLoop=1))
    assert siref > 3 * si64


@pytest.mark.parametrize("dims,spheres", CASES)
def test_standard_vcycle(t3d, dims, spheres):
    """The 'Standard' V-cycle (t3d_mg_set_vcycle(g, 1), TOP3D_XL.m:783-803) against
    the scipy restatement: finest K*u and the coarse levels' element-matrix
    operators in the residuals, to FP64 round-off."""
    p, m, E, Go, G = _setup(t3d, dims, spheres, 2)
    t3d.check(t3d.lib().t3d_mg_set_vcycle(G.h, 1), "t3d_mg_set_vcycle")
    r = m.free_to_full(np.random.default_rng(5).uniform(-1, 1, m.nF))
    z_o = Go.apply(r, standard=True)
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.abs(dz.download() - z_o).max() <= 1e-11 * np.abs(z_o).max()


@pytest.mark.parametrize("dims,spheres", CASES)
@pytest.mark.parametrize("standard", [True, False])
def test_mixed_vcycle_fp32(t3d, dims, spheres, standard):
    """The mixed solver's FP32 V-cycle (t3d_mixed_vcycle) of either cycle type
    against the scipy restatement's FP64 cycle on the same residual: within
    2e-5 of the largest entry (FP32 arithmetic on every level but the coarsest,
    whose inverse factor is FP64), masked DOFs exactly zero."""
    p, m, E, Go, G = _setup(t3d, dims, spheres, 3)
    if standard:
        t3d.check(t3d.lib().t3d_mg_set_vcycle(G.h, 1), "t3d_mg_set_vcycle")
    mx = t3d.Mixed(p, G)
    mx.setup()
    r = m.free_to_full(np.random.default_rng(7).uniform(-1, 1, m.nF))
    r32 = r.astype(np.float32)
    z_o = Go.apply(r32.astype(np.float64), standard=standard)
    dz = t3d.DeviceArray(m.nDOF, np.float32)
    mx.vcycle(t3d.DeviceArray.from_host(r32), dz)
    z = dz.download().astype(np.float64)
    err = np.abs(z - z_o).max() / np.abs(z_o).max()
    print(dims, "standard" if standard else "adapted", "rel err", err)
    assert err <= 2e-5
    fixed = m.free_to_full(np.ones(m.nF)) == 0
    assert np.all(z[fixed] == 0.0)


def test_standard_pcg_fp64(t3d):
    """FP64 PCG_free with the 'Standard' V-cycle on uniform SIMP moduli (rho = 0.12)
    follows the oracle's textbook PCG (iteration count within one, fewer than with
    the 'Adapted' cycle); the mixed solver with the same cycle in FP32 reaches the
    tolerance by its true residual in fewer iterations than the 'Adapted' cycle,
    with the compliance of the FP64 solve.  (On the random moduli with 1e-17 voids of the
    cases above the 'Standard' cycle with the Jacobi weight 0.6 is a poor
    preconditioner: the oracle needs 179 iterations against the 'Adapted' 79 on
    16 x 32 x 16, and PCG with it is not a stable comparison.)"""
    dims = (16, 32, 16)
    p, m = _both(t3d, dims, None)
    E = np.full(m.ne, 1e-9 + 0.12 ** 3 * (1 - 1e-9))
    g = O.MG(m, Emin=1e-9)
    Go = GalerkinMG(m, g, E)
    G = t3d.MG(p, Emin=1e-9, galerkin=True)
    G.setup(t3d.DeviceArray.from_host(E))
    t3d.check(t3d.lib().t3d_mg_set_vcycle(G.h, 1), "t3d_mg_set_vcycle")
    tol = 1e-8
    it_o, x_o = Go.pcg(m.b, tol, 500, standard=True)
    it_a, _ = Go.pcg(m.b, tol, 500)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, tol, 500)
    print("standard", it_g, it_o, "adapted", it_a)
    assert abs(it_g - it_o) <= 1 and it_g < it_a
    x_g = dx.download()
    assert np.linalg.norm(x_g - x_o) <= 1e-9 * np.linalg.norm(x_o) + np.linalg.norm(x_o) * tol * 10
    mx = t3d.Mixed(p, G)
    mx.setup()
    dxm = t3d.DeviceArray(m.nDOF)
    dxm.zero()
    it_m = mx.pcg(db, dxm, tol, 500, trueEvery=25)
    x_m = dxm.download()
    dy = t3d.DeviceArray(m.nDOF)
    p.kxu_f64(dxm, dy)
    b = db.download()
    res = np.linalg.norm(b - dy.download()) / np.linalg.norm(b)
    C_g, C_m = float(b @ x_g), float(b @ x_m)
    print("mixed standard", it_m, "true res", res, "compliance", C_g, C_m)
    # at tol 1e-8 the FP32 levels cost iterations (36 against 27 on this grid),
    # still far fewer than the 'Adapted' cycle needs
    assert res < tol and it_m < it_a
    assert abs(C_m - C_g) <= 1e-9 * abs(C_g)


@pytest.mark.parametrize("dims,spheres,hier", [((16, 32, 16), None, (0, 6)), ((24, 48, 24), None, (0, 6)),
                                               ((16, 32, 16), [(10.0, 8.0, 8.0, 4.5)], (0, 6))])
def test_dyadic_hierarchy(t3d, dims, spheres, hier):
    """The Galerkin form on a hierarchy coarsened by 2 at every level
    (nonDyadic = 0, the MATLAB original's nonDyadic_ = 0,
    matlab/TOP3D_XL.m:1110-1121): operators, V-cycle and both solvers as on
    the default hierarchy."""
    p, m, E, Go, G = _setup(t3d, dims, spheres, 9, hier)
    assert G.nlev == len(Go.A) and all(d[3] == 2 for d in G.dims[1:])
    for l in range(G.nlev):
        d = G.diag(l)
        assert np.abs(d - Go.D[l]).max() <= 1e-12 * np.abs(Go.D[l]).max(), l
        assert np.array_equal(G.mask(l).astype(bool), Go.mask[l]), l
    Kc, L = G.coarsest()
    assert np.abs(Kc - Go.Ac).max() <= 1e-12 * np.abs(Go.Ac).max()
    r = m.free_to_full(np.random.default_rng(3).uniform(-1, 1, m.nF))
    z_o = Go.apply(r)
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.abs(dz.download() - z_o).max() <= 1e-11 * np.abs(z_o).max()
    tol = 1e-8
    it_o, x_o = Go.pcg(m.b, tol, 500)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, tol, 500)
    assert abs(it_g - it_o) <= 1 and it_g < 500, (it_g, it_o)
    dE = t3d.DeviceArray.from_host(E)
    dce = t3d.DeviceArray(p.ne)
    C64 = t3d.compliance(p, dE, dx, dce)
    mx = t3d.Mixed(p, G)
    mx.setup()
    dx.zero()
    it_m = mx.pcg(db, dx, tol, 500)
    assert it_m < 500 and it_m <= 1.5 * it_g + 5, (it_m, it_g)
    dy = t3d.DeviceArray(m.nDOF)
    p.kxu_f64(dx, dy)
    assert np.linalg.norm(m.b - dy.download()) / np.linalg.norm(m.b) < tol
    assert abs(t3d.compliance(p, dE, dx, dce) - C64) <= 1e-6 * abs(C64)
