"""GPU parity on general voxel domains (t3d_problem_create_voxels: solid mask,
void voxels carry no element): the device model, the FP64 K*u, the multigrid
set-up, the V-cycle and PCG_free against the oracle, bit-exact; the mixed
solver and the batch on masked problems."""
import numpy as np
import pytest

import oracle as O
from test_voxels import MASKS, spheres_mask

pytestmark = pytest.mark.gpu

CASES = MASKS + [((16, 32, 16), [(10.0, 8.0, 8.0, 4.5), (22.0, 4.0, 10.0, 3.0)]),
                 ((24, 48, 24), [(30.0, 12.0, 12.0, 7.0)])]


def _both(t3d, dims, spheres):
    solid = spheres_mask(*dims, spheres)
    return t3d.Problem(*dims, solid=solid), O.Model(*dims, solid=solid), solid


@pytest.mark.parametrize("dims,spheres", CASES)
def test_model_and_kxu_bitexact(t3d, dims, spheres):
    p, m, solid = _both(t3d, dims, spheres)
    assert p.ne == m.ne == int(solid.sum())
    md = p.model()
    for k in ("eleMapBack", "eNodMat", "color", "Ke", "F", "isFree", "freeDofIndex", "eDofFree"):
        assert np.array_equal(md[k], getattr(m, k)), k
    rng = np.random.default_rng(7)
    E = rng.uniform(1e-3, 1.0, m.ne)
    E[::11] = 1e-17
    u = rng.uniform(-1, 1, m.nF)
    yo = m.kxu_free(E, u)
    yg = p.kxu_free_host(E, u)
    assert np.array_equal(yo, yg), np.abs(yo - yg).max()


@pytest.mark.parametrize("dims,spheres", CASES)
def test_mg_vcycle_pcg_bitexact(t3d, dims, spheres):
    p, m, _ = _both(t3d, dims, spheres)
    g = O.MG(m)
    G = t3d.MG(p)
    assert G.nlev == g.nlev
    for l in range(g.nlev):
        assert np.array_equal(G.mask(l), g.masks[l])
    rng = np.random.default_rng(3)
    E = rng.uniform(1e-3, 1.0, m.ne)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    for l in range(g.nlev):
        assert np.array_equal(G.diag(l), g.diag[l]), l
    nc, ok = G.info()
    assert nc == g.Nc and ok == g.cholOk
    r = m.free_to_full(rng.uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-3, 400)
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    assert np.array_equal(db.download(), m.b)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    it_g = t3d.pcg(p, G, db, dx, 1e-3, 400)
    assert it_g == it_o
    assert np.array_equal(dx.download(), x_o)


def test_mixed_and_batch_on_voxel_domains(t3d):
    """Mixed solves on masked problems reach the true residual (1e-3, and 1e-6
    for the first); the batch of the same masked problems with the single
    solver's tiles is bit-identical to solving them one by one."""
    cases = CASES[3:] + [((8, 24, 8), [(12.0, 4.0, 4.0, 2.5)])]
    v0s = [0.15, 0.25, 0.1]
    seq = []
    for (dims, spheres), v0 in zip(cases, v0s):
        solid = spheres_mask(*dims, spheres)
        p = t3d.Problem(*dims, solid=solid)
        dx = t3d.DeviceArray.from_host(np.full(p.ne, v0, np.float32))
        dE = t3d.DeviceArray(p.ne)
        t3d.simp(p, dx, 1.0, 1e-9, dE)
        M = t3d.MG(p, Emin=1e-9)
        M.setup(dE)
        mx = t3d.Mixed(p, M)
        mx.setup()
        db = t3d.DeviceArray(p.nDOF)
        p.load_vector(db)
        x = t3d.DeviceArray(p.nDOF)
        x.zero()
        b = db.download()
        y = t3d.DeviceArray(p.nDOF)
        if not seq:
            x.zero()
            assert mx.pcg(db, x, 1e-6, 4000) < 4000
            p.kxu_f64(x, y)
            assert np.linalg.norm(b - y.download()) <= 1e-6 * np.linalg.norm(b)
        x.zero()
        it = mx.pcg(db, x, 1e-3, 3000)
        assert it < 3000
        p.kxu_f64(x, y)
        assert np.linalg.norm(b - y.download()) <= 1e-3 * np.linalg.norm(b)
        seq.append((it, x.download()))
    dims = [c[0] for c in cases]
    solids = [spheres_mask(*c[0], c[1]) for c in cases]
    B = t3d.Batch(dims, solids=solids)
    dEs = []
    for p, v0 in zip(B.problems, v0s):
        dx = t3d.DeviceArray.from_host(np.full(p.ne, v0, np.float32))
        dE = t3d.DeviceArray(p.ne)
        t3d.simp(p, dx, 1.0, 1e-9, dE)
        dEs.append(dE)
    B.setup(dEs)
    B.set_tiles(16, 4)
    dbs, xs = [], []
    for p in B.problems:
        db = t3d.DeviceArray(p.nDOF)
        p.load_vector(db)
        dbs.append(db)
        x = t3d.DeviceArray(p.nDOF)
        x.zero()
        xs.append(x)
    its = B.solve(dbs, xs, 1e-3, 3000)
    for (it_s, x_s), it_b, x_b in zip(seq, its, xs):
        assert it_b == it_s
        assert np.array_equal(x_b.download(), x_s)
