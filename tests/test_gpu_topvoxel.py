"""A TopVoxel model (matlab/TOP3D_XL.m:458-552) on the device, bit-exact
against the oracle given the same domain and lists: the model arrays and load
vector, the multigrid set-up and V-cycle, PCG, and the optimisation loop with
the file's passive elements (TOP3D_XL.m:154) record by record."""
import numpy as np
import pytest

import oracle as O
from test_topvoxel import bracket, write_topvoxel

pytestmark = pytest.mark.gpu


@pytest.fixture()
def model_file(tmp_path):
    solid, passive, density, fix, loads = bracket()
    path = str(tmp_path / "bracket.TopVoxel")
    write_topvoxel(path, solid, passive, density, fix, loads)
    return path


def _oracle_model(vm):
    return O.Model(vm.nely, vm.nelx, vm.nelz, solid=vm.solid, fix=vm.fix, loads=(vm.loadNode, vm.loadF))


def test_model_mg_pcg_bitexact(t3d, model_file):
    vm = t3d.VoxelModel(model_file)
    p = t3d.Problem(vm.nely, vm.nelx, vm.nelz, model=vm)
    m = _oracle_model(vm)
    got = p.model()
    for k, ref in (("eleMapBack", m.eleMapBack), ("eNodMat", m.eNodMat), ("F", m.F), ("isFree", m.isFree),
                   ("freeDofIndex", m.freeDofIndex), ("eDofFree", m.eDofFree)):
        assert np.array_equal(got[k], ref), k
    db = t3d.DeviceArray(m.nDOF)
    p.load_vector(db)
    assert np.array_equal(db.download(), m.b)
    rng = np.random.default_rng(4)
    E = rng.uniform(1e-3, 1.0, m.ne)
    g = O.MG(m)
    G = t3d.MG(p)
    g.setup(E)
    G.setup(t3d.DeviceArray.from_host(E))
    assert G.info() == (g.Nc, g.cholOk)
    r = m.free_to_full(rng.uniform(-1, 1, m.nF))
    dz = t3d.DeviceArray(m.nDOF)
    G.apply(t3d.DeviceArray.from_host(r), dz)
    assert np.array_equal(dz.download(), g.apply(r))
    it_o, x_o = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-3, 200)
    dx = t3d.DeviceArray(m.nDOF)
    dx.zero()
    assert t3d.pcg(p, G, db, dx, 1e-3, 200) == it_o
    assert np.array_equal(dx.download(), x_o)


def test_loop_on_topvoxel_model_bitexact(t3d, model_file):
    vm = t3d.VoxelModel(model_file)
    kw = dict(nLoop=4, V0=0.2, cgMaxIt=200, Emin=1e-9, rMin=1.5)
    rec_g, x_g, Cs_g, si_g = t3d.top3d_run(vm.nely, vm.nelx, vm.nelz, model=vm, **kw)
    rec_o, x_o, Cs_o, si_o = O.run(vm.nely, vm.nelx, vm.nelz, solid=vm.solid, passive=vm.passive, fix=vm.fix,
                                   loads=(vm.loadNode, vm.loadF), **kw)
    assert Cs_g == Cs_o and si_g == si_o
    assert len(rec_g) == len(rec_o) == 4
    for a, b in zip(rec_o, rec_g):
        for key in ("pcgIters", "filterIters", "bisectSteps", "C", "Cdisp", "volFrac", "sharpness", "change",
                    "fval", "xCheck"):
            assert a[key] == b[key], (a["loop"], key, a[key], b[key])
    assert np.array_equal(x_g, x_o)
    # passive elements of the file stay at density 1
    p = t3d.Problem(vm.nely, vm.nelx, vm.nelz, model=vm)
    emb = p.ele_map_back().astype(np.int64)
    ny, nx = p.resY, p.resX
    ez, rem = emb // (ny * nx), emb % (ny * nx)
    ex, ey = rem // ny, ny - 1 - rem % ny
    pas = vm.passive.reshape(vm.nelz, vm.nelx, vm.nely)[ez, ex, ey].astype(bool)
    assert pas.any() and np.all(x_g[pas] == 1.0)


def test_mixed_galerkin_loop_on_topvoxel_model(t3d, model_file):
    """The chosen forms on the file's model: mixed solves with the Galerkin
    multigrid converge in every loop and follow the FP64 loop."""
    vm = t3d.VoxelModel(model_file)
    kw = dict(nLoop=3, V0=0.2, cgTol=1e-6, cgMaxIt=300, Emin=1e-9, rMin=1.5, mgMode=1)
    rec64, x64, _, _ = t3d.top3d_run(vm.nely, vm.nelx, vm.nelz, model=vm, precision=0, **kw)
    recm, xm, _, _ = t3d.top3d_run(vm.nely, vm.nelx, vm.nelz, model=vm, precision=1, **kw)
    for a, b in zip(rec64, recm):
        assert a["pcgIters"] < 300 and b["pcgIters"] < 300
        assert abs(a["C"] - b["C"]) <= 1e-5 * abs(a["C"])
    assert np.abs(xm - x64).max() <= 1e-3
