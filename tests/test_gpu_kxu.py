"""GPU parity: K_times_u_finest (cpp/src/fea.cpp:763-837) FP64 on the device vs the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 4, 2), (8, 16, 8), (30, 60, 30), (12, 20, 8)]


@pytest.mark.parametrize("dims", SHAPES)
def test_model_matches_oracle(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    md = p.model()
    for k in ("eleMapBack", "eNodMat", "color", "Ke", "F", "isFree", "freeDofIndex", "eDofFree"):
        assert np.array_equal(md[k], getattr(m, k)), k


@pytest.mark.parametrize("dims", SHAPES)
def test_kxu_f64_bitexact(t3d, dims):
    p = t3d.Problem(*dims)
    m = O.Model(*dims)
    rng = np.random.default_rng(7)
    E = rng.uniform(1e-3, 1.0, m.ne)
    E[::11] = 1e-17  # skipped elements (fea.cpp:743-744)
    u = rng.uniform(-1, 1, m.nF)
    yo = m.kxu_free(E, u)
    yg = p.kxu_free_host(E, u)
    assert np.array_equal(yo, yg), np.abs(yo - yg).max()


def test_dot_canonical(t3d):
    p = t3d.Problem(8, 16, 8)
    rng = np.random.default_rng(3)
    for n in (1, 5, 1024, 1025, 300001, 2 * 1024 * 1024 + 17):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        da, db = t3d.DeviceArray.from_host(a), t3d.DeviceArray.from_host(b)
        assert p.dot(da, db) == O.canon_dot(a, b)
