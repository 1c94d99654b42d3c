"""t3d_compliance_noalloc (the allocation-free variant of ComputeCompliance,
cpp/src/fea.cpp:587-677,851-864) against t3d_compliance through the C ABI:
ce and the total bit for bit, on two calls in a row (the second on a smaller
problem, so the shared scratch holds the first call's larger partials), then
once more on the first problem with new inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(t3d, dims, seed):
    p = t3d.Problem(*dims)
    rng = np.random.default_rng(seed)
    E = rng.uniform(1e-3, 1, p.ne)
    E[::11] = 1e-17
    U = rng.uniform(-1, 1, p.nDOF)
    return p, t3d.DeviceArray.from_host(E), t3d.DeviceArray.from_host(U)


def _check(t3d, p, dE, dU):
    ce_a = t3d.DeviceArray(p.ne)
    ce_b = t3d.DeviceArray(p.ne)
    C_a = t3d.compliance(p, dE, dU, ce_a)
    C_b = t3d.compliance_noalloc(p, dE, dU, ce_b)
    assert C_b == C_a
    assert np.array_equal(ce_b.download(), ce_a.download())


def test_compliance_noalloc_matches(t3d):
    big = _case(t3d, (48, 96, 48), 1)      # 221 184 elements: 216 tiles, two tree passes
    small = _case(t3d, (8, 16, 8), 2)      # 1 024 elements: one tile
    print("case", "48x96x48", flush=True)
    _check(t3d, *big)
    print("case", "8x16x8", flush=True)
    _check(t3d, *small)
    p, _, _ = big
    rng = np.random.default_rng(3)
    dE = t3d.DeviceArray.from_host(rng.uniform(1e-3, 1, p.ne))
    dU = t3d.DeviceArray.from_host(rng.uniform(-1, 1, p.nDOF))
    print("case", "48x96x48 again", flush=True)
    _check(t3d, p, dE, dU)
