"""CPU pins of the oracle that do not need the reference build.

Each check is chosen so that a plausible slip in the restatement (a wrong
index, a transposed operand, a dropped term) fails:
  * Ke against an exact rational integration of B^T D B (fea.cpp:117-179);
  * padding / level counts against the logged known answers;
  * K*u against a brute-force sparse assembly of K_ff from Ke;
  * compliance against u^T K u; PCG against a dense solve;
  * transfers against their linear-map invariants; Cholesky against numpy;
  * the PDE filter against a dense solve of its Helmholtz system;
  * correctly rounded cube against exact rational arithmetic;
  * the canonical reduction against its written definition on small inputs.
"""
import ctypes as C
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------------------
# Ke by exact integration of polynomials (independent of any quadrature code)
# ---------------------------------------------------------------------------
def _exact_ke(nu):
    # local node a: corner signs in natural coords (s,t,p) (fea.cpp:142-153 ordering)
    S = [-1, 1, 1, -1, -1, 1, 1, -1]
    T = [-1, -1, 1, 1, -1, -1, 1, 1]
    P = [-1, -1, -1, -1, 1, 1, 1, 1]
    # polynomials as dict {(i,j,k): coeff} in (s,t,p)

    def mul(a, b):
        out = {}
        for ka, va in a.items():
            for kb, vb in b.items():
                k = (ka[0] + kb[0], ka[1] + kb[1], ka[2] + kb[2])
                out[k] = out.get(k, 0) + va * vb
        return out

    def lin(sign, axis):
        e = [0, 0, 0]
        e[axis] = 1
        return {(0, 0, 0): Fr(1), tuple(e): Fr(sign)}

    def integ(poly):
        tot = Fr(0)
        for (i, j, k), v in poly.items():
            if i % 2 or j % 2 or k % 2:
                continue
            tot += v * Fr(2, i + 1) * Fr(2, j + 1) * Fr(2, k + 1)
        return tot

    # physical derivatives: d/dx = 2 d/ds for a unit cell; N_a = (1+Ss)(1+Tt)(1+Pp)/8
    dN = []
    for a in range(8):
        gx = {k: v * Fr(S[a], 8) * 2 for k, v in mul(lin(T[a], 1), lin(P[a], 2)).items()}
        gy = {k: v * Fr(T[a], 8) * 2 for k, v in mul(lin(S[a], 0), lin(P[a], 2)).items()}
        gz = {k: v * Fr(P[a], 8) * 2 for k, v in mul(lin(S[a], 0), lin(T[a], 1)).items()}
        dN.append((gx, gy, gz))
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
    D = [[lam + 2 * mu, lam, lam, 0, 0, 0], [lam, lam + 2 * mu, lam, 0, 0, 0],
         [lam, lam, lam + 2 * mu, 0, 0, 0], [0, 0, 0, mu, 0, 0], [0, 0, 0, 0, mu, 0], [0, 0, 0, 0, 0, mu]]
    B = [[{} for _ in range(24)] for _ in range(6)]
    for a in range(8):
        gx, gy, gz = dN[a]
        c = 3 * a
        B[0][c] = gx; B[1][c + 1] = gy; B[2][c + 2] = gz
        B[3][c + 1] = gz; B[3][c + 2] = gy
        B[4][c] = gz; B[4][c + 2] = gx
        B[5][c] = gy; B[5][c + 1] = gx
    detJ = Fr(1, 8)
    K = np.zeros((24, 24))
    for i in range(24):
        for j in range(i, 24):
            tot = Fr(0)
            for k in range(6):
                if not B[k][i]:
                    continue
                for l in range(6):
                    if D[k][l] == 0 or not B[l][j]:
                        continue
                    tot += D[k][l] * integ(mul(B[k][i], B[l][j]))
            K[i, j] = K[j, i] = float(tot * detJ)
    return K


def test_ke_exact_integration():
    m = O.Model(2, 4, 2)
    K = m.Ke.reshape(24, 24)
    Kx = _exact_ke(Fr(3, 10))
    assert np.abs(K - Kx).max() < 1e-15
    # rigid-body null space: 3 translations + 3 rotations about the cell centre
    # node positions in the element's natural frame (corner signs of fea.cpp:142-153)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    modes = []
    for c in range(3):
        v = np.zeros(24)
        v[c::3] = 1
        modes.append(v)
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        v = np.cross(np.tile(w, (8, 1)), X - 0.5).reshape(-1)
        modes.append(v)
    for v in modes:
        assert np.abs(K @ v).max() < 1e-14
    assert np.linalg.matrix_rank(K, tol=1e-10) == 18


# ---------------------------------------------------------------------------
# padding / hierarchy known answers (logged by the reference runs)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims,padded", [((30, 60, 30), (64, 32, 32)), ((50, 100, 50), (104, 56, 56)),
                                         ((96, 192, 96), (192, 96, 96)), ((192, 384, 192), (384, 192, 192)),
                                         ((2, 4, 2), (8, 8, 8))])
def test_padding_known_answers(dims, padded):
    import ctypes as C
    rx, ry, rz, L = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    nely, nelx, nelz = dims
    O.lib().o_padding(nelx, nely, nelz, C.byref(rx), C.byref(ry), C.byref(rz), C.byref(L))
    assert (rx.value, ry.value, rz.value) == padded


def test_hierarchy_known_answer_50x100x50():
    # out/log/GLOBAL_latest.log:51 -- "[MG] levels=3 coarsest=13x7x7 dofs=2688"
    m = O.Model(50, 100, 50)
    g = O.MG(m)
    assert g.nlev == 3
    rX, rY, rZ, _ = g.dims[-1]
    assert (rX, rY, rZ) == (13, 7, 7)
    assert 3 * g.nN[-1] == 2688


def test_model_invariants():
    m = O.Model(8, 16, 8)
    en = m.eNodMat.reshape(-1, 8)
    assert all(len(set(r)) == 8 for r in en)
    # colours are proper: no node shared inside a colour
    for c in range(8):
        nodes = en[m.color == c].reshape(-1)
        assert len(nodes) == len(set(nodes))
    # cuboid: nF = 3 * nelx * (nely+1) * (nelz+1) (x = 0 face clamped, padding void)
    assert m.nF == 3 * 16 * 9 * 9
    # loads: total -Z load = -(solid loaded nodes)/count
    assert abs(m.b.sum() - (-(9 * 2) / (2 * 9)) * 1.0) < 1e-12 or m.b[2::3].sum() < 0


# ---------------------------------------------------------------------------
# K*u against brute-force assembly
# ---------------------------------------------------------------------------
def _assemble(m, E):
    import scipy.sparse as sp
    K = m.Ke.reshape(24, 24)
    en = m.eNodMat.reshape(-1, 8)
    dofs = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 24)).reshape(-1)
    vals = (E[:, None, None] * K[None]).reshape(-1)
    mask = E > 1e-16
    sel = np.repeat(mask, 576)
    A = sp.csr_matrix((vals[sel], (rows[sel], cols[sel])), shape=(m.nDOF, m.nDOF))
    fd = m.freeDofIndex
    return A[fd][:, fd].tocsr()


@pytest.mark.parametrize("dims", [(2, 4, 2), (8, 16, 8), (6, 10, 4)])
def test_kxu_vs_assembled(dims):
    m = O.Model(*dims)
    rng = np.random.default_rng(0)
    E = rng.uniform(1e-3, 1, m.ne)
    E[::5] = 1e-20
    u = rng.uniform(-1, 1, m.nF)
    A = _assemble(m, E)
    y = m.kxu_free(E, u)
    ref = A @ u
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    # full layout form agrees bit-for-bit with the free form
    assert np.array_equal(m.full_to_free(m.kxu_full(E, m.free_to_full(u))), y)
    # symmetry v^T K u = u^T K v
    v = rng.uniform(-1, 1, m.nF)
    assert abs(v @ m.kxu_free(E, u) - u @ m.kxu_free(E, v)) < 1e-12 * abs(v @ y)


def test_compliance_and_pcg_vs_dense():
    import scipy.sparse.linalg as spla
    m = O.Model(8, 16, 8)
    rng = np.random.default_rng(1)
    E = rng.uniform(0.1, 1, m.ne)
    A = _assemble(m, E)
    bF = m.full_to_free(m.b)
    uF = spla.spsolve(A.tocsc(), bF)
    C, ce = m.compliance(E, m.free_to_full(uF))
    assert abs(C - uF @ (A @ uF)) < 1e-10 * abs(C)
    assert abs(C - bF @ uF) < 1e-9 * abs(C)
    g = O.MG(m)
    g.setup(E)
    # the additive multilevel-Jacobi preconditioner is not symmetric (restriction
    # is a normalised average, not P^T), so PCG stagnates near 1e-5 here
    it, x = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-5, 800)
    assert it < 800
    r = bF - A @ m.full_to_free(x)
    assert np.linalg.norm(r) < 1.1e-5 * np.linalg.norm(bF)
    assert np.abs(m.full_to_free(x) - uF).max() < 1e-3 * np.abs(uF).max()
    # unpreconditioned CG converges to the dense solution
    it, x = g.pcg(E, m.b, np.zeros(m.nDOF), 1e-9, 2000, precond=False)
    assert it < 2000
    assert np.abs(m.full_to_free(x) - uF).max() < 1e-6 * np.abs(uF).max()


def test_transfer_invariants():
    m = O.Model(30, 60, 30)
    g = O.MG(m)
    for lc in range(1, g.nlev):
        ones_f = np.ones(3 * g.nN[lc - 1])
        assert np.allclose(g.restrict_vec3(lc, ones_f), 1.0, rtol=0, atol=1e-15)
        ones_c = np.ones(3 * g.nN[lc])
        assert np.allclose(g.prolong_vec3(lc, ones_c), 1.0, rtol=0, atol=1e-15)
        # linearity
        rng = np.random.default_rng(lc)
        a, b = rng.standard_normal(3 * g.nN[lc]), rng.standard_normal(3 * g.nN[lc])
        assert np.allclose(g.prolong_vec3(lc, a + 2 * b), g.prolong_vec3(lc, a) + 2 * g.prolong_vec3(lc, b),
                           atol=1e-13)
        # prolongation is exact on coarse nodes
        nx, ny, nz, span = g.dims[lc]
        fx, fy, fz, _ = g.dims[lc - 1]
        xf = g.prolong_vec3(lc, a).reshape(fz + 1, fy + 1, fx + 1, 3)
        xc = a.reshape(nz + 1, ny + 1, nx + 1, 3)
        assert np.array_equal(xf[::span, ::span, ::span], xc)


def test_cholesky_vs_numpy():
    rng = np.random.default_rng(2)
    for N in (135, 300):   # row form (< 256) and column form
        Q = rng.standard_normal((N, N))
        A = Q @ Q.T + N * np.eye(N)
        L = np.tril(A).copy().reshape(-1)
        assert O.lib().o_chol_inplace(L, N) == 1
        L = L.reshape(N, N)
        assert np.abs(L - np.linalg.cholesky(A)).max() < 1e-10
        b = rng.standard_normal(N)
        x = np.zeros(N)
        O.lib().o_chol_solve(L.reshape(-1).copy(), b, x, N)
        assert np.abs(A @ x - b).max() < 1e-9


def test_filter_vs_dense():
    m = O.Model(4, 8, 4)
    f = O.Filter(m, 1.5)
    # assemble the Helmholtz node matrix
    en = m.eNodMat.reshape(-1, 8)
    A = np.zeros((m.nN, m.nN))
    Kf = f.kernel.reshape(8, 8)
    for e in range(m.ne):
        A[np.ix_(en[e], en[e])] += Kf
    x = np.random.default_rng(3).standard_normal(m.nN)
    assert np.abs(f.matvec(x) - A @ x).max() < 1e-13
    src = np.random.default_rng(4).uniform(-1, 0, m.ne).astype(np.float32)
    rhs = np.zeros(m.nN)
    for e in range(m.ne):
        rhs[en[e]] += float(src[e]) / 8
    used = A.diagonal() > 0
    sol = np.zeros(m.nN)
    sol[used] = np.linalg.solve(A[np.ix_(used, used)], rhs[used])
    want = np.array([sol[en[e]].sum() / 8 for e in range(m.ne)])
    got, its, warm = f.apply(src)
    # truncated CG with absolute tolerance 1e-3 on ||r||: compare loosely
    assert np.abs(got - want).max() < 5e-3 * np.abs(want).max()


def test_correctly_rounded_cube():
    rng = np.random.default_rng(5)
    L = O.lib()
    xs = np.concatenate([rng.uniform(0, 1, 3000), rng.uniform(1e-3, 2e-3, 500)]).astype(np.float32)
    for v in xs:
        ex = Fr(float(v)) ** 3
        got = np.float32(L.o_cr_cube(float(v)))
        lo = np.nextafter(got, np.float32(0))
        hi = np.nextafter(got, np.float32(2))
        d = abs(Fr(float(got)) - ex)
        assert d <= abs(Fr(float(lo)) - ex) and d <= abs(Fr(float(hi)) - ex)
        sq = np.float32(L.o_cr_square(float(v)))
        exs = Fr(float(v)) ** 2
        ds = abs(Fr(float(sq)) - exs)
        assert ds <= abs(Fr(float(np.nextafter(sq, np.float32(0)))) - exs)
        assert ds <= abs(Fr(float(np.nextafter(sq, np.float32(2)))) - exs)


def test_canonical_reduction_definition():
    rng = np.random.default_rng(6)
    for n in (1, 7, 1024, 1500, 5000):
        v = rng.standard_normal(n)
        # written definition
        def tile(items, t):
            s = [0.0] * 256
            for l in range(256):
                acc = 0.0
                for k in range(4):
                    i = t * 1024 + k * 256 + l
                    if i < len(items):
                        acc = acc + items[i]
                s[l] = acc
            w = 128
            while w >= 1:
                for l in range(w):
                    s[l] = s[l] + s[l + w]
                w //= 2
            return s[0]
        items = list(v)
        while len(items) > 1 or len(items) == 0:
            m = (len(items) + 1023) // 1024
            items = [tile(items, t) for t in range(m)]
        assert items[0] == O.canon_sum(v)
    a, b = rng.standard_normal(3000), rng.standard_normal(3000)
    assert O.canon_dot(a, b) == O.canon_sum(a * b)


def test_sequential_float_sum():
    rng = np.random.default_rng(7)
    v = rng.uniform(1e-3, 1, 20000).astype(np.float32)
    s = np.float32(0)
    for x in v:
        s = np.float32(s + x)
    assert O.lib().o_seq_sum(v, v.size) == s


def _oc(ne, x, dc, V0, passive=None):
    xn = np.zeros(ne, np.float32)
    ch = C.c_float()
    pm = None if passive is None else np.ascontiguousarray(passive, np.uint8)
    steps = O.lib().o_oc_update_passive(ne, x, dc, None if pm is None else pm.ctypes.data, V0, 0.2, xn,
                                        C.byref(ch))
    return xn, steps, ch.value


def test_oc_passive_pins():
    """Passive elements (matlab/TOP3D_XL.m:154,169) in the oracle's OC: no
    passive element reproduces the C++ port's OC bit for bit; passive elements
    end at 1; and since they count as 1 in every trial volume, the other
    elements follow the plain OC of the remaining elements with the volume
    target (V0 ne - |P|) / (ne - |P|), up to the bisection's 1e-6 tolerance
    (a passive element left out of the volume, or counted at its old density,
    moves them by far more)."""
    rng = np.random.default_rng(11)
    ne = 5000
    x = rng.uniform(0.01, 0.6, ne).astype(np.float32)   # targets within the move limits
    dc = -rng.uniform(0, 5, ne).astype(np.float32)
    V0 = np.float32(0.3)
    xa, sa, ca = _oc(ne, x, dc, V0)
    xb = np.zeros(ne, np.float32)
    chb = C.c_float()
    sb = O.lib().o_oc_update(ne, x, dc, V0, 0.2, xb, C.byref(chb))
    assert np.array_equal(xa, xb) and sa == sb and ca == chb.value
    assert np.array_equal(_oc(ne, x, dc, V0, np.zeros(ne, np.uint8))[0], xb)
    passive = (rng.random(ne) < 0.1).astype(np.uint8)
    xp, _, _ = _oc(ne, x, dc, V0, passive)
    assert np.all(xp[passive == 1] == 1.0)
    act = passive == 0
    nP = int(passive.sum())
    V0r = np.float32((float(V0) * ne - nP) / (ne - nP))
    xr, _, _ = _oc(int(act.sum()), np.ascontiguousarray(x[act]), np.ascontiguousarray(dc[act]), V0r)
    assert np.abs(xp[act] - xr).max() <= 1e-5
    wrong, _, _ = _oc(int(act.sum()), np.ascontiguousarray(x[act]), np.ascontiguousarray(dc[act]), V0)
    assert np.abs(wrong - xr).max() > 1e-3
    assert np.all(_oc(ne, x, dc, V0, np.ones(ne, np.uint8))[0] == 1.0)


def test_loop_voxels_passive_runs():
    """The oracle's loop on a voxel domain with passive voxels: the passive
    voxels hold density 1 after every update and the volume fraction tracks V0."""
    from test_voxels import spheres_mask
    dims = (8, 16, 8)
    solid = spheres_mask(*dims, [(8.0, 4.0, 4.0, 2.5)])
    passive = np.zeros_like(solid)
    passive[:, 15:, :] = 1          # the layer at the loaded end (ex = 15)
    passive &= solid
    recs, x, Cs, _ = O.run(*dims, solid=solid, passive=passive, nLoop=3, V0=0.3, rMin=1.5, Emin=1e-9)
    m = O.Model(*dims, solid=solid)
    ez, rem = m.eleMapBack // (m.resX * m.resY), m.eleMapBack % (m.resX * m.resY)
    ex, ey = rem // m.resY, m.resY - 1 - rem % m.resY
    pc = passive.reshape(-1)[(ez * dims[1] + ex) * dims[0] + ey].astype(bool)
    assert pc.sum() > 0 and np.all(x[pc] == 1.0)
    assert len(recs) == 3 and all(abs(r["volFrac"] - 0.3) < 2e-3 for r in recs)
