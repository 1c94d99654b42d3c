"""Galerkin multigrid of the MATLAB original, restated with scipy.sparse.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the device's Galerkin
form of the multigrid preconditioner (t3d_mg_set_mode(g, 1)).  The product
never imports this module.

What it restates (matlab/TOP3D_XL.m):
  * the finest operator with Dirichlet DOFs kept as unit rows, A_0 = M K M + (I - M)
    (Solving_ApplyBConEleStiffMat_B, :1450-1463, assembled);
  * the coarse operators A_{l+1} = P_l^T A_l P_l (Solving_AssembleFEAstencil,
    :945-1090, which forms the same product cell by cell);
  * Jacobi diagonals diag(A_l) (:1060-1087); on the finest level the C++
    port's diagonal (diagonal.cpp:11-63), which differs from diag(A_0) only
    where moduli are <= 1e-16;
  * restriction P^T and trilinear prolongation P (:809-837);
  * the 'Adapted' V-cycle (Solving_Vcycle, :772-807) with the coarsest level
    solved directly (:1089-1090).
The levels and the prolongation are those of the C++ port's hierarchy
(cpp/src/multigrid/hierarchy.cpp:84-138, transfers.cpp:378-445), taken from the
C oracle; `prolongation` is built here from the trilinear formula and checked
against the C oracle's prolongation in tests.  Everything here is written from
the definitions (global sparse products), not from the device's cell-by-cell
algorithm, so it pins that algorithm.  DOFs of a coarse level whose diagonal
is zero (covering only void) are masked: unit diagonal, zero residual.
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp


def finest_operator(model, E, threshold=True, untouched=False):
    """A_0 = M K M + (I - M) for element moduli E (compact order); moduli
    <= 1e-16 contribute nothing (fea.cpp:743-744) unless threshold=False.
    The unit diagonal covers fixed DOFs of nodes of some element (the MATLAB
    original has no other nodes); untouched=True adds it on the nodes of no
    element too (the padded grid's void nodes), which makes the matrix the
    nonsingular operator the solvers iterate on."""
    Ke = model.Ke.reshape(24, 24)
    en = model.eNodMat.reshape(model.ne, 8).astype(np.int64)
    edof = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(model.ne, 24)
    E = np.asarray(E, np.float64)
    Ee = np.where(E > 1e-16, E, 0.0) if threshold else E
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (Ee[:, None, None] * Ke[None]).reshape(model.ne, 576).ravel()
    K = sp.csr_matrix((vals, (rows, cols)), shape=(model.nDOF, model.nDOF))
    M = model.isFree.astype(np.float64)
    touched = np.zeros(model.nN, bool)
    touched[en.ravel()] = True
    unit = (1.0 - M) * (np.repeat(touched, 3) | untouched)
    return (sp.diags(M) @ K @ sp.diags(M) + sp.diags(unit)).tocsr()


def prolongation(fine_dims, coarse_dims):
    """Trilinear prolongation from level dims (resX, resY, resZ, span) of the
    coarse level to the fine level, node-major AoS DOFs (3 n + c)."""
    rc, s = coarse_dims[:3], coarse_dims[3]
    nxf, nyf, nzf = fine_dims[0] + 1, fine_dims[1] + 1, fine_dims[2] + 1
    nxc, nyc = rc[0] + 1, rc[1] + 1
    fz, fy, fx = np.meshgrid(np.arange(nzf), np.arange(nyf), np.arange(nxf), indexing="ij")
    fz, fy, fx = fz.ravel(), fy.ravel(), fx.ravel()

    def split(f, res):
        c = np.minimum(f // s, res - 1)
        return c, (f - c * s) / s

    cx, tx = split(fx, rc[0])
    cy, ty = split(fy, rc[1])
    cz, tz = split(fz, rc[2])
    fn = (fz * nyf + fy) * nxf + fx
    R, Cc, V = [], [], []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (tx if dx else 1 - tx) * (ty if dy else 1 - ty) * (tz if dz else 1 - tz)
                cn = ((cz + dz) * nyc + (cy + dy)) * nxc + (cx + dx)
                for c in range(3):
                    R.append(3 * fn + c)
                    Cc.append(3 * cn + c)
                    V.append(w)
    R, Cc, V = np.concatenate(R), np.concatenate(Cc), np.concatenate(V)
    keep = V != 0
    nc = (rc[0] + 1) * (rc[1] + 1) * (rc[2] + 1)
    return sp.csr_matrix((V[keep], (R[keep], Cc[keep])), shape=(3 * nxf * nyf * nzf, 3 * nc))


class GalerkinMG:
    """The Galerkin hierarchy on the levels of a C-oracle MG object `mg`
    (oracle.MG), for moduli E; weight as the device's (double)(float)w."""

    def __init__(self, model, mg, E, weight=0.6):
        self.model = model
        self.w = float(np.float32(weight))
        self.P = [None] + [prolongation(mg.dims[l - 1], mg.dims[l]) for l in range(1, mg.nlev)]
        self.A = [finest_operator(model, E)]
        self.A0 = finest_operator(model, E, untouched=True)
        for l in range(1, mg.nlev):
            self.A.append((self.P[l].T @ self.A[-1] @ self.P[l]).tocsr())
        self.D, self.mask = [], []
        for l, A in enumerate(self.A):
            d = A.diagonal().copy()
            if l == 0:   # the C++ port's finest Jacobi diagonal (diagonal.cpp:11-63, no modulus threshold)
                m = ~model.isFree.astype(bool)
                d = finest_operator(model, E, threshold=False).diagonal().copy()
                d = np.where(m | ~(d > 0), 1.0, d)
            else:
                m = ~(d > 0)
                d = np.where(m, 1.0, d)
            self.D.append(d)
            self.mask.append(m)
        Ac = self.A[-1].toarray()
        dead = self.mask[-1]
        Ac[dead, :] = 0.0
        Ac[:, dead] = 0.0
        Ac[dead, dead] = 1.0
        self.Ac = Ac
        self.cf = sla.cho_factor(Ac, lower=True)

    def apply(self, r, standard=False):
        """'Adapted' V-cycle (TOP3D_XL.m:772-807); the output is masked on every level.
        standard=True: the 'Standard' V-cycle (:783-803), the residual recomputed with
        the level operator before restriction and after prolongation."""
        if standard:
            return self._apply_standard(r)
        L = len(self.A) - 1
        rs = [np.where(self.mask[0], 0.0, r)]
        for l in range(L):
            rs.append(np.where(self.mask[l + 1], 0.0, self.P[l + 1].T @ rs[l]))
        x = np.where(self.mask[L], 0.0, sla.cho_solve(self.cf, rs[L]))
        for l in range(L - 1, -1, -1):
            t = self.w * rs[l] / self.D[l]
            x = np.where(self.mask[l], 0.0, (t + self.P[l + 1] @ x) + t)
        return x

    def _apply_standard(self, r):
        L = len(self.A) - 1
        m = self.mask
        rs = [np.where(m[0], 0.0, r)]
        xs = []
        for l in range(L):
            x = np.where(m[l], 0.0, self.w * rs[l] / self.D[l])
            d = np.where(m[l], 0.0, rs[l] - self.A[l] @ x)
            xs.append(x)
            rs.append(np.where(m[l + 1], 0.0, self.P[l + 1].T @ d))
        x = np.where(m[L], 0.0, sla.cho_solve(self.cf, rs[L]))
        for l in range(L - 1, -1, -1):
            xl = np.where(m[l], 0.0, xs[l] + self.P[l + 1] @ x)
            x = np.where(m[l], 0.0, xl + self.w * (rs[l] - self.A[l] @ xl) / self.D[l])
        return x

    def pcg(self, b, tol, maxit, standard=False):
        """Textbook PCG on A_0 (with unit rows on every fixed DOF) with this
        preconditioner: (iterations, x)."""
        A = self.A0
        x = np.zeros_like(b)
        r = b.copy()
        z = self.apply(r, standard)
        p = z.copy()
        rz = r @ z
        nb = np.linalg.norm(b)
        for it in range(maxit):
            Ap = A @ p
            a = rz / (p @ Ap)
            x += a * p
            r -= a * Ap
            if np.linalg.norm(r) < tol * nb:
                return it + 1, x
            z = self.apply(r, standard)
            rzn = r @ z
            p = z + rzn / rz * p
            rz = rzn
        return maxit, x
