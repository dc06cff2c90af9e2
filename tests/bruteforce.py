"""Independent numpy/scipy constructions used to pin the oracle and the GPU path.

Nothing here reuses the oracle's index tables: every operator is rebuilt from
physical node coordinates, hat functions and Gauss quadrature on a dense grid,
then compared with the oracle's compact-index (eNodMat / transferMat / Ks)
restatement of TOP3D_XL.m.

Conventions (TOP3D_XL.m:1629-1711): element (iy,ix,iz) of the padded grid,
column-major with y fastest; node (jy,jx,jz) likewise on the (ny+1,nx+1,nz+1)
grid; the physical y axis points opposite to the row index (diagram at
1635-1648), so node (jy,jx,jz) sits at (x,y,z) = (jx, -jy, jz).
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import scipy.sparse as sp


# ----------------------------------------------------------------- geometry
def level_count_and_padding(shape, control=50000):
    """Number of MG levels and padded dims (rule of TOP3D_XL.m:1651-1662,1085-1089)."""
    nely, nelx, nelz = shape
    nvox = nely * nelx * nelz
    n = 0
    while nvox >= control:
        n += 1
        nvox = int(math.floor(nvox / 8 + 0.5))
    n = max(3, n)
    pw = 2 ** n
    pad = tuple(int(math.ceil(d / pw) * pw) for d in (nely, nelx, nelz))
    return n, pad   # n levels after the non-dyadic merge; spans 4,2,2,...


def padded_mask(vox):
    _, (ay, ax, az) = level_count_and_padding(vox.shape)
    m = np.zeros((ay, ax, az), dtype=bool)
    m[: vox.shape[0], : vox.shape[1], : vox.shape[2]] = vox
    return m


def coarsen_mask(mask, s):
    ny, nx, nz = (d // s for d in mask.shape)
    return mask.reshape(ny, s, nx, s, nz, s).any(axis=(1, 3, 5))


def node_mask(emask):
    ny, nx, nz = emask.shape
    nm = np.zeros((ny + 1, nx + 1, nz + 1), dtype=bool)
    for dy, dx, dz in itertools.product((0, 1), repeat=3):
        nm[dy:dy + ny, dx:dx + nx, dz:dz + nz] |= emask
    return nm


def flat_f(a):
    """Column-major flatten (y fastest)."""
    return a.reshape(-1, order="F")


def compact_nodes(emask):
    """Dense ids (column-major) of solid nodes, ascending."""
    return np.nonzero(flat_f(node_mask(emask)))[0]


def compact_elems(emask):
    return np.nonzero(flat_f(emask))[0]


# ------------------------------------------------------- element matrices
def D_iso(E=1.0, nu=0.3):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    return D


def hex8_gauss_stiffness(coords, E=1.0, nu=0.3):
    """Ke of a trilinear hex with the given 8 node coordinates (order as given).

    Built from the isoparametric map with 2x2x2 Gauss quadrature (exact for a
    brick).  Natural coordinates of node a are recovered from the box."""
    coords = np.asarray(coords, float)
    lo, hi = coords.min(0), coords.max(0)
    nat = np.where(np.isclose(coords, lo), -1.0, 1.0)  # natural coords of each node
    D = D_iso(E, nu)
    g = 1 / math.sqrt(3)
    K = np.zeros((24, 24))
    for xi, eta, zeta in itertools.product((-g, g), repeat=3):
        dN = np.zeros((8, 3))
        for a in range(8):
            sa, ta, pa = nat[a]
            dN[a, 0] = 0.125 * sa * (1 + ta * eta) * (1 + pa * zeta)
            dN[a, 1] = 0.125 * ta * (1 + sa * xi) * (1 + pa * zeta)
            dN[a, 2] = 0.125 * pa * (1 + sa * xi) * (1 + ta * eta)
        J = dN.T @ coords            # 3x3: d(x)/d(nat)
        dNx = dN @ np.linalg.inv(J).T  # d N / d x
        B = np.zeros((6, 24))
        for a in range(8):
            bx, by, bz = dNx[a]
            B[0, 3 * a] = bx
            B[1, 3 * a + 1] = by
            B[2, 3 * a + 2] = bz
            B[3, 3 * a + 1], B[3, 3 * a + 2] = bz, by
            B[4, 3 * a], B[4, 3 * a + 2] = bz, bx
            B[5, 3 * a], B[5, 3 * a + 1] = by, bx
        K += B.T @ D @ B * abs(np.linalg.det(J))
    return K


# Local hex8 node order of the reference as (dy, dx, dz) offsets of the node
# grid relative to element (iy,ix,iz): node 1 is one row down (1697).
REF_LOCAL = [(1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 0), (1, 0, 1), (1, 1, 1), (0, 1, 1), (0, 0, 1)]


def ref_ke_gauss(nu=0.3):
    """Gauss-quadrature Ke in the reference's local node order, from physical
    coordinates (x, y, z) = (dx, -dy, dz)."""
    coords = [(dx, -dy, dz) for (dy, dx, dz) in REF_LOCAL]
    return hex8_gauss_stiffness(coords, 1.0, nu)


# -------------------------------------------------------- global operators
class Grid:
    """Dense-grid brute force of one voxel model (finest level + hierarchy)."""

    def __init__(self, vox, fixed_dense=None, fixed_state=None, load_dense=None, load_vec=None):
        self.vox = np.asarray(vox, bool)
        self.nlev, _ = level_count_and_padding(self.vox.shape)
        self.emask = [padded_mask(self.vox)]
        spans = [0] + [4] + [2] * (self.nlev - 2)
        self.spans = spans
        for l in range(1, self.nlev):
            self.emask.append(coarsen_mask(self.emask[-1], spans[l]))
        self.nodes = [compact_nodes(m) for m in self.emask]
        self.elems = [compact_elems(m) for m in self.emask]
        self.nN = [len(n) for n in self.nodes]
        self.Ke = ref_ke_gauss()
        ny, nx, nz = self.emask[0].shape
        self.dims = (ny, nx, nz)
        NY, NX = ny + 1, nx + 1
        self.node_index = -np.ones((ny + 1) * (nx + 1) * (nz + 1), dtype=np.int64)
        self.node_index[self.nodes[0]] = np.arange(self.nN[0])
        # BCs
        nely, nelx, nelz = self.vox.shape
        if fixed_dense is None:   # built-in cuboid, 536-557
            fd = [jy + NY * (0 + NX * jz) for jz in range(nelz + 1) for jy in range(nely + 1)]
            fixed_state = np.ones((len(fd), 3), bool)
            ztop = int(math.floor(nelz / 6 + 1 + 0.5))
            ld = [jy + NY * (nelx + NX * jz) for jz in range(ztop) for jy in range(nely + 1)]
            load_vec = np.tile(np.array([0.0, 0.0, -1.0]) / len(ld), (len(ld), 1))
            fixed_dense, load_dense = fd, ld
        self.fixed_dense = np.asarray(fixed_dense)
        self.fixed_state = np.asarray(fixed_state, bool).reshape(-1, 3)
        self.load_dense = np.asarray(load_dense)
        self.load_vec = np.asarray(load_vec, float).reshape(-1, 3)
        # fixed nodes that are loaded are released (1303-1304)
        keep = ~np.isin(self.fixed_dense, self.load_dense)
        self.fixed_dense, self.fixed_state = self.fixed_dense[keep], self.fixed_state[keep]
        n = 3 * self.nN[0]
        self.fixed = np.zeros(n, bool)
        for d, st in zip(self.fixed_dense, self.fixed_state):
            k = self.node_index[d]
            for c in range(3):
                if st[c]:
                    self.fixed[3 * k + c] = True
        self.F = np.zeros(n)
        for d, v in zip(self.load_dense, self.load_vec):
            k = self.node_index[d]
            self.F[3 * k: 3 * k + 3] = v

    # dense element -> 8 dense node ids in reference order
    def elem_nodes(self, level, dense_e):
        ny, nx, nz = self.emask[level].shape
        NY, NX = ny + 1, nx + 1
        iy, ix, iz = dense_e % ny, (dense_e // ny) % nx, dense_e // (ny * nx)
        return [(iy + dy) + NY * ((ix + dx) + NX * (iz + dz)) for (dy, dx, dz) in REF_LOCAL]

    def K(self, E):
        """Global (unconstrained) stiffness on compact solid nodes."""
        rows, cols, vals = [], [], []
        for k, de in enumerate(self.elems[0]):
            nodes = self.node_index[self.elem_nodes(0, de)]
            dofs = np.repeat(3 * nodes, 3) + np.tile(np.arange(3), 8)
            rows.append(np.repeat(dofs, 24))
            cols.append(np.tile(dofs, 24))
            vals.append((self.Ke * E[k]).reshape(-1))
        n = 3 * self.nN[0]
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))

    def K_bc(self, E):
        """P_free K P_free + I_fixed: the operator of 703-715 + 1338-1437."""
        K = self.K(E)
        free = sp.diags((~self.fixed).astype(float))
        return (free @ K @ free + sp.diags(self.fixed.astype(float))).tocsr()

    def P(self, level):
        """Trilinear prolongation from level (coarse) to level-1 (fine) on solid nodes."""
        s = self.spans[level]
        fy, fx, fz = self.emask[level - 1].shape
        cy, cx, cz = self.emask[level].shape
        FNY, FNX = fy + 1, fx + 1
        CNY, CNX = cy + 1, cx + 1
        cidx = -np.ones(CNY * CNX * (cz + 1), dtype=np.int64)
        cidx[self.nodes[level]] = np.arange(self.nN[level])
        rows, cols, vals = [], [], []
        for i, d in enumerate(self.nodes[level - 1]):
            jy, jx, jz = d % FNY, (d // FNY) % FNX, d // (FNY * FNX)
            ty, tx, tz = jy / s, jx / s, jz / s
            for Jy in {math.floor(ty), math.ceil(ty)}:
                for Jx in {math.floor(tx), math.ceil(tx)}:
                    for Jz in {math.floor(tz), math.ceil(tz)}:
                        w = (1 - abs(ty - Jy)) * (1 - abs(tx - Jx)) * (1 - abs(tz - Jz))
                        if w == 0:
                            continue
                        J = cidx[Jy + CNY * (Jx + CNX * Jz)]
                        assert J >= 0, "hat function reaches a void coarse node"
                        rows.append(i)
                        cols.append(J)
                        vals.append(w)
        Pn = sp.csr_matrix((vals, (rows, cols)), shape=(self.nN[level - 1], self.nN[level]))
        return sp.kron(Pn, sp.eye(3)).tocsr()

    def hierarchy(self, E):
        """Galerkin operators A_l, prolongations P_l, diagonals D_l."""
        A = [self.K_bc(E)]
        P = [None]
        for l in range(1, self.nlev):
            P.append(self.P(l))
            A.append((P[l].T @ A[l - 1] @ P[l]).tocsr())
        return A, P

    def vcycle_matrix(self, E, omega=0.6):
        A, P = self.hierarchy(E)
        L = self.nlev
        Minv = np.linalg.inv(A[L - 1].toarray())
        for l in range(L - 2, -1, -1):
            d = A[l].diagonal()
            Pl = P[l + 1].toarray()
            Minv = np.diag(2 * omega / d) + Pl @ Minv @ Pl.T
        return Minv

    def standard_vcycle_matrix(self, E, omega=0.6):
        """'Standard' V(1,1) operator from its error propagation: with
        S = I - w D^-1 A, M_l = (I - S (I - P M_(l+1) P^T A) S) A^-1."""
        A, P = self.hierarchy(E)
        L = self.nlev
        M = np.linalg.inv(A[L - 1].toarray())
        for l in range(L - 2, -1, -1):
            Al = A[l].toarray()
            n = Al.shape[0]
            S = np.eye(n) - omega * (1.0 / np.diag(Al))[:, None] * Al
            Pl = P[l + 1].toarray()
            Err = S @ (np.eye(n) - Pl @ M @ Pl.T @ Al) @ S
            M = (np.eye(n) - Err) @ np.linalg.inv(Al)
        return M


def pcg_numpy(A, M, b, x0, tol, maxit):
    """Restated PCG (600-639) with dense/sparse operators."""
    nb = np.linalg.norm(b)
    y = x0.copy()
    r = b - A @ y
    z = M(r)
    p = z.copy()
    x1 = z @ r
    its = 0
    for its in range(1, maxit + 1):
        q = A @ p
        lam = x1 / (p @ q)
        y = y + lam * p
        r = r - lam * q
        if np.linalg.norm(r) / nb < tol:
            break
        z = M(r)
        x2 = z @ r
        p = z + x2 / x1 * p
        x1 = x2
    return y, its


# ---------------------------------------------------------------- PDE filter
def laplace_mass_unit_cube():
    """Exact Laplacian stiffness and consistent mass of a unit trilinear brick,
    nodes in the reference order (any orientation: both are reflection invariant)."""
    # 1-D linear element on [0,1]: k = [[1,-1],[-1,1]], m = [[1/3,1/6],[1/6,1/3]]
    k1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
    m1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    Kl = np.zeros((8, 8))
    Mm = np.zeros((8, 8))
    for a, (ay, ax, az) in enumerate(REF_LOCAL):
        for b, (by, bx, bz) in enumerate(REF_LOCAL):
            kx, ky, kz = k1[ax, bx], k1[ay, by], k1[az, bz]
            mx, my, mz = m1[ax, bx], m1[ay, by], m1[az, bz]
            Kl[a, b] = kx * my * mz + mx * ky * mz + mx * my * kz
            Mm[a, b] = mx * my * mz
    return Kl, Mm
