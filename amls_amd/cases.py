"""Synthetic pencils (K, M) for the AMLS hot path.

Two families live here:

* Graph Laplacians with closed-form spectra.  They follow the reference's
  generators `src/io/case_generator.cpp:15-146` (path graph with lumped or
  consistent mass, 2-D and 3-D grid Laplacians with M = I) and are the
  known-answer inputs of its unit and regression tests.
* 3-D linear elasticity on a regular grid of unit hexahedra (8-node trilinear
  or 27-node triquadratic), isotropic (E, nu), consistent mass, one clamped face
  (z = 0), with per-element stiffness and density scales drawn from a seed.
  The reference has no such generator (its io module only builds Laplacians);
  this one provides the workloads the measurements are quoted on.

All matrices are returned as `scipy.sparse.csc_matrix` with int64 indices and
both triangles stored, i.e. the layout of the reference's
`Eigen::SparseMatrix<double, ColMajor, int64_t>` (`src/core/types.hpp:11-21`).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import scipy.sparse as sp


# --------------------------------------------------------------------------
# Closed-form Laplacian cases (src/io/case_generator.cpp:15-146)
# --------------------------------------------------------------------------

def _csc64(A) -> sp.csc_matrix:
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int64)
    A.indices = A.indices.astype(np.int64)
    return A


def _csc_pattern(R, C, V, n) -> sp.csc_matrix:
    """CSC from unique (row, col) pairs, keeping explicit zeros."""
    order = np.lexsort((R, C))
    R, C, V = R[order], C[order], V[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, C + 1, 1)
    np.cumsum(indptr, out=indptr)
    A = sp.csc_matrix((V, R.astype(np.int64), indptr), shape=(n, n))
    A.has_sorted_indices = True
    return A


def path_graph(n: int, consistent_mass: bool = False):
    """Path-graph Laplacian K = tridiag(-1, 2, -1); M = I or the consistent
    linear-element mass tridiag(1/6, 2/3, 1/6) (`case_generator.cpp:15-40`)."""
    main = np.full(n, 2.0)
    off = np.full(n - 1, -1.0)
    K = sp.diags([off, main, off], [-1, 0, 1], shape=(n, n))
    if consistent_mass:
        M = sp.diags([np.full(n - 1, 1.0 / 6.0), np.full(n, 2.0 / 3.0),
                      np.full(n - 1, 1.0 / 6.0)], [-1, 0, 1], shape=(n, n))
    else:
        M = sp.identity(n)
    return _csc64(K), _csc64(M)


def path_graph_spectrum(n: int, consistent_mass: bool = False) -> np.ndarray:
    """`case_generator.cpp:42-54`: 2 - 2cos(pi j/(n+1)), or
    6(1 - cos t)/(2 + cos t) for the consistent mass."""
    th = math.pi * np.arange(1, n + 1) / (n + 1)
    if consistent_mass:
        w = 6.0 * (1.0 - np.cos(th)) / (2.0 + np.cos(th))
    else:
        w = 2.0 - 2.0 * np.cos(th)
    return np.sort(w)


def grid_graph_2d(nx: int, ny: int):
    """5-point Laplacian on an nx x ny grid, vertex id j*nx + i, M = I
    (`case_generator.cpp:56-77`)."""
    Tx = sp.diags([-np.ones(nx - 1), 2 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    Ty = sp.diags([-np.ones(ny - 1), 2 * np.ones(ny), -np.ones(ny - 1)], [-1, 0, 1])
    K = sp.kron(sp.identity(ny), Tx) + sp.kron(Ty, sp.identity(nx))
    return _csc64(K), _csc64(sp.identity(nx * ny))


def grid_graph_2d_spectrum(nx: int, ny: int) -> np.ndarray:
    """`case_generator.cpp:79-93`."""
    tx = 2.0 - 2.0 * np.cos(math.pi * np.arange(1, nx + 1) / (nx + 1))
    ty = 2.0 - 2.0 * np.cos(math.pi * np.arange(1, ny + 1) / (ny + 1))
    return np.sort((ty[:, None] + tx[None, :]).ravel())


def grid_graph_3d(nx: int, ny: int, nz: int):
    """7-point Laplacian, vertex id (k*ny + j)*nx + i, M = I
    (`case_generator.cpp:95-122`)."""
    def T(m):
        return sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    K = (sp.kron(Iz, sp.kron(Iy, T(nx))) + sp.kron(Iz, sp.kron(T(ny), Ix))
         + sp.kron(T(nz), sp.kron(Iy, Ix)))
    return _csc64(K), _csc64(sp.identity(nx * ny * nz))


def grid_graph_3d_spectrum(nx: int, ny: int, nz: int) -> np.ndarray:
    """`case_generator.cpp:124-146`."""
    ax = lambda m: 2.0 - 2.0 * np.cos(math.pi * np.arange(1, m + 1) / (m + 1))
    w = (ax(nz)[:, None, None] + ax(ny)[None, :, None] + ax(nx)[None, None, :])
    return np.sort(w.ravel())


# --------------------------------------------------------------------------
# Hexahedral elasticity
# --------------------------------------------------------------------------

def _lagrange_1d(order: int):
    """1-D Lagrange basis on [-1, 1] with equispaced nodes; returns
    (values(x) -> (order+1,), derivatives(x) -> (order+1,))."""
    if order == 1:
        f = lambda x: np.array([(1 - x) / 2, (1 + x) / 2])
        d = lambda x: np.array([-0.5, 0.5])
    elif order == 2:
        f = lambda x: np.array([x * (x - 1) / 2, 1 - x * x, x * (x + 1) / 2])
        d = lambda x: np.array([x - 0.5, -2 * x, x + 0.5])
    else:
        raise ValueError("order must be 1 or 2")
    return f, d


def _gauss(npt: int):
    if npt == 2:
        g = 1.0 / math.sqrt(3.0)
        return np.array([-g, g]), np.array([1.0, 1.0])
    if npt == 3:
        g = math.sqrt(0.6)
        return np.array([-g, 0.0, g]), np.array([5 / 9, 8 / 9, 5 / 9])
    raise ValueError(npt)


def hex_element_matrices(order: int, E: float = 1.0, nu: float = 0.3,
                         rho: float = 1.0, h: float = 1.0):
    """Reference element stiffness and consistent mass of a cube of edge h.

    Local node (a, b, c), a fastest, a,b,c in 0..order; local dof
    3*(a + (order+1)*(b + (order+1)*c)) + comp.  Full Gauss quadrature
    (2^3 for trilinear, 3^3 for triquadratic), exact for the mass and the
    stiffness of an undistorted cube."""
    p1 = order + 1
    nn = p1 ** 3
    f, d = _lagrange_1d(order)
    xg, wg = _gauss(order + 1)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    jac = h / 2.0                 # dx/dxi
    detJ = jac ** 3
    Ke = np.zeros((3 * nn, 3 * nn))
    Me = np.zeros((3 * nn, 3 * nn))
    for iz, zz in enumerate(xg):
        for iy, yy in enumerate(xg):
            for ix, xx in enumerate(xg):
                w = wg[ix] * wg[iy] * wg[iz] * detJ
                fx, fy, fz = f(xx), f(yy), f(zz)
                dx, dy, dz = d(xx) / jac, d(yy) / jac, d(zz) / jac
                N = np.einsum("c,b,a->cba", fz, fy, fx).ravel()
                Nx = np.einsum("c,b,a->cba", fz, fy, dx).ravel()
                Ny = np.einsum("c,b,a->cba", fz, dy, fx).ravel()
                Nz = np.einsum("c,b,a->cba", dz, fy, fx).ravel()
                B = np.zeros((6, 3 * nn))
                B[0, 0::3] = Nx
                B[1, 1::3] = Ny
                B[2, 2::3] = Nz
                B[3, 0::3] = Ny
                B[3, 1::3] = Nx
                B[4, 1::3] = Nz
                B[4, 2::3] = Ny
                B[5, 0::3] = Nz
                B[5, 2::3] = Nx
                Ke += w * (B.T @ D @ B)
                NN = np.outer(N, N) * (w * rho)
                for c in range(3):
                    Me[c::3, c::3] += NN
    return 0.5 * (Ke + Ke.T), 0.5 * (Me + Me.T)


@dataclass
class HexProblem:
    """A generated elasticity pencil plus what the adjoint needs."""
    K: sp.csc_matrix
    M: sp.csc_matrix
    coords: np.ndarray            # (n, 3) coordinates of each free dof's node
    conn: np.ndarray              # (n_elem, ndof_e) free-dof index or -1 (clamped)
    Ke0: np.ndarray               # reference element stiffness (E = 1)
    Me0: np.ndarray               # reference element mass (rho = 1)
    kscale: np.ndarray            # per-element stiffness scale
    rscale: np.ndarray            # per-element density scale
    order: int
    dims: tuple
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.K.shape[0]


def hex_connectivity(order: int, nx: int, ny: int, nz: int):
    """Element -> free-dof table for a grid with the z = 0 face clamped."""
    p1 = order + 1
    NX, NY = order * nx + 1, order * ny + 1
    a = np.arange(p1)
    la, lb, lc = np.meshgrid(a, a, a, indexing="ij")
    # local node order: a fastest -> flatten with c slowest
    la = la.transpose(2, 1, 0).ravel()
    lb = lb.transpose(2, 1, 0).ravel()
    lc = lc.transpose(2, 1, 0).ravel()
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex = ex.transpose(2, 1, 0).ravel()
    ey = ey.transpose(2, 1, 0).ravel()
    ez = ez.transpose(2, 1, 0).ravel()
    gx = order * ex[:, None] + la[None, :]
    gy = order * ey[:, None] + lb[None, :]
    gz = order * ez[:, None] + lc[None, :]
    node = ((gz - 1) * NY + gy) * NX + gx
    node = np.where(gz >= 1, node, -1)
    dof = np.where(node[:, :, None] >= 0, 3 * node[:, :, None] + np.arange(3)[None, None, :], -1)
    return dof.reshape(dof.shape[0], -1).astype(np.int64)


def hex_elasticity(order: int, nx: int, ny: int, nz: int, *, E: float = 1.0,
                   nu: float = 0.3, rho: float = 1.0,
                   kscale: Optional[np.ndarray] = None,
                   rscale: Optional[np.ndarray] = None) -> HexProblem:
    """Assemble K and M on an nx x ny x nz grid of unit hexes, z = 0 clamped."""
    ne = nx * ny * nz
    Ke0, Me0 = hex_element_matrices(order, 1.0, nu, 1.0)
    Ke0 = Ke0 * E
    Me0 = Me0 * rho
    ks = np.ones(ne) if kscale is None else np.asarray(kscale, dtype=np.float64)
    rs = np.ones(ne) if rscale is None else np.asarray(rscale, dtype=np.float64)
    conn = hex_connectivity(order, nx, ny, nz)
    NX, NY, NZ = order * nx + 1, order * ny + 1, order * nz + 1
    n = 3 * NX * NY * (NZ - 1)
    nd = conn.shape[1]
    rows = np.repeat(conn, nd, axis=1)          # (ne, nd*nd) row dof
    cols = np.tile(conn, (1, nd))               # (ne, nd*nd) col dof
    # assemble the upper triangle only, then mirror it: stored values are
    # bitwise symmetric and the structural pattern (every dof pair sharing an
    # element, explicit zeros included) is the same for K and M, as an
    # FE assembly through triplets produces it
    keep = (rows >= 0) & (cols >= 0) & (rows <= cols)
    kv = (ks[:, None] * Ke0.ravel()[None, :])[keep]
    mv = (rs[:, None] * Me0.ravel()[None, :])[keep]
    r, c = rows[keep], cols[keep]
    U = sp.coo_matrix((np.ones_like(kv), (r, c)), shape=(n, n)).tocsr()
    U.sum_duplicates()
    # accumulate values in a fixed order (element-major) on the upper pattern
    key = r * n + c
    ukey = (np.repeat(np.arange(n, dtype=np.int64), np.diff(U.indptr)) * n
            + U.indices.astype(np.int64))
    pos = np.searchsorted(ukey, key)
    kU = np.zeros(len(ukey))
    mU = np.zeros(len(ukey))
    np.add.at(kU, pos, kv)
    np.add.at(mU, pos, mv)
    ur = ukey // n
    uc = ukey % n
    off = ur != uc
    R = np.concatenate([ur, uc[off]])
    C = np.concatenate([uc, ur[off]])
    K = _csc_pattern(R, C, np.concatenate([kU, kU[off]]), n)
    M = _csc_pattern(R, C, np.concatenate([mU, mU[off]]), n)
    gz, gy, gx = np.meshgrid(np.arange(1, NZ), np.arange(NY), np.arange(NX), indexing="ij")
    nodexyz = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1).astype(np.float64) / order
    coords = np.repeat(nodexyz, 3, axis=0)
    return HexProblem(K=K, M=M, coords=coords, conn=conn, Ke0=Ke0, Me0=Me0,
                      kscale=ks, rscale=rs, order=order, dims=(nx, ny, nz))


# --------------------------------------------------------------------------
# The named workloads (seeds fixed here, reported with every measurement)
# --------------------------------------------------------------------------

SEEDS = {"hex27_10": 1234, "batch64": 4321, "hex27_14": 2744, "fd_params": 20}


def case_hex8_10() -> HexProblem:
    """8-node hex 10x10x10, E = 1, nu = 0.3, rho = 1: n = 3,630."""
    pb = hex_elasticity(1, 10, 10, 10)
    pb.meta.update(name="hex8_10x10x10", nLevels=3, nev=20)
    return pb


def case_hex27_10(seed: int = SEEDS["hex27_10"]) -> HexProblem:
    """27-node hex 10x10x10, E scale ~ lognormal(0, 0.3): n = 26,460."""
    rng = np.random.default_rng(seed)
    ks = rng.lognormal(0.0, 0.3, 1000)
    pb = hex_elasticity(2, 10, 10, 10, kscale=ks)
    pb.meta.update(name="hex27_10x10x10", seed=seed, nLevels=6, nev=100)
    return pb


def case_hex27_14(seed: int = SEEDS["hex27_14"]) -> HexProblem:
    """27-node hex 14^3 with 2,744 stiffness and 2,744 density scales: n = 70,644."""
    rng = np.random.default_rng(seed)
    ks = rng.lognormal(0.0, 0.3, 14 ** 3)
    rs = rng.uniform(0.5, 2.0, 14 ** 3)
    pb = hex_elasticity(2, 14, 14, 14, kscale=ks, rscale=rs)
    pb.meta.update(name="hex27_14x14x14", seed=seed, nLevels=6, nev=50)
    return pb


def case_batch64(seed: int = SEEDS["batch64"], count: int = 64):
    """64 independent 8-node hex problems, nx,ny,nz ~ U{6..16},
    E ~ lognormal(0, 0.5), rho ~ U[0.5, 2]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny, nz = (int(v) for v in rng.integers(6, 17, size=3))
        ne = nx * ny * nz
        ks = rng.lognormal(0.0, 0.5, ne)
        rs = rng.uniform(0.5, 2.0, ne)
        pb = hex_elasticity(1, nx, ny, nz, kscale=ks, rscale=rs)
        pb.meta.update(name=f"hex8_{nx}x{ny}x{nz}", nev=20)
        out.append(pb)
    return out
