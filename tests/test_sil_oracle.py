"""Pins of the oracle's shift-invert Lanczos (SIL) local solver and of the
dense / SIL switch (src/amls/shift_invert_lanczos.hpp:32-203,
src/amls/eigenpair_solver.hpp:32-82), by what the mathematics fixes:
a full Krylov space reproduces dense eigh, Ritz values interlace from above,
a start vector confined to half of a doubled spectrum finds each double
eigenvalue once, plus the reference's sigma / ncv / pick rules, its start
vector generator, and which nodes of the five workloads take SIL."""
import ctypes
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from amls_amd import cases
from oracle import amls_oracle as O
from tests.helpers import tree_of


def lap1d(n, scale=1.0):
    return scale * (np.diag(2.0 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1))


def sparse_pencil(n, seed):
    """SPD pencil of order n with a 3D-grid-like sparsity (density < 0.15)."""
    rng = np.random.default_rng(seed)
    A = lap1d(n) + np.diag(rng.uniform(0.0, 0.5, n))
    for _ in range(2 * n):                     # a few long-range couplings
        i, j = rng.integers(0, n, 2)
        if i != j:
            v = rng.uniform(0, 0.1)
            A[i, j] -= v
            A[j, i] -= v
            A[i, i] += v
            A[j, j] += v
    Md = np.diag(rng.uniform(0.5, 1.5, n)) + 0.05 * (np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))
    return A, Md


def test_glibc_rand_matches_libc():
    """The restated generator equals this platform's rand() after srand()."""
    try:
        libc = ctypes.CDLL("libc.so.6")
    except OSError:
        pytest.skip("no glibc")
    for seed in (12345, 1, 0, 987654321):
        libc.srand(seed)
        ref = np.array([libc.rand() for _ in range(700)])
        assert np.array_equal(O.glibc_rand(seed, 700), ref), seed


def test_start_vector_is_eigen_random_formula():
    r = O.glibc_rand(12345, 50).astype(float)
    v = O.sil_start_vector(50)
    assert np.array_equal(v, -1.0 + (2.0 * r) / 2147483647.0)
    assert v.min() >= -1.0 and v.max() <= 1.0


def test_full_krylov_space_is_exact():
    """ncv = n: the Lanczos basis spans everything, every Ritz pair is an
    eigenpair of (K - sigma M, M) mapped back through lambda = 1/nu + sigma."""
    n = 220
    K, M = sparse_pencil(n, 1)
    lz = O.ShiftInvertLanczos(K, M, -1e-9)
    m = lz.iterate(n)
    assert m == n
    lam = np.sort([lz.lam(j) for j in range(m)])
    w = sla.eigh(K, M, eigvals_only=True)
    assert np.abs(lam - w).max() <= 1e-10 * w.max()
    X = np.stack([lz.ritz_vector(j) for j in range(m)], axis=1)
    assert np.abs(X.T @ M @ X - np.eye(m)).max() <= 1e-9
    th = np.array([lz.lam(j) for j in range(m)])
    r = K @ X - (M @ X) * th
    assert np.linalg.norm(r, axis=0).max() <= 1e-8 * np.linalg.norm(K, 2)


def test_ritz_values_interlace_from_above():
    """ncv < n: the j-th smallest Ritz value is >= the j-th smallest
    eigenvalue (Courant-Fischer on the Krylov subspace), and the extremal
    ones have converged after 48 steps."""
    n = 300
    K, M = sparse_pencil(n, 2)
    w = sla.eigh(K, M, eigvals_only=True)
    lz = O.ShiftInvertLanczos(K, M, O._sil_sigma(K))
    m = lz.iterate(48)
    ritz = np.sort([lz.lam(j) for j in range(m)])
    assert np.all(ritz >= w[:m] * (1 - 1e-12))
    assert np.abs(ritz[:4] - w[:4]).max() <= 1e-9 * w[3]


def test_double_eigenvalues(monkeypatch):
    """K = blkdiag(A, A), M = I, every eigenvalue double.  A start vector
    (x1, 0) keeps the Krylov space {(p(A) x1, 0)} exactly (the factor is
    block-diagonal, so the zero half stays zero in floating point too): each
    double eigenvalue is found once and the interval solve keeps half of the
    dense count.  With the reference's full start vector the second copies
    are seeded by rounding, which shift-invert amplifies, and the lowest ten
    double values are all found within the 96 steps."""
    A = lap1d(160, 1.0)
    K = sla.block_diag(A, A)
    M = np.eye(320)
    w = np.sort(sla.eigh(A, eigvals_only=True))
    hi = 0.5 * (w[9] + w[10])             # ten distinct values below hi
    wd = sla.eigh(K, M, eigvals_only=True, subset_by_value=(-np.inf, hi))
    assert wd.size == 20
    orig = O.sil_start_vector
    monkeypatch.setattr(O, "sil_start_vector",
                        lambda n, seed=12345: np.concatenate([orig(n // 2, seed), np.zeros(n - n // 2)]))
    ws, V, lz = O.sil_solve_interval(K, M, -1e300, hi)
    assert ws.size == 10 and np.all(V[160:] == 0.0)
    assert np.abs(np.sort(ws) - w[:10]).max() <= 1e-9 * w[9]
    # the reference's interval result is ordered by ascending nu
    assert np.all(np.diff(ws) < 0)
    monkeypatch.setattr(O, "sil_start_vector", orig)
    ws, V, lz = O.sil_solve_interval(K, M, -1e300, hi)
    assert ws.size == 20
    assert np.abs(np.sort(ws) - wd).max() <= 1e-8 * wd.max()


def test_residual_estimate_reads_unstored_beta():
    """iterate() breaks before storing beta_{m-1}; res() is zero for every
    Ritz value, whether the loop stopped at ncv or at a breakdown."""
    K, M = sparse_pencil(250, 3)
    lz = O.ShiftInvertLanczos(K, M, -1e-9)
    for ncv in (20, 96):
        m = lz.iterate(ncv)
        assert m == ncv and lz.beta[m - 1] == 0.0
        assert all(lz.res(j) == 0.0 for j in range(m))


def test_sigma_and_ncv_rules():
    n = 260
    K, M = sparse_pencil(n, 4)
    scale = np.abs(np.diag(K)).max()
    assert O._sil_sigma(K) == -max(1e-10, 1e-9 * scale)
    assert O._sil_sigma(K, -1e300) == O._sil_sigma(K)
    assert O._sil_sigma(K, -2.0) == -2.0 - max(1e-10, 1e-9 * scale)
    assert O._sil_sigma(1e-3 * np.eye(3)) == -1e-10
    # index: ncv = max(4k + 16, 48), no doubling (every estimate is zero)
    for k, ncv in ((4, 48), (20, 96), (60, 256)):
        w, V, lz = O.sil_solve_index(K, M, 1, k)
        assert lz.m == min(n, ncv) and w.size == k
        assert np.all(np.diff(w) > 0)
    # interval: 96, doubled while every Ritz value lies below hi
    w, V, lz = O.sil_solve_interval(K, M, -1e300, 1e30)
    assert lz.m == n and w.size == n           # 96 -> 192 -> 260 (= n)
    wd = sla.eigh(K, M, eigvals_only=True)
    assert np.abs(np.sort(w) - wd).max() <= 1e-10 * wd.max()
    w, V, lz = O.sil_solve_interval(K, M, -1e300, 0.5 * wd[5] + 0.5 * wd[6])
    assert lz.m == 96 and w.size == 6
    # pick: (lo, hi] only
    w, V, lz = O.sil_solve_interval(K, M, 0.5 * (wd[1] + wd[2]), 0.5 * (wd[5] + wd[6]))
    assert w.size == 4


def test_dispatch_rules():
    """useSil_: n >= 200 and nnz < 0.15 n^2; index path only when k < 4n/5,
    dense fallback when SIL returns fewer than k pairs or none."""
    K, M = sparse_pencil(250, 5)
    assert O.use_sil(K)
    assert not O.use_sil(K[:199, :199])
    dense = np.ones((250, 250)) + 250 * np.eye(250)
    assert not O.use_sil(dense)
    assert O.use_sil(dense, sil_mode=1) and not O.use_sil(K, sil_mode=-1)
    w, V, path = O.local_eig_dispatch(K, M, "fixed", 10)
    assert path == "sil" and w.size == 10
    w, V, path = O.local_eig_dispatch(K, M, "fixed", 200)       # k >= 4n/5
    assert path == "dense" and w.size == 200
    wd = sla.eigh(K, M, eigvals_only=True)
    w, V, path = O.local_eig_dispatch(K, M, "threshold", 0.5 * wd[0])   # no Ritz value below
    assert path == "dense" and w.size == 0
    w, V, path = O.local_eig_dispatch(K, M, "threshold", 0.5 * (wd[7] + wd[8]))
    assert path == "sil" and w.size == 8
    # the SIL pairs agree with dense eigh where converged, and are M-orthonormal
    assert np.abs(np.sort(w) - wd[:8]).max() <= 1e-9 * wd[7]
    assert np.abs(V.T @ M @ V - np.eye(8)).max() <= 1e-9


def _leaf_sil(pb, nLevels, leafSize):
    """useSil_ on every leaf of the library's tree, from the leaf's own K
    block (a leaf's K-hat is the original block)."""
    tree, t = tree_of(pb.K, pb.M, pb.coords, nLevels, leafSize)
    Kr = pb.K.tocsr()[tree.perm][:, tree.perm]
    out = []
    for p in range(tree.N):
        if t["level"][p] != 0:
            continue
        o, n = int(tree.offset[p]), int(tree.size[p])
        blk = Kr[o:o + n, o:o + n]
        blk.eliminate_zeros()
        out.append((p, n, blk.nnz / float(n * n), n >= 200 and blk.nnz < 0.15 * n * n))
    return tree, t, out


def test_sil_nodes_of_the_workloads(blas1):
    """Which nodes the reference's switch sends to SIL, per workload:
    every leaf of the 8-node 10^3 problem (densities 0.10-0.13); none of the
    27-node 10^3 and 14^3 leaves (densities >= 0.18); separators are dense
    after fill on the 8-node problem; the 64-problem batch has SIL leaves in
    many problems.  (The separators of the 27-node problems are checked on
    the device, from the elimination's non-zero counts.)"""
    c1 = cases.case_hex8_10()
    tree, t, leaves = _leaf_sil(c1, 3, 500)
    assert len(leaves) == 8 and all(s for (_, _, _, s) in leaves)
    assert all(0.09 < d < 0.15 for (_, _, d, _) in leaves)
    perm = tree.perm
    ge = O.BlockGE(c1.K.tocsr()[perm][:, perm], c1.M.tocsr()[perm][:, perm], tree).run()
    for p in range(tree.N):
        if t["level"][p] > 0:
            assert not O.use_sil(ge.Khat[p]), p
    for pb in (cases.case_hex27_10(),):
        _, _, leaves = _leaf_sil(pb, 6, 500)
        assert leaves and not any(s for (_, _, _, s) in leaves)
        assert min(d for (_, _, d, _) in leaves) > 0.17
    batch = cases.case_batch64()
    nsil = [sum(s for (_, _, _, s) in _leaf_sil(pb, 6, 500)[2]) for pb in batch[:12]]
    assert sum(nsil) > 0
    print("SIL leaves in the first 12 batch problems:", nsil)


def test_sil_leaves_hex27_14():
    """The 27-node 14^3 problem's leaves are larger (540-1,029 dofs) and
    sparser: 56 of its 64 leaves fall below the switch's density of 0.15."""
    pb = cases.case_hex27_14()
    _, _, leaves = _leaf_sil(pb, 6, 500)
    assert len(leaves) == 64
    assert sum(s for (_, _, _, s) in leaves) == 56
    assert all(s == (d < 0.15) for (_, n, d, s) in leaves)


def test_oracle_c1_takes_sil_on_leaves(blas1):
    """The oracle's full two-pass solve of the 8-node 10^3 workload runs SIL
    on all eight leaves in both passes; separators stay dense.  The kept
    local modes (2 to 6 per leaf) all converge within SIL's 48 and 96 steps,
    double eigenvalues included, so p and the eigenvalues equal the
    all-dense oracle's to rounding."""
    pb = cases.case_hex8_10()
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 500)
    r = O.amls_solve(pb.K, pb.M, tree, nev=20, recoverVectors=False)
    leaves = [p for p in range(tree.N) if t["level"][p] == 0]
    assert all(r.cond.pathA[p] == "sil" for p in leaves)
    assert all(r.cond.path[p] == "sil" for p in leaves)
    assert all(r.cond.path[p] == "dense" for p in range(tree.N) if p not in leaves)
    d = O.amls_solve(pb.K, pb.M, tree, nev=20, recoverVectors=False, sil_mode=-1)
    assert r.p == d.p and np.array_equal(r.k, d.k)
    assert np.abs(r.theta - d.theta).max() <= 1e-12 * d.theta.max()


def test_oracle_sil_differs_from_dense_when_unconverged(blas1):
    """An 8-node 8 x 8 x 9 problem with lognormal element stiffness (no exact
    multiplicities), 2 levels, eta = 10: the four 432-dof leaves keep more
    modes than 96 Lanczos steps converge, so SIL keeps fewer (and higher)
    Ritz values than dense and p drops; a 1e-13 perturbation of the start
    vector moves them by no more than 1e-12 (well-conditioned, so a second
    implementation can match them)."""
    pb = sil_case()
    tree, t = tree_of(pb.K, pb.M, pb.coords, 2, 500)
    r = O.amls_solve(pb.K, pb.M, tree, nev=20, recoverVectors=False, thresholdFactor=10.0)
    d = O.amls_solve(pb.K, pb.M, tree, nev=20, recoverVectors=False, thresholdFactor=10.0,
                     sil_mode=-1)
    leaves = [p for p in range(tree.N) if t["level"][p] == 0]
    assert all(r.cond.path[p] == "sil" for p in leaves)
    assert r.p < d.p
    assert np.all(r.k[leaves] < d.k[leaves])
    assert np.abs(r.theta - d.theta).max() > 1e-5 * d.theta.max()
    p = leaves[0]
    Kh, Mh = r.ge.Khat[p], r.ge.Mhat[p]
    w0, _, _ = O.sil_solve_interval(Kh, Mh, -1e300, r.lambda_cut)
    orig = O.sil_start_vector
    rng = np.random.default_rng(1)
    try:
        O.sil_start_vector = lambda n, seed=12345: orig(n, seed) * (1 + 1e-13 * rng.standard_normal(n))
        w1, _, _ = O.sil_solve_interval(Kh, Mh, -1e300, r.lambda_cut)
    finally:
        O.sil_start_vector = orig
    assert w1.size == w0.size
    assert np.abs(np.sort(w1) - np.sort(w0)).max() <= 1e-12 * w0.max()


def sil_case():
    rng = np.random.default_rng(77)
    dims = (8, 8, 9)
    ne = int(np.prod(dims))
    return cases.hex_elasticity(1, *dims, kscale=rng.lognormal(0, 0.5, ne),
                                rscale=rng.uniform(0.5, 2.0, ne))
