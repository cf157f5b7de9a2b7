"""CPU oracle for the AMLS hot path -- test infrastructure, not product code.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker / the timed CPU baseline.  Nothing in
amls_amd/ imports it.

It is a numpy/scipy restatement of the reference's CPU algorithm, run on a
given separator tree (the reference partitions with METIS, which is absent;
both the oracle and the GPU path therefore take the tree produced by
libamls_hip's own nested dissection):

* block Gaussian elimination in post-order with the reference's block
  formulas and zero-row compression -- src/amls/gaussian_elimination.cpp
  :458-508 (factor), :510-598 (X, Y on non-zero rows), :653-676 (WX),
  :678-715 (K Schur), :717-781 (M congruence);
* local eigenproblems with LAPACK ?sygvx semantics (itype 1, lower,
  RANGE 'I' / 'V', abstol 0) -- src/linalg/dense_eigensolver.cpp:44-133,
  dispatched as in src/amls/subspace_condensation.cpp:45-100;
* truncation policies -- src/amls/spectral_truncation.hpp:35-88;
* post-order projection (mid update, right projection, left projection) --
  src/amls/subspace_condensation.cpp:129-261;
* reduced dense pencil with identity diagonal blocks and node-id ordered
  offsets -- src/amls/solver.cpp:72-96, :145-161, subspace_condensation.cpp
  :120-127;
* two-pass truncation (k0 formula, lambda_hat, lambda_cut) --
  src/amls/solver.cpp:237-267;
* back-transformation, un-permutation and M-normalisation --
  src/amls/subspace_condensation.cpp:263-319, src/amls/solver.cpp:296-311.

* the local dense / shift-invert-Lanczos dispatch (n >= 200 and fewer than
  0.15 n^2 non-zeros in K-hat; index path only for il = 1 and k < 4n/5;
  dense fallback when SIL returns fewer than k pairs or none) --
  src/amls/eigenpair_solver.hpp:32-82, and SIL itself (sigma rule, ncv
  schedule, M-inner-product Lanczos with two full modified-Gram-Schmidt
  passes, residual estimate, pick rule, Ritz vectors) --
  src/amls/shift_invert_lanczos.hpp:32-203.

Deviations, all value-preserving: the sparse Eigen LDL^T of leaves and the
dense Eigen LDLT of separators are both done as a dense Cholesky (same
factorisation of an SPD matrix, different rounding).  The SIL start vector
is std::srand(12345) + Eigen's Vec::Random as glibc and Eigen <= 3.4 produce
it (glibc_rand restates glibc's generator; the reference's own platform
library may draw other numbers, so its unconverged Ritz values are pinned
only for that generator).

SIL's residual estimate beta_{m-1} s_{m-1,j} reads beta_{m-1}, which
iterate() never stores (it leaves the loop first), so it is always zero:
solveIndex stops at its first ncv, solveInterval stops unless no Ritz value
or every Ritz value lies below hi, and the pick keeps every Ritz value in
(lo, hi].  This module does the same, literally.

Parity pins for this module live in tests/test_oracle.py: the closed-form
spectra and structural properties of the reference's unit and regression
tests, full-keep exactness, brute-force dense Schur complements and
congruences for every node block, the committed golden eigenvalue fixtures,
and planted mistakes (the `plant` arguments) that those checks must reject.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp


@dataclass
class Tree:
    """Separator tree given by the GPU library's analysis, in post-order
    positions.  treeId = the reference's node id (preorder), used for the
    reduced-basis ordering."""
    treeId: np.ndarray
    offset: np.ndarray
    size: np.ndarray
    parent: np.ndarray
    perm: np.ndarray

    @property
    def N(self) -> int:
        return len(self.size)

    def ancestors(self, p: int) -> List[int]:
        out = []
        q = int(self.parent[p])
        while q >= 0:
            out.append(q)
            q = int(self.parent[q])
        return out

    @staticmethod
    def from_dict(d: dict) -> "Tree":
        return Tree(d["treeId"], d["offset"], d["size"], d["parent"], d["perm"])


@dataclass
class OracleResult:
    theta: np.ndarray
    Z: Optional[np.ndarray]
    p: int
    k: np.ndarray                  # per post-order node
    mu: List[np.ndarray]
    Zloc: List[np.ndarray]
    lambda_cut: float = -1.0
    p0: int = 0
    timing: Dict[str, float] = field(default_factory=dict)
    ge: object = None
    cond: object = None


def _nzrows(A: np.ndarray) -> np.ndarray:
    """Rows with a non-zero entry (the reference's nonzeroColIdx on expandT /
    the row mask of a sparse block, gaussian_elimination.cpp:63-83, 519-532)."""
    return np.flatnonzero(np.any(A != 0.0, axis=1))


class BlockGE:
    """GaussianElimination restated on dense blocks (post-order, serial)."""

    def __init__(self, K: sp.spmatrix, M: sp.spmatrix, tree: Tree):
        self.tree = tree
        N = tree.N
        off, size = tree.offset, tree.size
        K = sp.csr_matrix(K)
        M = sp.csr_matrix(M)
        self.anc = [tree.ancestors(p) for p in range(N)]
        # diagonal blocks and ancestor couplings (a, p): n_a x n_p
        self.Kd, self.Md = [], []
        self.Kc: Dict[tuple, np.ndarray] = {}
        self.Mc: Dict[tuple, np.ndarray] = {}
        for p in range(N):
            o, n = int(off[p]), int(size[p])
            self.Kd.append(K[o:o + n, o:o + n].toarray())
            self.Md.append(M[o:o + n, o:o + n].toarray())
        Kcsc, Mcsc = K.tocsc(), M.tocsc()
        for p in range(N):
            o, n = int(off[p]), int(size[p])
            Kcol = Kcsc[:, o:o + n].tocsr()
            Mcol = Mcsc[:, o:o + n].tocsr()
            for a in self.anc[p]:
                oa, na = int(off[a]), int(size[a])
                self.Kc[(a, p)] = Kcol[oa:oa + na].toarray()
                self.Mc[(a, p)] = Mcol[oa:oa + na].toarray()
        self.Khat: List[Optional[np.ndarray]] = [None] * N
        self.Mhat: List[Optional[np.ndarray]] = [None] * N
        self.fac = [None] * N

    def solve(self, p: int, B: np.ndarray) -> np.ndarray:
        return sla.cho_solve(self.fac[p], B, check_finite=False)

    def run(self, plant: Optional[str] = None):
        """gaussian_elimination.cpp:458-826 in post-order (the serial
        run()/runLevel order; the DAG order computes the same sums).

        plant: a deliberate mistake for the tests that show the parity checks
        catch it -- "drop_wx" (no +K WX term in the M congruence),
        "schur_sign" (K Schur update added instead of subtracted),
        "coup_transpose" (the (K_lo,i Y_hi)^T coupling term untransposed on a
        square block, dropped otherwise)."""
        tr = self.tree
        for p in range(tr.N):
            self.Khat[p] = self.Kd[p].copy()
            self.Mhat[p] = self.Md[p].copy()
            anc = self.anc[p]
            fac, deltas = ge_node(self.Kd[p], self.Mhat[p], anc,
                                  [self.Kc[(a, p)] for a in anc], [self.Mc[(a, p)] for a in anc],
                                  int(tr.treeId[p]), plant)
            self.fac[p] = fac
            self.apply(deltas)
        return self

    def apply(self, deltas):
        """Add the updates ge_node returned into the ancestors' blocks, in
        order."""
        for kind, key, (ri, ci), delta in deltas:
            tgt = {"Kd": self.Kd, "Md": self.Md, "Kc": self.Kc, "Mc": self.Mc}[kind][key]
            tgt[np.ix_(ri, ci)] += delta


def ge_node(Kh: np.ndarray, W: np.ndarray, anc: List[int], Kc_p: List[np.ndarray],
            Mc_p: List[np.ndarray], node_id: int, plant: Optional[str] = None):
    """One node of BlockGE.run (gaussian_elimination.cpp:458-826): the
    factor of K-hat_pp and the updates the node sends to its ancestors'
    blocks, as (block kind, key, (rows, cols), delta) to add in order.  It
    reads only the node's own blocks (K-hat_pp, M-hat_pp and its couplings
    Kc_p[j] = K_(anc[j], p), Mc_p[j] = M_(anc[j], p)), so nodes of one tree
    level, which have no ancestor in common with each other's own blocks,
    can run concurrently with the same results."""
    try:
        c, low = sla.cho_factor(Kh, lower=True, check_finite=False)
    except np.linalg.LinAlgError as e:   # D > 0 check (sparse_direct.cpp:13-22)
        raise np.linalg.LinAlgError(f"node {node_id}: not SPD") from e
    d = np.diag(c)
    if not np.all(np.isfinite(d)) or np.any(d <= 0):
        raise np.linalg.LinAlgError(f"node {node_id}: not SPD")
    fac = (c, low)
    deltas = []
    if not anc:
        return fac, deltas
    solve = lambda B: sla.cho_solve(fac, B, check_finite=False)
    n_p = Kh.shape[0]
    Kc = dict(zip(anc, Kc_p))
    Mc = dict(zip(anc, Mc_p))
    X, Y, WX, rK, rM = {}, {}, {}, {}, {}
    for a in anc:                               # solveXYAnc_ (:510-598)
        Ka, Ma = Kc[a], Mc[a]
        na = Ka.shape[0]
        rK[a], rM[a] = _nzrows(Ka), _nzrows(Ma)
        X[a] = np.zeros((n_p, na))
        Y[a] = np.zeros((n_p, na))
        if rK[a].size:
            X[a][:, rK[a]] = solve(Ka[rK[a]].T)
        if rM[a].size:
            Y[a][:, rM[a]] = solve(Ma[rM[a]].T)
        WX[a] = np.zeros((n_p, na))             # makeWX_ (:653-676)
        if rK[a].size:
            WX[a][:, rK[a]] = solve(W @ X[a][:, rK[a]])
    for a in anc:                               # schurKDiagTask_ (:678-693)
        r = rK[a]
        if r.size:
            upd = Kc[a][r] @ X[a][:, r]
            deltas.append(("Kd", a, (r, r), upd if plant == "schur_sign" else -upd))
    for u in range(len(anc)):                   # schurKCoupTask_ (:695-715)
        for v in range(u + 1, len(anc)):
            lo, hi = anc[u], anc[v]
            rh, rl = rK[hi], rK[lo]
            if rh.size and rl.size:
                deltas.append(("Kc", (hi, lo), (rh, rl), -(Kc[hi][rh] @ X[lo][:, rl])))
    for a in anc:                               # congDiagTask_ (:717-749)
        B = Kc[a]
        r, rm = rK[a], rM[a]
        if r.size and rm.size:
            T1 = B[r] @ Y[a][:, rm]
            deltas.append(("Md", a, (r, rm), -T1))
            deltas.append(("Md", a, (rm, r), -T1.T))
        if r.size and plant != "drop_wx":
            deltas.append(("Md", a, (r, r), B[r] @ WX[a][:, r]))
    for u in range(len(anc)):                   # congCoupTask_ (:751-781)
        for v in range(u + 1, len(anc)):
            lo, hi = anc[u], anc[v]
            rh, rl = rK[hi], rK[lo]
            rmh, rml = rM[hi], rM[lo]
            if rh.size and rml.size:
                deltas.append(("Mc", (hi, lo), (rh, rml), -(Kc[hi][rh] @ Y[lo][:, rml])))
            if rl.size and rmh.size:
                T2 = Kc[lo][rl] @ Y[hi][:, rmh]
                if plant != "coup_transpose":
                    deltas.append(("Mc", (hi, lo), (rmh, rl), -T2.T))
                elif T2.shape[0] == T2.shape[1]:
                    deltas.append(("Mc", (hi, lo), (rmh, rl), -T2))
            if rh.size and rl.size:
                deltas.append(("Mc", (hi, lo), (rh, rl), Kc[hi][rh] @ WX[lo][:, rl]))
    return fac, deltas


# --------------------------------------------------------------------------
# Truncation (src/amls/spectral_truncation.hpp:35-88) and local solves
# --------------------------------------------------------------------------

def local_eig(Kh: np.ndarray, Mh: np.ndarray, kind: str, value: float):
    """EigenpairSolver::solveLocalIndex / solveLocalInterval with the dense
    ?sygvx path (dense_eigensolver.cpp:44-133): 'fixed' -> RANGE 'I' 1..k
    (k <= 0 keeps all: solveLocalAll), 'threshold' -> RANGE 'V' (-inf, cut]."""
    n = Kh.shape[0]
    if n == 0:
        return np.zeros(0), np.zeros((0, 0))
    if kind == "fixed":
        k = int(value)
        if k <= 0:
            w, V = sla.eigh(Kh, Mh, driver="gvd")
            return w, V
        iu = min(k, n)
        return sla.eigh(Kh, Mh, subset_by_index=[0, iu - 1], driver="gvx")
    cut = float(value)
    if not (cut > -1e300):
        return np.zeros(0), np.zeros((n, 0))
    try:
        w, V = sla.eigh(Kh, Mh, subset_by_value=(-np.inf, cut), driver="gvx")
    except ValueError:
        w, V = np.zeros(0), np.zeros((n, 0))
    return w, V


# --------------------------------------------------------------------------
# Shift-invert Lanczos local solver (src/amls/shift_invert_lanczos.hpp) and
# the dense / SIL dispatch of EigenpairSolver (src/amls/eigenpair_solver.hpp)
# --------------------------------------------------------------------------

RAND_MAX = 2147483647


def glibc_rand(seed: int, count: int) -> np.ndarray:
    """The first `count` values of glibc's rand() after srand(seed): the
    TYPE_3 additive feedback generator (r_i = r_{i-31} + r_{i-3} mod 2^32
    after a Park-Miller seeding of 31 words and 310 discarded outputs,
    output r_i >> 1).  This is what std::srand / std::rand give the
    reference's Lanczos start vector on Linux (shift_invert_lanczos.hpp
    :49-50)."""
    r = np.zeros(344 + count, dtype=np.int64)
    r[0] = seed if seed != 0 else 1
    for i in range(1, 31):
        hi, lo = divmod(int(r[i - 1]), 127773)
        w = 16807 * lo - 2836 * hi
        r[i] = w + 2147483647 if w < 0 else w
    r[31:34] = r[0:3]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return r[344:] >> 1


def sil_start_vector(n: int, seed: int = 12345) -> np.ndarray:
    """Vec::Random(n) after std::srand(seed) (shift_invert_lanczos.hpp:49-50):
    Eigen (<= 3.4) draws x + (y - x) rand() / RAND_MAX with (x, y) = (-1, 1),
    coefficient 0 first."""
    r = glibc_rand(seed, n).astype(np.float64)
    return -1.0 + (2.0 * r) / float(RAND_MAX)


class ShiftInvertLanczos:
    """sil::ShiftInvertLanczos (shift_invert_lanczos.hpp:32-122): M-inner-
    product Lanczos on C = (K - sigma M)^-1 M with two passes of full
    (modified Gram-Schmidt) reorthogonalisation per step."""

    def __init__(self, Kd: np.ndarray, Md: np.ndarray, sigma: float):
        self.n = Kd.shape[0]
        self.shift = sigma
        # sparseView(1e-14) (:37-38) keeps every non-zero entry
        self.M = sp.csr_matrix(Md)
        A = Kd - sigma * Md
        try:                              # EigenSimplicialLDLTSolver: D > 0
            self.fac = sla.cho_factor(A, lower=True, check_finite=False)
        except np.linalg.LinAlgError as e:
            raise np.linalg.LinAlgError("SIL: K - sigma M is not SPD") from e
        self.m = 0

    def _applyC(self, x: np.ndarray) -> np.ndarray:
        return sla.cho_solve(self.fac, self.M @ x, check_finite=False)

    def iterate(self, ncv: int, seed: int = 12345) -> int:
        """iterate (:44-85): at most ncv steps; T's eigenpairs ascending."""
        n, M = self.n, self.M
        ncv = min(ncv, n)
        alpha = np.zeros(ncv)
        beta = np.zeros(ncv)
        q = sil_start_vector(n, seed)
        q = q / math.sqrt(max(float(q @ (M @ q)), 1e-300))
        Q = [q]
        m = 0
        for j in range(ncv):
            w = self._applyC(Q[j])
            alpha[j] = float(Q[j] @ (M @ w))
            for _ in range(2):            # :61-64
                for l in range(j + 1):
                    w = w - float(Q[l] @ (M @ w)) * Q[l]
            bj = math.sqrt(max(float(w @ (M @ w)), 0.0))
            m = j + 1
            if bj < 1e-13 or m == ncv:    # :67-68, before beta[j] is stored
                break
            beta[j] = bj
            Q.append(w / bj)
        self.m = m
        self.alpha, self.beta = alpha, beta
        self.Q = np.stack(Q[:m], axis=1)
        T = np.diag(alpha[:m]) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        self.nu, self.S = np.linalg.eigh(T)   # SelfAdjointEigenSolver: ascending
        return m

    def lam(self, j: int) -> float:
        """lambda (:90): 1 / nu + sigma."""
        return 1.0 / float(self.nu[j]) + self.shift

    def res(self, j: int) -> float:
        """res (:92-95): |beta_{m-1} s_{m-1, j}|.  iterate() leaves the
        loop before storing beta_{m-1}, so this reads the zero it was
        initialised with: every convergence test below passes."""
        bm = self.beta[self.m - 1] if self.m > 1 else 0.0
        return abs(bm * float(self.S[self.m - 1, j]))

    def ritz_vector(self, j: int) -> np.ndarray:
        """ritzVector (:97-102): x = Q s_j, M-normalised with Q."""
        return self.Q @ self.S[:, j]


def _sil_sigma(Kd: np.ndarray, lo: Optional[float] = None) -> float:
    """sigma of solveInterval (:128-135) / solveIndex (:175-178): just left of
    min(0, lo) by max(1e-10, 1e-9 max_j |K_jj|)."""
    scale = float(np.abs(np.diag(Kd)).max()) if Kd.size else 0.0
    eps = max(1e-10, 1e-9 * (scale if scale > 0.0 else 1.0))
    if lo is None:
        return -eps
    lo_sane = lo if lo > -1e100 else 0.0
    return min(0.0, lo_sane) - eps


def sil_solve_interval(Kd, Md, lo: float, hi: float, tol: float = 1e-9):
    """sil::solveInterval (:126-168): Ritz pairs with lambda in (lo, hi],
    ncv 96 doubling to min(n, 1536).  Returned in the reference's order:
    ascending nu, i.e. descending lambda."""
    n = Kd.shape[0]
    lz = ShiftInvertLanczos(Kd, Md, _sil_sigma(Kd, lo))
    ncv = min(n, 96)
    cap = min(n, 1536)
    while True:
        m = lz.iterate(ncv)
        cnt = sum(1 for j in range(m) if lz.lam(j) <= hi)
        if cnt < m and cnt > 0 and lz.res(cnt - 1) <= tol * max(1.0, float(lz.nu[m - 1])):
            break
        if ncv >= cap or m < ncv:
            break
        ncv = min(cap, ncv * 2)
    pick = []
    for j in range(lz.m):
        lam = lz.lam(j)
        if lam > hi or lam <= lo:
            continue
        if lz.res(j) > 100 * tol * max(1.0, float(lz.nu[lz.m - 1])):
            continue
        pick.append(j)
    w = np.array([lz.lam(j) for j in pick])
    V = np.stack([lz.ritz_vector(j) for j in pick], axis=1) if pick else np.zeros((n, 0))
    return w, V, lz


def sil_solve_index(Kd, Md, il: int, iu: int, tol: float = 1e-10):
    """sil::solveIndex (:171-203): the smallest k = iu - il + 1 pairs
    (il = 1), ncv = min(n, max(4 k + 16, 48)) doubling to min(n, 1536)
    while the top k Ritz values are unconverged; ascending lambda."""
    n = Kd.shape[0]
    k = iu - il + 1
    lz = ShiftInvertLanczos(Kd, Md, _sil_sigma(Kd))
    cap = min(n, 1536)
    ncv = min(n, max(4 * k + 16, 48))
    while True:
        m = lz.iterate(ncv)
        want = min(k, m)
        ok = all(lz.res(m - 1 - t) <= tol * max(1.0, float(lz.nu[m - 1])) for t in range(want))
        if ok or ncv >= cap or m < ncv:
            break
        ncv = min(cap, ncv * 2)
    m = lz.m
    kk = min(k, m)
    js = [m - 1 - t for t in range(kk)]
    w = np.array([lz.lam(j) for j in js])
    V = np.stack([lz.ritz_vector(j) for j in js], axis=1) if js else np.zeros((n, 0))
    return w, V, lz


def use_sil(Kd: np.ndarray, sil_mode: int = 0) -> bool:
    """EigenpairSolver::useSil_ (eigenpair_solver.hpp:68-82): n >= 200 and
    fewer than 0.15 n^2 non-zero entries in the dense K block."""
    if sil_mode > 0:
        return True
    if sil_mode < 0:
        return False
    n = Kd.shape[0]
    if n < 200:
        return False
    return float(np.count_nonzero(Kd)) < 0.15 * float(n) * float(n)


def local_eig_dispatch(Kh: np.ndarray, Mh: np.ndarray, kind: str, value: float,
                       sil_mode: int = 0):
    """EigenpairSolver::solveLocalIndex / solveLocalInterval / solveLocalAll
    (eigenpair_solver.hpp:32-63) as called by spectralLevel with the
    truncation's hint (subspace_condensation.cpp:67-80,
    spectral_truncation.hpp:50-82).  Returns (w, V, path) with path 'sil' or
    'dense'."""
    n = Kh.shape[0]
    if kind == "fixed" and int(value) > 0:
        il, iu = 1, int(value)                      # FixedCount indexHint
        if il == 1 and use_sil(Kh, sil_mode):
            k = iu - il + 1
            if k < (n * 4) // 5:
                w, V, _ = sil_solve_index(Kh, Mh, il, iu)
                if w.size == k:
                    return w, V, "sil"
        w, V = local_eig(Kh, Mh, "fixed", iu)
        return w, V, "dense"
    if kind == "threshold":
        lo, hi = -1e300, float(value)               # Threshold intervalHint
        if use_sil(Kh, sil_mode):
            w, V, _ = sil_solve_interval(Kh, Mh, lo, hi)
            if w.size:
                return w, V, "sil"
        w, V = local_eig(Kh, Mh, "threshold", hi)
        return w, V, "dense"
    w, V = local_eig(Kh, Mh, "fixed", 0)            # no hint: solveLocalAll
    return w, V, "dense"


class Condensation:
    """SubspaceCondensation restated (src/amls/subspace_condensation.cpp)."""

    def __init__(self, ge: BlockGE):
        self.ge = ge

    def run(self, kind: str, value: float, pinned_k: Optional[np.ndarray] = None,
            plant: Optional[str] = None, sil_mode: int = 0):
        """plant: "drop_mid" skips the mid update M_a,kk -= K_a,i Khat^-1 M_i,kk
        (a deliberate mistake for the tests of the parity checks).
        sil_mode: EigenpairSolver::silMode (eigenpair_solver.hpp:28-29), 0 =
        the density switch, 1 = always SIL, -1 = always dense."""
        ge, tr = self.ge, self.ge.tree
        N = tr.N
        self.k = np.zeros(N, dtype=np.int64)
        self.mu: List[np.ndarray] = [np.zeros(0)] * N
        self.Z: List[np.ndarray] = [np.zeros((int(tr.size[p]), 0)) for p in range(N)]
        self.path = ["dense"] * N
        # spectralLevel (:45-100): local spectra + select
        for p in range(N):
            if pinned_k is not None:
                w, V = local_eig(ge.Khat[p], ge.Mhat[p], "fixed", max(int(pinned_k[p]), 0)) \
                    if pinned_k[p] > 0 else (np.zeros(0), np.zeros((int(tr.size[p]), 0)))
            else:
                w, V, self.path[p] = local_eig_dispatch(ge.Khat[p], ge.Mhat[p], kind, value,
                                                        sil_mode)
            if kind == "threshold" and pinned_k is None:
                kk = 0
                while kk < w.size and w[kk] <= value:
                    kk += 1
            else:
                kk = w.size
            self.k[p] = kk
            self.mu[p] = w[:kk].copy()
            self.Z[p] = V[:, :kk].copy()
        # projectNode_ in post-order (:129-261)
        mbar: Dict[tuple, np.ndarray] = {}
        rows: Dict[int, List[int]] = {}
        blocks = []
        for i in range(N):
            anc = ge.anc[i]
            rowKK = rows.get(i, [])
            upd, made, blk = project_node(
                ge.fac[i], ge.Mhat[i], self.Z[i], anc, [ge.Kc[(a, i)] for a in anc],
                [ge.Mc[(a, i)] for a in anc], rowKK, [mbar[(i, kk)] for kk in rowKK], plant)
            self.apply_projection(i, upd, made, blk, mbar, rows, blocks)
        self.blocks = blocks
        # finalize (:120-127): offsets in node-id order
        order = np.argsort(tr.treeId)
        self.projOffset = np.zeros(N, dtype=np.int64)
        acc = 0
        for p in order:
            self.projOffset[p] = acc
            acc += int(self.k[p])
        self.p = acc
        return self

    @staticmethod
    def apply_projection(i, upd, made, blk, mbar, rows, blocks):
        """Fold what project_node returned for node i into the running
        contribution table: mid updates, the node's new blocks towards its
        ancestors, its reduced-M blocks; node i's own entries are consumed."""
        for key, delta in upd:
            mbar[key] += delta
        for a, arr in made:
            mbar[(a, i)] = arr
            rows.setdefault(a, []).append(i)
        for kk, B in blk:
            blocks.append((i, kk, B))
        for kk in rows.get(i, []):
            del mbar[(i, kk)]
        rows.pop(i, None)

    def assemble(self):
        """assembleMprojDense / assembleKprojDiag (solver.cpp:72-96)."""
        P = self.p
        Mp = np.eye(P)
        for (r, c, B) in self.blocks:
            r0, c0 = self.projOffset[r], self.projOffset[c]
            Mp[r0:r0 + B.shape[0], c0:c0 + B.shape[1]] = B
            Mp[c0:c0 + B.shape[1], r0:r0 + B.shape[0]] = B.T
        Kp = np.zeros((P, P))
        for p in range(len(self.k)):
            o = self.projOffset[p]
            for j in range(int(self.k[p])):
                Kp[o + j, o + j] = self.mu[p][j]
        return Kp, Mp

    def solve_projected(self, want: int, vectors: bool = True):
        """solveProjected (solver.cpp:145-161)."""
        if self.p <= 0:
            raise RuntimeError("reduced problem is empty")
        Kp, Mp = self.assemble()
        want = min(want, self.p)
        if want < self.p:
            return sla.eigh(Kp, Mp, subset_by_index=[0, want - 1], driver="gvx",
                            eigvals_only=not vectors)
        return sla.eigh(Kp, Mp, driver="gvd", eigvals_only=not vectors)

    def backtransform(self, Q: np.ndarray) -> np.ndarray:
        """backtransform (:263-319): reverse post-order z_i = Z_i q_i -
        Khat^-1 sum_a K_{a,i}^T z_a."""
        ge, tr = self.ge, self.ge.tree
        m = Q.shape[1]
        z = [None] * tr.N
        for i in reversed(range(tr.N)):
            ki = int(self.k[i])
            ni = int(tr.size[i])
            o = self.projOffset[i]
            yi = self.Z[i] @ Q[o:o + ki] if ki > 0 else np.zeros((ni, m))
            rhs = np.zeros((ni, m))
            for a in ge.anc[i]:
                rhs += ge.Kc[(a, i)].T @ z[a]
            z[i] = yi - ge.solve(i, rhs)
        out = np.zeros((int(tr.offset[-1] + tr.size[-1]), m))
        for i in range(tr.N):
            o, ni = int(tr.offset[i]), int(tr.size[i])
            out[o:o + ni] = z[i]
        return out


def project_node(fac, Mhat_i: np.ndarray, Z_i: np.ndarray, anc: List[int], Kc_i: List[np.ndarray],
                 Mc_i: List[np.ndarray], rowKK: List[int], mbar_i: List[np.ndarray],
                 plant: Optional[str] = None):
    """One node of the post-order projection (subspace_condensation.cpp
    :129-261): the mid updates M_(a, kk) -= K_(a, i) Khat_i^-1 M_(i, kk) it
    sends up for every earlier node kk in its row (as ((a, kk), delta)), its
    own blocks M_(a, i) = M_(a, i) Z_i - K_(a, i) Khat_i^-1 M-hat_i Z_i, and
    the reduced blocks Z_i^T M_(i, kk).  It reads only node i's own blocks
    and contributions, so the nodes of one tree level can run concurrently."""
    solve = lambda B: sla.cho_solve(fac, B, check_finite=False)
    ki = Z_i.shape[1]
    upd, made, blk = [], [], []
    if anc:
        if rowKK and plant != "drop_mid":
            Vs = solve(np.hstack(mbar_i))
            off = 0
            for kk, mb in zip(rowKK, mbar_i):
                kc = mb.shape[1]
                for a, Ka in zip(anc, Kc_i):
                    upd.append(((a, kk), -(Ka @ Vs[:, off:off + kc])))
                off += kc
        if ki > 0:
            WMz = solve(Mhat_i @ Z_i)
            for a, Ka, Ma in zip(anc, Kc_i, Mc_i):
                made.append((a, Ma @ Z_i - Ka @ WMz))
    if ki > 0:
        for kk, mb in zip(rowKK, mbar_i):
            blk.append((kk, Z_i.T @ mb))
    return upd, made, blk


def amls_solve(K: sp.spmatrix, M: sp.spmatrix, tree: Tree, *, nev: int = 100,
               thresholdFactor: float = 2.0, fixedKeep: int = 0, shift: float = 0.0,
               twoPass: bool = True, recoverVectors: bool = True,
               pinned_k: Optional[np.ndarray] = None, sil_mode: int = 0) -> OracleResult:
    """AmlsSolver::solve (src/amls/solver.cpp:100-319) on the given tree.
    sil_mode as in Condensation.run (0: the reference's density switch)."""
    t0 = time.perf_counter()
    if nev <= 0:
        raise ValueError("nev must be positive")
    n = K.shape[0]
    if n == 0:
        raise ValueError("empty matrix")
    nev = min(nev, n)
    Ks = sp.csr_matrix(K) + shift * sp.csr_matrix(M) if shift != 0.0 else sp.csr_matrix(K)
    perm = np.asarray(tree.perm)
    Kr = Ks[perm][:, perm]
    Mr = sp.csr_matrix(M)[perm][:, perm]
    t1 = time.perf_counter()
    ge = BlockGE(Kr, Mr, tree).run()
    t2 = time.perf_counter()
    cond = Condensation(ge)
    lam_cut = -1.0
    p0 = 0
    if pinned_k is not None:
        cond.run("fixed", 0, pinned_k=pinned_k)
    elif fixedKeep > 0:
        cond.run("fixed", fixedKeep, sil_mode=sil_mode)
    else:
        if not twoPass:
            raise ValueError("twoPass=False requires fixedKeep > 0")
        N = tree.N
        k0 = max(4, (3 * nev + N - 1) // N)
        cond.run("fixed", k0, sil_mode=sil_mode)
        p0 = cond.p
        cond.pathA = list(cond.path)
        wA = cond.solve_projected(min(nev, cond.p), vectors=False)
        if wA.size < nev:
            raise RuntimeError("pass A reduced problem smaller than nev")
        lam_cut = thresholdFactor * float(wA[nev - 1])
        cond.run("threshold", lam_cut, sil_mode=sil_mode)
    t3 = time.perf_counter()
    want = min(nev, cond.p)
    res = cond.solve_projected(want, vectors=recoverVectors)
    t4 = time.perf_counter()
    Z = None
    if recoverVectors:
        theta, Q = res
        z = cond.backtransform(Q)
        X = np.zeros_like(z)
        X[perm] = z
        MX = sp.csr_matrix(M) @ X
        nrm = np.einsum("ij,ij->j", X, MX)
        X /= np.sqrt(np.maximum(nrm, 1e-300))
        Z = X
    else:
        theta = res
    t5 = time.perf_counter()
    return OracleResult(theta=np.asarray(theta), Z=Z, p=cond.p, k=cond.k.copy(),
                        mu=cond.mu, Zloc=cond.Z, lambda_cut=lam_cut, p0=p0,
                        timing={"partition": t1 - t0, "gauss": t2 - t1, "condense": t3 - t2,
                                "reduced": t4 - t3, "backtransform": t5 - t4, "total": t5 - t0},
                        ge=ge, cond=cond)


# --------------------------------------------------------------------------
# Subspace iteration with the AMLS factors (AMLS-SIMb)
# --------------------------------------------------------------------------

def amls_kinv(ge: BlockGE, R: np.ndarray) -> np.ndarray:
    """K^-1 R in the permuted ordering through the block factors:
    K = L Khat L^T with L_{a,i} = K_{a,i} Khat_ii^-1 (the AMLS transformation
    U = L^-T of papers/Improving_eigenpairs_...md:384-427, eqs. 30-32)."""
    tr = ge.tree
    r = [R[int(tr.offset[p]):int(tr.offset[p] + tr.size[p])].copy() for p in range(tr.N)]
    w = [None] * tr.N
    for p in range(tr.N):                  # forward, post-order
        w[p] = ge.solve(p, r[p])
        for a in ge.anc[p]:
            r[a] -= ge.Kc[(a, p)] @ w[p]
    x = [None] * tr.N
    for p in reversed(range(tr.N)):        # backward, reverse post-order
        rhs = np.zeros_like(w[p])
        for a in ge.anc[p]:
            rhs += ge.Kc[(a, p)].T @ x[a]
        x[p] = w[p] - ge.solve(p, rhs)
    return np.vstack(x)


def subspace_iteration(K, M, ge: BlockGE, Q0: np.ndarray, steps: int, nev: int,
                       perm: np.ndarray, every_step: bool = False):
    """Q <- K^-1 M Q, Rayleigh-Ritz with the original K, M after every step
    (Algorithm 5 of the paper, with an RR step per iteration).  Q0 and the
    result are in the original ordering.  every_step: also return the list
    of the lowest nev Ritz values after each step."""
    Kr = sp.csr_matrix(K)[perm][:, perm]
    Mr = sp.csr_matrix(M)[perm][:, perm]
    Q = Q0[perm]
    theta = None
    hist = []
    for _ in range(steps):
        R = Mr @ Q
        Y = amls_kinv(ge, R)
        Kc = R.T @ Y
        Mc = Y.T @ (Mr @ Y)
        Kc = 0.5 * (Kc + Kc.T)
        Mc = 0.5 * (Mc + Mc.T)
        theta, V = sla.eigh(Kc, Mc)
        Q = Y @ V
        hist.append(theta[:nev].copy())
    X = np.zeros_like(Q)
    X[perm] = Q
    if every_step:
        return theta[:nev], X[:, :nev], hist
    return theta[:nev], X[:, :nev]


# --------------------------------------------------------------------------
# FP64 ground truth (src/reference/reference_solver.cpp:17-95)
# --------------------------------------------------------------------------

def reference_eigs(K, M, nev: int, dense_max: int = 5000, vectors: bool = False):
    """ReferenceSolver::smallestEigenvalues: dense generalized eigh for
    n <= dense_max, else shift-invert Lanczos at sigma = 0 with
    ncv = min(3 nev + 20, n) and tol 1e-12 (Spectra semantics, here ARPACK
    through scipy)."""
    from scipy.sparse.linalg import eigsh
    n = K.shape[0]
    nev = min(nev, n)
    if n <= dense_max:
        w, V = sla.eigh(K.toarray(), M.toarray(), subset_by_index=[0, nev - 1])
        return (w, V) if vectors else w
    w, V = eigsh(sp.csc_matrix(K), k=nev, M=sp.csc_matrix(M), sigma=0.0, which="LM",
                 ncv=min(3 * nev + 20, n), tol=1e-12)
    o = np.argsort(w)
    return (w[o], V[:, o]) if vectors else w[o]


def rel_err(theta: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """ErrorMetrics::eigenvalueRelErr (src/reference/error_metrics.cpp:7-18)."""
    m = min(theta.size, ref.size)
    d = np.abs(ref[:m])
    e = np.abs(theta[:m] - ref[:m])
    return np.where(d > 0, e / np.where(d > 0, d, 1), e)


def residual(K, M, theta: float, z: np.ndarray) -> float:
    """ErrorMetrics::residual (src/reference/error_metrics.cpp:29-37)."""
    Kz, Mz = K @ z, M @ z
    num = np.linalg.norm(Kz - theta * Mz)
    den = np.linalg.norm(Kz) + abs(theta) * np.linalg.norm(Mz)
    return num / den if den > 0 else 0.0


# --------------------------------------------------------------------------
# Eigenvalue adjoint (new; no counterpart in the reference)
# --------------------------------------------------------------------------

def eig_adjoint(conn: np.ndarray, Ke0: np.ndarray, Me0: np.ndarray, theta: np.ndarray,
                Z: np.ndarray, shift: float = 0.0):
    """d theta_j / d s_e = x_je^T Ke0 x_je, d theta_j / d r_e =
    -(theta_j - shift) x_je^T Me0 x_je for K = sum_e s_e P_e^T Ke0 P_e,
    M = sum_e r_e P_e^T Me0 P_e and M-normalised x_j (the first-order
    perturbation of a simple eigenvalue of (K + shift M, M),
    d theta = x^T (dK + shift dM - theta dM) x).  Returns (grad_s, grad_r),
    nelem x nev."""
    conn = np.asarray(conn)
    nev = len(theta)
    Zp = np.vstack([Z[:, :nev], np.zeros((1, nev))])      # row -1 -> zeros
    X = Zp[np.where(conn >= 0, conn, Z.shape[0])]           # nelem x npe x nev
    gs = np.einsum("eij,ik,ekj->ej", X, Ke0, X)
    lam = np.asarray(theta) - shift
    gr = -np.einsum("eij,ik,ekj->ej", X, Me0, X) * lam[None, :]
    return gs, gr


# --------------------------------------------------------------------------
# Operator interfaces: ColCompOps (src/linalg/colcomp.hpp:124-423) and the
# dense LDL^T of DenseLDLTSolver (src/linalg/sparse_direct.cpp:38-63)
# --------------------------------------------------------------------------

@dataclass
class ColComp:
    """ColComp (colcomp.hpp:26-75): data holds the stored columns of a
    logical rows x full block, cols (ascending) their logical numbers; cols
    None is the identity map (ident)."""
    data: np.ndarray
    full: int
    cols: Optional[np.ndarray] = None

    @property
    def ident(self) -> bool:
        return self.cols is None

    def stored(self) -> int:
        return self.full if self.ident else len(self.cols)

    def empty(self) -> bool:
        return self.stored() == 0

    def logical(self) -> np.ndarray:
        """The rows x full block the compressed storage stands for."""
        if self.ident:
            return self.data.copy()
        out = np.zeros((self.data.shape[0], self.full))
        out[:, self.cols] = self.data
        return out


class NumpyColCompOps:
    """OpenBlasColCompOps (colcomp.hpp:174-423) restated on numpy arrays.
    Operands: Dense = ndarray, ColComp, Sparse = scipy sparse.  The
    transposed terms restrict (B C)^T to the target's stored columns with
    S = B.data(A.cols, :) (the rows of B's stored block that the target
    keeps, all of B's stored columns); the reference's compressed-B branch
    of accMul (:396-397) writes sub(B.data, A.cols, B.cols), which reads
    B.data's columns at logical numbers."""

    @staticmethod
    def _rows(X, idx):
        return X[np.asarray(idx)]

    def gemm(self, ta: bool, tb: bool, alpha, A, B, beta, C):
        if isinstance(A, ColComp):
            if not A.ident and len(A.cols) == 0:
                return beta * C
            if A.ident:
                A = A.data
            elif not ta:
                Bg = self._rows(B, A.cols)
                return alpha * (A.data @ (Bg.T if tb else Bg)) + beta * C
            else:
                A = A.data          # stored-row-space result (:132-137)
        opA = A.T if ta else A
        opB = B.T if tb else B
        return alpha * (opA @ opB) + beta * C

    def subMul(self, A, B, C):
        """A -= B C; returns the updated A (ndarray or ColComp)."""
        return self._acc(A, B, C, -1.0, False)

    def addMul(self, A, B, C):
        return self._acc(A, B, C, +1.0, False)

    def subTransMul(self, A, B, C):
        return self._acc(A, B, C, -1.0, True)

    def addTransMul(self, A, B, C):
        return self._acc(A, B, C, +1.0, True)

    def _acc(self, A, B, C, sign, trans):
        if sp.issparse(B):
            B = ColComp(np.asarray(B.todense()), B.shape[1])
        if B.empty():
            return A
        if isinstance(A, ColComp) and A.empty():
            return A
        dense_target = not isinstance(A, ColComp)
        data = A if dense_target else A.data
        acols = None if dense_target or A.ident else A.cols
        if not trans:
            Cg = C if B.ident else self._rows(C, B.cols)
            if acols is not None:
                Cg = Cg[:, acols]
            data = data + sign * (B.data @ Cg)
        else:
            Cg = C if B.ident else self._rows(C, B.cols)
            S = B.data if acols is None else B.data[acols]
            data = data + sign * (Cg.T @ S.T)
        if dense_target:
            return data
        return ColComp(data, A.full, A.cols)

    def mul(self, B, C):
        if sp.issparse(B):
            return np.asarray(B @ C)
        return self.gemm(False, False, 1.0, B, C, 0.0, np.zeros((B.data.shape[0], C.shape[1])))

    def expandT(self, B):
        if sp.issparse(B):
            return np.asarray(B.T.todense())
        return B.logical().T


def eigen_ldlt(A: np.ndarray):
    """Eigen's LDLT<Lower> as DenseLDLTSolver uses it (Eigen 3, unblocked
    left-looking factorisation with diagonal pivoting): step k takes the
    first largest |diagonal| of positions k..n-1 of the working matrix,
    swaps it into k (a symmetric permutation of the lower-stored matrix),
    then forms column k from the columns before it.  Returns perm
    (perm[new] = old), unit lower L and D with A[perm][:, perm] = L D L^T."""
    n = A.shape[0]
    mat = np.tril(np.asarray(A, dtype=np.float64)).copy()
    perm = np.arange(n)
    for k in range(n):
        idx = k + int(np.argmax(np.abs(np.diag(mat)[k:])))
        if idx != k:
            mat[[k, idx], :k] = mat[[idx, k], :k]
            tail = mat[idx + 1:, k].copy()
            mat[idx + 1:, k] = mat[idx + 1:, idx]
            mat[idx + 1:, idx] = tail
            mat[k, k], mat[idx, idx] = mat[idx, idx], mat[k, k]
            for i in range(k + 1, idx):
                mat[i, k], mat[idx, i] = mat[idx, i], mat[i, k]
            perm[k], perm[idx] = perm[idx], perm[k]
        if k > 0:
            temp = np.diag(mat)[:k] * mat[k, :k]
            mat[k, k] -= mat[k, :k] @ temp
            mat[k + 1:, k] -= mat[k + 1:, :k] @ temp
        if k + 1 < n and mat[k, k] != 0.0:
            mat[k + 1:, k] /= mat[k, k]
    return perm, np.tril(mat, -1) + np.eye(n), np.diag(mat).copy()


def diagonal_selection_order(diag: np.ndarray) -> np.ndarray:
    """The pivot order eigen_ldlt ends with, from the input diagonal alone:
    the first largest |d| of positions k..n-1 swapped into k, k = 0..n-1."""
    a = np.abs(np.asarray(diag, dtype=np.float64)).copy()
    pos = np.arange(a.size)
    for k in range(a.size):
        j = k + int(np.argmax(a[k:]))
        a[[k, j]] = a[[j, k]]
        pos[[k, j]] = pos[[j, k]]
    return pos
