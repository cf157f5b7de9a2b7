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

Deviations, all value-preserving: the sparse Eigen LDL^T of leaves and the
dense Eigen LDLT of separators are both done as a dense Cholesky (same
factorisation of an SPD matrix, different rounding); the SIL local solver
(src/amls/shift_invert_lanczos.hpp) is replaced by the dense solver it
approximates (it is only selected for nodes with density < 0.15, which the
hexahedral workloads do not have).

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
        for p in range(N):
            o, n = int(off[p]), int(size[p])
            Kcol = K[:, o:o + n]
            Mcol = M[:, o:o + n]
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
            Kh = self.Kd[p]
            self.Khat[p] = Kh.copy()
            self.Mhat[p] = self.Md[p].copy()
            try:
                c, low = sla.cho_factor(Kh, lower=True, check_finite=False)
            except np.linalg.LinAlgError as e:   # D > 0 check (sparse_direct.cpp:13-22)
                raise np.linalg.LinAlgError(f"node {int(tr.treeId[p])}: not SPD") from e
            d = np.diag(c)
            if not np.all(np.isfinite(d)) or np.any(d <= 0):
                raise np.linalg.LinAlgError(f"node {int(tr.treeId[p])}: not SPD")
            self.fac[p] = (c, low)
            anc = self.anc[p]
            if not anc:
                continue
            n_p = int(tr.size[p])
            W = self.Mhat[p]
            X, Y, WX, rK, rM = {}, {}, {}, {}, {}
            for a in anc:                               # solveXYAnc_ (:510-598)
                Ka, Ma = self.Kc[(a, p)], self.Mc[(a, p)]
                na = Ka.shape[0]
                rK[a], rM[a] = _nzrows(Ka), _nzrows(Ma)
                X[a] = np.zeros((n_p, na))
                Y[a] = np.zeros((n_p, na))
                if rK[a].size:
                    X[a][:, rK[a]] = self.solve(p, Ka[rK[a]].T)
                if rM[a].size:
                    Y[a][:, rM[a]] = self.solve(p, Ma[rM[a]].T)
                WX[a] = np.zeros((n_p, na))             # makeWX_ (:653-676)
                if rK[a].size:
                    WX[a][:, rK[a]] = self.solve(p, W @ X[a][:, rK[a]])
            ix = np.ix_
            for a in anc:                               # schurKDiagTask_ (:678-693)
                r = rK[a]
                if r.size:
                    upd = self.Kc[(a, p)][r] @ X[a][:, r]
                    self.Kd[a][ix(r, r)] += upd if plant == "schur_sign" else -upd
            for u in range(len(anc)):                   # schurKCoupTask_ (:695-715)
                for v in range(u + 1, len(anc)):
                    lo, hi = anc[u], anc[v]
                    rh, rl = rK[hi], rK[lo]
                    if rh.size and rl.size:
                        self.Kc[(hi, lo)][ix(rh, rl)] -= self.Kc[(hi, p)][rh] @ X[lo][:, rl]
            for a in anc:                               # congDiagTask_ (:717-749)
                B = self.Kc[(a, p)]
                r, rm = rK[a], rM[a]
                if r.size and rm.size:
                    T1 = B[r] @ Y[a][:, rm]
                    self.Md[a][ix(r, rm)] -= T1
                    self.Md[a][ix(rm, r)] -= T1.T
                if r.size and plant != "drop_wx":
                    self.Md[a][ix(r, r)] += B[r] @ WX[a][:, r]
            for u in range(len(anc)):                   # congCoupTask_ (:751-781)
                for v in range(u + 1, len(anc)):
                    lo, hi = anc[u], anc[v]
                    rh, rl = rK[hi], rK[lo]
                    rmh, rml = rM[hi], rM[lo]
                    tgt = self.Mc[(hi, lo)]
                    if rh.size and rml.size:
                        tgt[ix(rh, rml)] -= self.Kc[(hi, p)][rh] @ Y[lo][:, rml]
                    if rl.size and rmh.size:
                        T2 = self.Kc[(lo, p)][rl] @ Y[hi][:, rmh]
                        if plant != "coup_transpose":
                            tgt[ix(rmh, rl)] -= T2.T
                        elif T2.shape[0] == T2.shape[1]:
                            tgt[ix(rmh, rl)] -= T2
                    if rh.size and rl.size:
                        tgt[ix(rh, rl)] += self.Kc[(hi, p)][rh] @ WX[lo][:, rl]
        return self


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


class Condensation:
    """SubspaceCondensation restated (src/amls/subspace_condensation.cpp)."""

    def __init__(self, ge: BlockGE):
        self.ge = ge

    def run(self, kind: str, value: float, pinned_k: Optional[np.ndarray] = None,
            plant: Optional[str] = None):
        """plant: "drop_mid" skips the mid update M_a,kk -= K_a,i Khat^-1 M_i,kk
        (a deliberate mistake for the tests of the parity checks)."""
        ge, tr = self.ge, self.ge.tree
        N = tr.N
        self.k = np.zeros(N, dtype=np.int64)
        self.mu: List[np.ndarray] = [np.zeros(0)] * N
        self.Z: List[np.ndarray] = [np.zeros((int(tr.size[p]), 0)) for p in range(N)]
        # spectralLevel (:45-100): local spectra + select
        for p in range(N):
            if pinned_k is not None:
                w, V = local_eig(ge.Khat[p], ge.Mhat[p], "fixed", max(int(pinned_k[p]), 0)) \
                    if pinned_k[p] > 0 else (np.zeros(0), np.zeros((int(tr.size[p]), 0)))
            else:
                w, V = local_eig(ge.Khat[p], ge.Mhat[p], kind, value)
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
            ki = int(self.k[i])
            rowKK = rows.get(i, [])
            if anc:
                if rowKK and plant != "drop_mid":
                    RHS = np.hstack([mbar[(i, kk)] for kk in rowKK])
                    Vs = ge.solve(i, RHS)
                    off = 0
                    for kk in rowKK:
                        kc = mbar[(i, kk)].shape[1]
                        for a in anc:
                            mbar[(a, kk)] -= ge.Kc[(a, i)] @ Vs[:, off:off + kc]
                        off += kc
                if ki > 0:
                    Zi = self.Z[i]
                    WMz = ge.solve(i, ge.Mhat[i] @ Zi)
                    for a in anc:
                        mbar[(a, i)] = ge.Mc[(a, i)] @ Zi - ge.Kc[(a, i)] @ WMz
                        rows.setdefault(a, []).append(i)
            if ki > 0:
                for kk in rowKK:
                    blocks.append((i, kk, self.Z[i].T @ mbar[(i, kk)]))
            for kk in rowKK:
                del mbar[(i, kk)]
            rows.pop(i, None)
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


def amls_solve(K: sp.spmatrix, M: sp.spmatrix, tree: Tree, *, nev: int = 100,
               thresholdFactor: float = 2.0, fixedKeep: int = 0, shift: float = 0.0,
               twoPass: bool = True, recoverVectors: bool = True,
               pinned_k: Optional[np.ndarray] = None) -> OracleResult:
    """AmlsSolver::solve (src/amls/solver.cpp:100-319) on the given tree."""
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
        cond.run("fixed", fixedKeep)
    else:
        if not twoPass:
            raise ValueError("twoPass=False requires fixedKeep > 0")
        N = tree.N
        k0 = max(4, (3 * nev + N - 1) // N)
        cond.run("fixed", k0)
        p0 = cond.p
        wA = cond.solve_projected(min(nev, cond.p), vectors=False)
        if wA.size < nev:
            raise RuntimeError("pass A reduced problem smaller than nev")
        lam_cut = thresholdFactor * float(wA[nev - 1])
        cond.run("threshold", lam_cut)
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
                       perm: np.ndarray):
    """Q <- K^-1 M Q, Rayleigh-Ritz with the original K, M after every step
    (Algorithm 5 of the paper, with an RR step per iteration).  Q0 and the
    result are in the original ordering."""
    Kr = sp.csr_matrix(K)[perm][:, perm]
    Mr = sp.csr_matrix(M)[perm][:, perm]
    Q = Q0[perm]
    theta = None
    for _ in range(steps):
        R = Mr @ Q
        Y = amls_kinv(ge, R)
        Kc = R.T @ Y
        Mc = Y.T @ (Mr @ Y)
        Kc = 0.5 * (Kc + Kc.T)
        Mc = 0.5 * (Mc + Mc.T)
        theta, V = sla.eigh(Kc, Mc)
        Q = Y @ V
    X = np.zeros_like(Q)
    X[perm] = Q
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
                Z: np.ndarray):
    """d theta_j / d s_e = x_je^T Ke0 x_je, d theta_j / d r_e =
    -theta_j x_je^T Me0 x_je for K = sum_e s_e P_e^T Ke0 P_e,
    M = sum_e r_e P_e^T Me0 P_e and M-normalised x_j (the first-order
    perturbation of a simple eigenvalue, d theta = x^T (dK - theta dM) x).
    Returns (grad_s, grad_r), nelem x nev."""
    conn = np.asarray(conn)
    nev = len(theta)
    Zp = np.vstack([Z[:, :nev], np.zeros((1, nev))])      # row -1 -> zeros
    X = Zp[np.where(conn >= 0, conn, Z.shape[0])]           # nelem x npe x nev
    gs = np.einsum("eij,ik,ekj->ej", X, Ke0, X)
    gr = -np.einsum("eij,ik,ekj->ej", X, Me0, X) * np.asarray(theta)[None, :]
    return gs, gr
