"""Node-level parallel run of the oracle's AMLS -- test infrastructure, not
product code (the CPU baseline of bench.py; tests/ check it against the
serial oracle).

The reference runs AMLS on all host cores with BLAS pinned to one thread and
its parallelism at the level of tree nodes (src/amls/solver.cpp:136-143,
199-212; src/amls/thread_pool.hpp:16-43; src/core/dag.hpp:38-152), adding
the updates of concurrent nodes into shared ancestor blocks under locks
(src/amls/gaussian_elimination.hpp:267-270); its fused path runs the whole
pipeline -- elimination, local spectra, projection -- once per truncation
pass (solver.cpp:162-212, 237-267).  This module does the same with the
oracle's per-node functions (ge_node, local_eig_dispatch, project_node) in a
pool of forked worker processes with single-threaded BLAS (scipy's LAPACK
wrappers hold the interpreter lock, so threads would not run them
concurrently).  Every block lives in one shared anonymous mapping created
before the pool, so a task carries only its node number:
  - elimination: tree levels from the deepest up, the nodes of a level
    concurrently; a node adds its updates into its ancestors' blocks under
    one lock per ancestor (the same sums as the serial post-order, in
    another order);
  - local spectra: all nodes concurrently;
  - projection: levels from the deepest up; the contributions a node writes
    belong to it alone, so no lock is needed.
The reduced problems and the back-transformation run in the parent.
"""
from __future__ import annotations

import mmap
import multiprocessing as mp
import os
import time
from concurrent.futures import ProcessPoolExecutor
from typing import Dict, List

import numpy as np
import scipy.sparse as sp

from . import amls_oracle as O

_S = None   # the shared state; set in the parent before the pool forks


def host_workers(cap: int = 16) -> int:
    """Worker count: the CPUs this process may run on, at most cap."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


def levels_of(tree: O.Tree) -> List[List[int]]:
    """Post-order nodes grouped by depth, deepest level first."""
    depth = np.zeros(tree.N, dtype=np.int64)
    for p in reversed(range(tree.N)):           # parents come after children
        q = int(tree.parent[p])
        depth[p] = 0 if q < 0 else depth[q] + 1
    return [[p for p in range(tree.N) if depth[p] == d] for d in range(int(depth.max()), -1, -1)]


class _Shared:
    """FP64 arrays carved out of one anonymous shared mapping."""

    def __init__(self, shapes: List[tuple]):
        total = sum(int(np.prod(s)) for s in shapes) * 8
        self.buf = mmap.mmap(-1, max(total, 8))
        self.off = 0

    def take(self, shape) -> np.ndarray:
        cnt = int(np.prod(shape))
        a = np.frombuffer(self.buf, dtype=np.float64, count=cnt, offset=self.off).reshape(shape)
        self.off += cnt * 8
        return a


class _State:
    def __init__(self, ge: O.BlockGE, locks):
        tr = ge.tree
        self.tree, self.anc, self.locks = tr, ge.anc, locks
        size = [int(s) for s in tr.size]
        pairs = list(ge.Kc.keys())
        shapes = [(n, n) for n in size] * 4                          # Kd, Md, fac, Z
        shapes += [ge.Kc[k].shape for k in pairs] * 3                # Kc, Mc, contributions
        sh = _Shared(shapes)
        self.init = (ge.Kd, ge.Md, ge.Kc, ge.Mc)
        self.Kd = [sh.take((n, n)) for n in size]
        self.Md = [sh.take((n, n)) for n in size]
        self.fac = [sh.take((n, n)) for n in size]
        self.Z = [sh.take((n, n)) for n in size]
        self.Kc = {k: sh.take(ge.Kc[k].shape) for k in pairs}
        self.Mc = {k: sh.take(ge.Kc[k].shape) for k in pairs}
        self.mbar = {k: sh.take(ge.Kc[k].shape) for k in pairs}
        self.k = np.zeros(tr.N, dtype=np.int64)
        self.ge = ge
        self.reset()

    def reset(self):
        """Blocks back to the permuted K and M (the start of a pass)."""
        Kd, Md, Kc, Mc = self.init
        for p in range(self.tree.N):
            self.Kd[p][...] = Kd[p]
            self.Md[p][...] = Md[p]
        for k in self.Kc:
            self.Kc[k][...] = Kc[k]
            self.Mc[k][...] = Mc[k]

    def view(self) -> O.BlockGE:
        """The oracle's BlockGE over the shared blocks (for the parent's
        reduced solve, back-transformation and subspace iteration)."""
        ge = self.ge
        ge.Kd, ge.Md, ge.Kc, ge.Mc = self.Kd, self.Md, self.Kc, self.Mc
        ge.Khat, ge.Mhat = self.Kd, self.Md
        ge.fac = [(self.fac[p], True) for p in range(self.tree.N)]
        return ge


def _init_worker():
    from threadpoolctl import threadpool_limits
    global _LIMITS
    _LIMITS = threadpool_limits(1)


def _ge_task(p: int) -> int:
    S = _S
    anc = S.anc[p]
    fac, deltas = O.ge_node(S.Kd[p], S.Md[p], anc, [S.Kc[(a, p)] for a in anc],
                            [S.Mc[(a, p)] for a in anc], int(S.tree.treeId[p]))
    S.fac[p][...] = fac[0]
    tgt = {"Kd": S.Kd, "Md": S.Md, "Kc": S.Kc, "Mc": S.Mc}
    by_owner: Dict[int, list] = {}
    for d in deltas:
        owner = d[1] if d[0] in ("Kd", "Md") else d[1][0]
        by_owner.setdefault(owner, []).append(d)
    for owner in sorted(by_owner):
        with S.locks[owner]:
            for kind, key, (ri, ci), delta in by_owner[owner]:
                tgt[kind][key][np.ix_(ri, ci)] += delta
    return p


def _spec_task(p: int, kind: str, value: float, sil_mode: int):
    S = _S
    w, V, path = O.local_eig_dispatch(S.Kd[p], S.Md[p], kind, value, sil_mode)
    if kind == "threshold":
        kk = int(np.searchsorted(w, value, side="right"))
    else:
        kk = w.size
    S.Z[p][:, :kk] = V[:, :kk]
    return w[:kk].copy(), kk, path


def _proj_task(i: int, rowKK: List[int], k: List[int]):
    S = _S
    anc = S.anc[i]
    ki = k[i]
    upd, made, blk = O.project_node((S.fac[i], True), S.Md[i], S.Z[i][:, :ki], anc,
                                    [S.Kc[(a, i)] for a in anc], [S.Mc[(a, i)] for a in anc],
                                    rowKK, [S.mbar[(i, kk)][:, :k[kk]] for kk in rowKK])
    for (a, kk), delta in upd:
        S.mbar[(a, kk)][:, :k[kk]] += delta
    for a, arr in made:
        S.mbar[(a, i)][:, :ki] = arr
    return [a for a, _ in made], blk


class ParallelAmls:
    """AmlsSolver::solve's fused two-pass pipeline on a process pool."""

    def __init__(self, K, M, tree: O.Tree, workers: int):
        global _S
        self.tree, self.workers = tree, workers
        perm = np.asarray(tree.perm)
        self.Kr = sp.csr_matrix(K)[perm][:, perm]
        self.Mr = sp.csr_matrix(M)[perm][:, perm]
        ge = O.BlockGE(self.Kr, self.Mr, tree)
        ctx = mp.get_context("fork")
        _S = self.S = _State(ge, [ctx.Lock() for _ in range(tree.N)])
        self.pool = ProcessPoolExecutor(max_workers=workers, mp_context=ctx,
                                        initializer=_init_worker)

    def close(self):
        self.pool.shutdown()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def ge(self) -> O.BlockGE:
        self.S.reset()
        for level in levels_of(self.tree):
            list(self.pool.map(_ge_task, level))
        return self.S.view()

    def condense(self, kind: str, value: float, sil_mode: int = 0) -> O.Condensation:
        S, tr = self.S, self.tree
        N = tr.N
        cond = O.Condensation(S.view())
        cond.k = np.zeros(N, dtype=np.int64)
        cond.mu, cond.Z, cond.path = [None] * N, [None] * N, [None] * N
        futs = [self.pool.submit(_spec_task, p, kind, value, sil_mode) for p in range(N)]
        for p, f in enumerate(futs):
            w, kk, cond.path[p] = f.result()
            cond.k[p] = kk
            cond.mu[p] = w
            cond.Z[p] = S.Z[p][:, :kk]
        k = [int(v) for v in cond.k]
        rows: Dict[int, List[int]] = {}
        blocks = []
        for level in levels_of(tr):
            futs = [self.pool.submit(_proj_task, i, list(rows.get(i, [])), k) for i in level]
            for i, f in zip(level, futs):
                made, blk = f.result()
                for a in made:
                    rows.setdefault(a, []).append(i)
                for kk, B in blk:
                    blocks.append((i, kk, B))
                rows.pop(i, None)
        cond.blocks = blocks
        order = np.argsort(tr.treeId)
        cond.projOffset = np.zeros(N, dtype=np.int64)
        acc = 0
        for p in order:
            cond.projOffset[p] = acc
            acc += int(cond.k[p])
        cond.p = acc
        return cond

    def solve(self, *, nev: int, thresholdFactor: float = 2.0, recoverVectors: bool = True,
              sil_mode: int = 0) -> O.OracleResult:
        """Two-pass AMLS as the reference's fused path runs it: pass A
        (elimination, FixedCount(k0) spectra, projection, reduced solve for
        lambda-hat), pass B (the whole pipeline again with the threshold
        eta lambda-hat), the reduced solve, back-transformation and
        M-normalisation (src/amls/solver.cpp:162-311)."""
        tree = self.tree
        t0 = time.perf_counter()
        k0 = max(4, (3 * nev + tree.N - 1) // tree.N)
        self.ge()
        cond = self.condense("fixed", k0, sil_mode)
        p0 = cond.p
        wA = cond.solve_projected(min(nev, cond.p), vectors=False)
        if wA.size < nev:
            raise RuntimeError("pass A reduced problem smaller than nev")
        lam_cut = thresholdFactor * float(wA[nev - 1])
        t1 = time.perf_counter()
        ge = self.ge()
        cond = self.condense("threshold", lam_cut, sil_mode)
        t2 = time.perf_counter()
        res = cond.solve_projected(min(nev, cond.p), vectors=recoverVectors)
        t3 = time.perf_counter()
        Z = None
        if recoverVectors:
            theta, Q = res
            z = cond.backtransform(Q)
            perm = np.asarray(tree.perm)
            X = np.zeros_like(z)
            X[perm] = z
            Mo = self.Mr
            nrm = np.einsum("ij,ij->j", z, Mo @ z)
            Z = X / np.sqrt(np.maximum(nrm, 1e-300))
        else:
            theta = res
        t4 = time.perf_counter()
        return O.OracleResult(theta=np.asarray(theta), Z=Z, p=cond.p, k=cond.k.copy(), mu=cond.mu,
                              Zloc=cond.Z, lambda_cut=lam_cut, p0=p0,
                              timing={"passA": t1 - t0, "passB": t2 - t1, "reduced": t3 - t2,
                                      "backtransform": t4 - t3, "total": t4 - t0},
                              ge=ge, cond=cond)


# ---------------------------------------------------------------------------
# Subspace iteration (AMLS-SIMb) on the same kind of pool
# ---------------------------------------------------------------------------

_SI = None   # shared blocks of the subspace iteration; set before its pool forks


def _si_spmm(which: str, src: str, dst: str, c0: int, c1: int):
    S = _SI
    A = S.Kr if which == "K" else S.Mr
    getattr(S, dst)[:, c0:c1] = A @ getattr(S, src)[:, c0:c1]


def _si_kinv(c0: int, c1: int):
    S = _SI
    S.Y[:, c0:c1] = O.amls_kinv(S.ge, np.asfortranarray(S.R[:, c0:c1]))


def _si_gram(a: str, b: str, out: str, c0: int, c1: int):
    S = _SI
    getattr(S, out)[:, c0:c1] = getattr(S, a).T @ getattr(S, b)[:, c0:c1]


def _si_rotate(c0: int, c1: int):
    S = _SI
    S.Q[:, c0:c1] = S.Y @ S.V[:, c0:c1]


class ParallelSubspaceIteration:
    """O.subspace_iteration (Rayleigh-Ritz after every step) with every
    product split by columns over a pool of forked workers (one BLAS thread
    each): M Q, K^-1 through the AMLS factors, the two projections, and the
    rotation Y V; the small dense eigenproblem runs in the parent."""

    def __init__(self, ge: O.BlockGE, Kr, Mr, q: int, workers: int):
        global _SI
        n = Kr.shape[0]
        self.n, self.q, self.ge, self.Kr, self.Mr = n, q, ge, Kr, Mr
        sh = _Shared([(n, q)] * 4 + [(q, q)] * 3)

        def big():
            return sh.take((q, n)).T            # column-major n x q
        self.Q, self.R, self.Y, self.T = big(), big(), big(), big()
        self.Kc, self.Mc, self.V = (sh.take((q, q)).T for _ in range(3))
        self.chunks = [(int(c0), int(min(q, c0 + -(-q // workers))))
                       for c0 in range(0, q, -(-q // workers))]
        _SI = self
        self.pool = ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork"),
                                        initializer=_init_worker)

    def close(self):
        self.pool.shutdown()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _each(self, f, *args):
        futs = [self.pool.submit(f, *args, c0, c1) for c0, c1 in self.chunks]
        for fu in futs:
            fu.result()

    def run(self, Q0: np.ndarray, steps: int, nev: int, perm: np.ndarray):
        import scipy.linalg as sla
        self.Q[...] = Q0[perm]
        theta = None
        for _ in range(steps):
            self._each(_si_spmm, "M", "Q", "R")
            self._each(_si_kinv)
            self._each(_si_gram, "R", "Y", "Kc")
            self._each(_si_spmm, "M", "Y", "T")
            self._each(_si_gram, "Y", "T", "Mc")
            Kc = 0.5 * (self.Kc + self.Kc.T)
            Mc = 0.5 * (self.Mc + self.Mc.T)
            theta, V = sla.eigh(Kc, Mc)
            self.V[...] = V
            self._each(_si_rotate)
        X = np.zeros((self.n, nev))
        X[perm] = self.Q[:, :nev]
        return theta[:nev], X
