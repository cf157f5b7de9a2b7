"""amls_amd: the AMLS eigensolver hot path on MI355X (gfx950).

Python mirror of the reference's C++ API (src/amls/solver.hpp:18-52):

    from amls_amd import AmlsSolver, AmlsOptions, Pencil
    sol = AmlsSolver().solve(Pencil(K, M), AmlsOptions(nev=100))
    sol.theta, sol.Z, sol.p, sol.timing

All numerics run in libamls_hip.so through its C-ABI (include/amls_hip.h);
this module only converts arrays.  `HipContext` exposes the per-phase entry
points used by the parity tests.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import AmlsError, AmlsOptionsC, dptr, iptr

__all__ = ["AmlsOptions", "AmlsSolver", "Solution", "Pencil", "HipContext", "AmlsError",
           "DenseFactor", "ContributionStore"]

PRECISIONS = {"fp64": 0, "mixed": 1}


@dataclass
class AmlsOptions:
    """AmlsOptions of src/amls/solver.hpp:18-30 (same names and defaults),
    plus the GPU forms: precision 'fp64' | 'mixed', siSteps, siVectors, and
    silMode (EigenpairSolver::silMode, src/amls/eigenpair_solver.hpp:28-29:
    0 density switch, 1 always shift-invert Lanczos, -1 always dense)."""
    nLevels: int = 6
    leafSize: int = 500
    thresholdFactor: float = 2.0
    fixedKeep: int = 0
    nev: int = 100
    denseProjectedMax: int = 5000
    shift: float = 0.0
    twoPass: bool = True
    numThreads: int = 0
    recoverVectors: bool = True
    verbosity: int = 1
    precision: str = "fp64"
    siSteps: int = 0
    siVectors: int = 0
    silMode: int = 0

    def to_c(self) -> AmlsOptionsC:
        o = AmlsOptionsC()
        o.nLevels = self.nLevels
        o.leafSize = self.leafSize
        o.thresholdFactor = self.thresholdFactor
        o.fixedKeep = self.fixedKeep
        o.nev = self.nev
        o.denseProjectedMax = self.denseProjectedMax
        o.shift = self.shift
        o.twoPass = int(bool(self.twoPass))
        o.numThreads = self.numThreads
        o.recoverVectors = int(bool(self.recoverVectors))
        o.verbosity = self.verbosity
        o.precision = PRECISIONS[self.precision]
        o.siSteps = self.siSteps
        o.siVectors = self.siVectors
        o.silMode = self.silMode
        return o


@dataclass
class Solution:
    """Solution of src/amls/solver.hpp:38-43."""
    theta: np.ndarray
    Z: Optional[np.ndarray]
    p: int
    timing: Dict[str, float] = field(default_factory=dict)


@dataclass
class Pencil:
    """Pencil{K, M} (src/core/types.hpp:24-28); coords: optional (n, 3) dof
    coordinates that let the partitioner cut along node planes."""
    K: sp.spmatrix
    M: sp.spmatrix
    coords: Optional[np.ndarray] = None


def _csc(A):
    A = A if isinstance(A, sp.csc_matrix) else sp.csc_matrix(A)
    # one entry per (row, col), rows ascending (the library sums duplicates
    # too; doing it here keeps the arrays it receives small).  The caller's
    # matrix is copied only when it is not in that form already.
    if not A.has_canonical_format:
        A = A.copy()
        A.sum_duplicates()
    return (np.ascontiguousarray(A.indptr, dtype=np.int64),
            np.ascontiguousarray(A.indices, dtype=np.int64),
            np.ascontiguousarray(A.data, dtype=np.float64))


def _batch_args(pencils):
    """ctypes arguments (count, n[], K colptr/rowidx/val[], M ..., coords[])."""
    arrs = [(_csc(pc.K), _csc(pc.M),
             None if pc.coords is None else np.ascontiguousarray(pc.coords, dtype=np.float64))
            for pc in pencils]
    cnt = len(arrs)
    ns = np.array([pc.K.shape[0] for pc in pencils], dtype=np.int64)

    def ptrs(get):
        out = (C.c_void_p * cnt)()
        for i, a in enumerate(arrs):
            v = get(a)
            out[i] = None if v is None else v.ctypes.data
        return C.cast(out, C.POINTER(C.c_void_p)), out

    cols = [ptrs(lambda a, i=i, j=j: a[i][j]) for i in (0, 1) for j in range(3)]
    cc = ptrs(lambda a: a[2])
    keep = (arrs, ns, cols, cc)
    return (cnt, iptr(ns), *[c[0] for c in cols], cc[0], keep)


class HipContext:
    """One amls_ctx.  device < 0: host-only (partition and analysis only)."""

    def __init__(self, device: int = 0):
        self._L = _lib.lib()
        h = C.c_void_p()
        st = self._L.amls_hip_create(device, 0, C.byref(h))
        if st != 0:
            raise AmlsError(st, f"amls_hip_create(device={device}) failed")
        self._h = h
        self.device = device
        self.n = 0
        self.count = 1   # problems of the last set_problem / set_problem_batch

    def close(self):
        if getattr(self, "_h", None):
            self._L.amls_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st: int):
        if st != 0:
            msg = self._L.amls_hip_last_error(self._h)
            raise AmlsError(st, msg.decode() if msg else "")

    # ---- analysis --------------------------------------------------------
    def set_problem(self, K, M, coords=None, nLevels: int = 6, leafSize: int = 500,
                    shift: float = 0.0):
        kp, ki, kx = _csc(K)
        mp, mi, mx = _csc(M)
        self._keep = (kp, ki, kx, mp, mi, mx)
        cc = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        self._chk(self._L.amls_hip_set_problem(self._h, K.shape[0], iptr(kp), iptr(ki), dptr(kx),
                                               iptr(mp), iptr(mi), dptr(mx), dptr(cc),
                                               nLevels, leafSize, shift))
        self.n = K.shape[0]
        self.count = 1

    def set_problem_batch(self, pencils, nLevels: int = 6, leafSize: int = 500,
                          shift: float = 0.0):
        """Independent pencils solved together (one forest, level-batched)."""
        args = _batch_args(pencils)
        self._keep = args
        self._chk(self._L.amls_hip_set_problem_batch(self._h, *args[:9], nLevels, leafSize,
                                                     shift))
        self.n = int(sum(pc.K.shape[0] for pc in pencils))
        self.count = len(pencils)

    def batch_info(self):
        cnt = C.c_int64()
        self._chk(self._L.amls_hip_batch_info(self._h, C.byref(cnt), None, None))
        rows = np.zeros(cnt.value + 1, dtype=np.int64)
        nodes = np.zeros(cnt.value + 1, dtype=np.int64)
        self._chk(self._L.amls_hip_batch_info(self._h, None, iptr(rows), iptr(nodes)))
        return rows, nodes

    def trace_begin(self, path: str):
        """Record a Chrome trace (host phases and GPU launches) into path."""
        self._chk(self._L.amls_hip_trace(self._h, str(path).encode(), None))

    def trace_end(self) -> int:
        """Write the trace started by trace_begin; the number of spans."""
        n = C.c_int64()
        self._chk(self._L.amls_hip_trace(self._h, None, C.byref(n)))
        return n.value

    def setup_timing(self) -> dict:
        """Seconds of the last set_problem: analysis (canonical columns, graph,
        nested dissection, symbolic tables), tables (device-side index maps and
        operator layouts built on the host), upload, total."""
        t = np.zeros(4)
        self._chk(self._L.amls_hip_setup_timing(self._h, dptr(t)))
        return dict(zip(("analysis", "tables", "upload", "total"), t.tolist()))

    def set_partition(self, method: str = "auto"):
        """Bisection of the following set_problem calls: "auto" (planes with
        coordinates, else spectral cut + FM refinement on low-degree graphs,
        else BFS level sets) or "levels" (BFS level sets only)."""
        self._chk(self._L.amls_hip_set_partition(self._h, {"auto": 0, "levels": 1}[method]))

    def tree_stats(self) -> dict:
        """Statistics of the current separator tree: per level (0 = leaves)
        the node count and the min / max / total node and boundary sizes, and
        in all the separator total, the front storage sum (n_p + b_p)^2 and
        the dense elimination flops sum n_p^3 / 3 + n_p^2 b_p + n_p b_p^2
        (the factor, X and the Schur update of each front)."""
        return tree_stats(self.tree())

    def tree(self) -> dict:
        nn, ml, n = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._L.amls_hip_tree_info(self._h, C.byref(nn), C.byref(ml), C.byref(n)))
        N = nn.value
        out = {k: np.zeros(N, dtype=np.int64)
               for k in ("treeId", "offset", "size", "parent", "level", "bsize")}
        perm = np.zeros(n.value, dtype=np.int64)
        self._chk(self._L.amls_hip_tree_get(self._h, iptr(out["treeId"]), iptr(out["offset"]),
                                            iptr(out["size"]), iptr(out["parent"]),
                                            iptr(out["level"]), iptr(out["bsize"]), iptr(perm)))
        out["perm"] = perm
        out["maxLevel"] = ml.value
        return out

    # ---- phases ----------------------------------------------------------
    def ge(self, precision: str = "fp64"):
        self._chk(self._L.amls_hip_ge(self._h, PRECISIONS[precision]))

    def spectral_prepare(self):
        self._chk(self._L.amls_hip_spectral_prepare(self._h))

    def set_sil_mode(self, mode: int):
        """Local solver switch of the following spectral_prepare / spectral
        calls: 0 density switch, 1 always shift-invert Lanczos, -1 never."""
        self._chk(self._L.amls_hip_set_sil_mode(self._h, int(mode)))

    def sil_info(self):
        """(path, steps) per post-order node after spectral(): path 0 dense,
        1 shift-invert Lanczos, 2 Lanczos fell back to dense; steps = the
        Lanczos steps whose Ritz pairs were kept."""
        N = self.tree()["size"].shape[0]
        path = np.zeros(N, dtype=np.int32)
        steps = np.zeros(N, dtype=np.int64)
        self._chk(self._L.amls_hip_get_sil(self._h, path.ctypes.data_as(C.POINTER(C.c_int32)),
                                           iptr(steps)))
        return path, steps

    def set_rr_mode(self, mode: int):
        """Solver of the final Rayleigh-Ritz problem of the subspace
        iteration: 0 Chebyshev-filtered subspace iteration where it applies
        (dense fallback), 1 always the dense reduction."""
        self._chk(self._L.amls_hip_set_rr_mode(self._h, int(mode)))

    def rr_info(self):
        """(filter iterations, largest relative Ritz residual) of the last
        final Rayleigh-Ritz; 0 iterations when the dense path gave it."""
        it = C.c_int32()
        res = C.c_double()
        self._chk(self._L.amls_hip_get_rr_info(self._h, C.byref(it), C.byref(res)))
        return it.value, res.value

    def spectral(self, kind: str, value: float) -> np.ndarray:
        N = self.tree()["size"].shape[0]
        k = np.zeros(N, dtype=np.int64)
        self._chk(self._L.amls_hip_spectral(self._h, {"fixed": 0, "threshold": 1}[kind],
                                            float(value), iptr(k)))
        return k

    def spectral_pinned(self, k) -> np.ndarray:
        """Every node keeps exactly min(k[p], n_p) modes from the dense
        reduction (post-order counts); returns the counts kept."""
        N = self.tree()["size"].shape[0]
        kin = np.ascontiguousarray(np.asarray(k, dtype=np.int64).reshape(-1))
        if kin.size != N:
            raise ValueError(f"one count per node expected ({N})")
        kout = np.zeros(N, dtype=np.int64)
        self._chk(self._L.amls_hip_spectral_pinned(self._h, iptr(kin), iptr(kout)))
        return kout

    def truncation_margin(self):
        """(margin, node) of the last threshold truncation: the smallest
        relative distance of a node's last kept or first dropped local
        eigenvalue to the cut; (inf, -1) after a fixed-count selection."""
        m = C.c_double()
        p = C.c_int64()
        self._chk(self._L.amls_hip_truncation_margin(self._h, C.byref(m), C.byref(p)))
        return m.value, p.value

    def project(self, precision: str = "fp64") -> int:
        p = C.c_int64()
        self._chk(self._L.amls_hip_project(self._h, PRECISIONS[precision], C.byref(p)))
        return p.value

    def _per_problem(self, v: np.ndarray, k: int) -> np.ndarray:
        # one problem: k values; a batch: count x k, problem-major
        return v if self.count == 1 else v.reshape(self.count, k)

    def reduced_solve(self, want: int, vectors: bool = True) -> np.ndarray:
        th = np.zeros(self.count * want, dtype=np.float64)
        self._chk(self._L.amls_hip_reduced_solve(self._h, want, int(vectors), dptr(th)))
        return self._per_problem(th, want)

    def backtransform(self, want: int) -> np.ndarray:
        Z = np.zeros((self.n, want), dtype=np.float64, order="F")
        self._chk(self._L.amls_hip_backtransform(self._h, dptr(Z)))
        return Z

    def subspace_iter(self, steps: int, q: int, nev: int):
        th = np.zeros(self.count * nev, dtype=np.float64)
        Z = np.zeros((self.n, nev), dtype=np.float64, order="F")
        self._chk(self._L.amls_hip_subspace_iter(self._h, steps, q, nev, dptr(th), dptr(Z)))
        return self._per_problem(th, nev), Z

    # ---- measurement -----------------------------------------------------
    def autotune(self) -> np.ndarray:
        """Measures the GEMM's split-k choices on this device and installs the
        fastest per shape; returns the table, rows (output tiles, k, split)."""
        tab = np.zeros(3 * 64, dtype=np.int64)
        rows = C.c_int64()
        self._chk(self._L.amls_hip_autotune(self._h, iptr(tab), tab.size, C.byref(rows)))
        return tab[:3 * rows.value].reshape(-1, 3).copy()

    def set_tuning(self, table):
        """Installs a split table (rows of output tiles, k, split); an empty
        table restores the built-in rule."""
        t = np.ascontiguousarray(np.asarray(table, dtype=np.int64).reshape(-1, 3))
        self._chk(self._L.amls_hip_set_tuning(self._h, iptr(t) if t.size else None, t.shape[0]))

    def profile(self, enable: bool = True, by_phase: bool = False, by_shape: bool = False):
        """HIP-event timing of every launch class; with by_phase the classes
        are keyed "phase/class" (ge, spectral, select, project, reduced,
        backtransform, si), with by_shape GEMM batches also by their shape."""
        mode = (3 if by_shape else 2 if by_phase else 1) if enable else 0
        self._chk(self._L.amls_hip_profile(self._h, mode))

    def kernel_stats(self) -> dict:
        out = {}
        i = 0
        while True:
            name = C.create_string_buffer(64)
            la, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
            st = self._L.amls_hip_kernel_stats(self._h, i, name, 64, C.byref(la), C.byref(ms),
                                               C.byref(fl), C.byref(by))
            if st != 0:
                break
            out[name.value.decode()] = {"launches": la.value, "ms": ms.value,
                                        "flops": fl.value, "bytes": by.value}
            i += 1
        return out

    def solve_numeric(self, opt: "AmlsOptions", vectors: bool = False):
        """Numeric solve on the problem (or batch) already set (inputs resident
        in HBM).  After set_problem_batch, theta is count x nev and p holds
        the count reduced sizes."""
        nev = max(opt.nev, 1)
        theta = np.zeros(self.count * nev)
        Z = np.zeros((self.n, nev), order="F") if vectors else None
        p = np.zeros(self.count, dtype=np.int64)
        timing = np.zeros(8)
        oc = opt.to_c()
        self._chk(self._L.amls_hip_solve_numeric(self._h, C.byref(oc), dptr(theta), dptr(Z),
                                                 iptr(p), dptr(timing)))
        pv = int(p[0]) if self.count == 1 else p
        return (self._per_problem(theta, nev), Z, pv,
                dict(zip(AmlsSolver.TIMING_KEYS, timing.tolist())))

    def adjoint(self, conn, Ke0, Me0, theta, Z=None):
        """Element sensitivities (grad_s, grad_r), each nelem x nev:
        d theta_j / d s_e = x_je^T Ke0 x_je and d theta_j / d r_e =
        -(theta_j - shift) x_je^T Me0 x_je, theta as returned by the solve
        of the problem set on this context (shift: its K <- K + shift M).
        Z None uses the eigenvectors of the last solve kept in device
        memory."""
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        nelem, npe = conn.shape
        K0 = np.asfortranarray(Ke0, dtype=np.float64)
        M0 = np.asfortranarray(Me0, dtype=np.float64)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        nev = th.size
        Zc = None if Z is None else np.asfortranarray(Z[:, :nev], dtype=np.float64)
        gs = np.zeros((nelem, nev), order="F")
        gr = np.zeros((nelem, nev), order="F")
        self._chk(self._L.amls_hip_adjoint(self._h, nelem, npe, iptr(conn), dptr(K0), dptr(M0),
                                           nev, dptr(th), dptr(Zc), dptr(gs), dptr(gr)))
        return gs, gr

    # ---- inspection ------------------------------------------------------
    def node_factor(self, p: int, n_p: int):
        LD = np.zeros((n_p, n_p), order="F")
        Mh = np.zeros((n_p, n_p), order="F")
        self._chk(self._L.amls_hip_get_node_factor(self._h, p, dptr(LD), dptr(Mh)))
        return LD, Mh

    def local(self, p: int, n_p: int, k_p: int):
        mu = np.zeros(max(k_p, 1))
        Z = np.zeros((n_p, max(k_p, 1)), order="F")
        self._chk(self._L.amls_hip_get_local(self._h, p, dptr(mu), dptr(Z)))
        return mu[:k_p], Z[:, :k_p]

    def reduced(self, ptot: int, nnodes: int):
        Mt = np.zeros((ptot, ptot), order="F")
        moff = np.zeros(nnodes, dtype=np.int64)
        self._chk(self._L.amls_hip_get_reduced(self._h, dptr(Mt), iptr(moff)))
        return Mt, moff

    # ---- operator-granular ---------------------------------------------
    def gemm(self, A, B, C0=None, alpha=1.0, beta=0.0, transA=False, transB=False,
             precision="fp64"):
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        m = A.shape[1] if transA else A.shape[0]
        k = A.shape[0] if transA else A.shape[1]
        n = B.shape[0] if transB else B.shape[1]
        Cm = np.zeros((m, n), order="F") if C0 is None else np.asfortranarray(C0, dtype=np.float64).copy(order="F")
        self._chk(self._L.amls_hip_gemm(self._h, PRECISIONS[precision], int(transA), int(transB),
                                        m, n, k, alpha, dptr(A), max(A.shape[0], 1), dptr(B),
                                        max(B.shape[0], 1), beta, dptr(Cm), max(m, 1)))
        return Cm

    def ldlt_solve(self, A, B):
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        n = A.shape[0]
        LD = np.zeros((n, n), order="F")
        X = np.zeros_like(B, order="F")
        self._chk(self._L.amls_hip_ldlt_solve(self._h, n, dptr(A), B.shape[1], dptr(B),
                                              dptr(LD), dptr(X)))
        return LD, X

    def gemm_gather(self, A, B, C0=None, m=None, n=None, k=None, alpha=1.0, beta=0.0,
                    transA=False, transB=False, aRows=None, aCols=None, bRows=None, bCols=None,
                    precision="fp64"):
        """C = alpha op(A[aRows][:, aCols]) op(B[bRows][:, bCols]) + beta C, the
        gathers done on the device (the GEMM of the ColCompOps methods)."""
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        idx = [None if v is None else np.ascontiguousarray(v, dtype=np.int32)
               for v in (aRows, aCols, bRows, bCols)]
        ar = A.shape[0] if idx[0] is None else idx[0].size
        ac = A.shape[1] if idx[1] is None else idx[1].size
        br = B.shape[0] if idx[2] is None else idx[2].size
        bc = B.shape[1] if idx[3] is None else idx[3].size
        m = (ac if transA else ar) if m is None else m
        k = (ar if transA else ac) if k is None else k
        n = (br if transB else bc) if n is None else n
        Cm = (np.zeros((m, n), order="F") if C0 is None
              else np.asfortranarray(C0, dtype=np.float64).copy(order="F"))
        self._chk(self._L.amls_hip_gemm_gather(
            self._h, PRECISIONS[precision], int(transA), int(transB), m, n, k, alpha, dptr(A),
            max(A.shape[0], 1), A.shape[1], _lib.i32ptr(idx[0]), _lib.i32ptr(idx[1]), dptr(B),
            max(B.shape[0], 1), B.shape[1], _lib.i32ptr(idx[2]), _lib.i32ptr(idx[3]), beta,
            dptr(Cm), max(m, 1)))
        return Cm

    def ldlt_factor(self, A, pivoting: bool = True) -> "DenseFactor":
        """ISparseDirectSolver::factorize on the device: the lower triangle of
        the sparse or dense A, with Eigen LDLT's diagonal pivoting when
        `pivoting` (DenseLDLTSolver) or none (SimplicialLDLT's natural
        order)."""
        kp, ki, kx = _csc(sp.csc_matrix(A))
        h = C.c_void_p()
        self._chk(self._L.amls_hip_ldlt_factor(self._h, A.shape[0], iptr(kp), iptr(ki), dptr(kx),
                                               int(pivoting), C.byref(h)))
        return DenseFactor(self, h, A.shape[0])

    def dense_eig(self, K, M, range_: str = "all", lo: float = 0.0, hi: float = 0.0,
                  il: int = 1, iu: int = 0, vectors: bool = True):
        """IDenseEigensolver::solveAll / solveInterval (lo, hi] / solveIndex
        [il, iu] (1-based) on K x = lambda M x: (w ascending, V M-orthonormal)."""
        K = np.asfortranarray(K, dtype=np.float64)
        M = np.asfortranarray(M, dtype=np.float64)
        n = K.shape[0]
        r = {"all": 0, "value": 1, "index": 2}[range_]
        w = np.zeros(max(n, 1))
        V = np.zeros((n, max(n, 1)), order="F") if vectors else None
        m = C.c_int64()
        self._chk(self._L.amls_hip_dense_eig(self._h, n, dptr(K), dptr(M), r, lo, hi, il, iu,
                                             C.byref(m), dptr(w), dptr(V)))
        mm = m.value
        return w[:mm].copy(), (None if V is None else V[:, :mm].copy(order="F"))

    def syevx_top(self, A, k: int):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        lam = np.zeros(max(k, 1))
        V = np.zeros((n, max(k, 1)), order="F")
        self._chk(self._L.amls_hip_syevx_top(self._h, n, dptr(A), k, dptr(lam), dptr(V)))
        return lam[:k], V[:, :k]


class DenseFactor:
    """A device LDL^T factor (amls_factor): solve(B) = A^-1 B
    (solveWithFactor, src/linalg/sparse_direct.hpp:23), factor() = (LD of
    P A P^T, perm with perm[new] = old)."""

    def __init__(self, ctx: HipContext, handle, n: int):
        self.ctx, self.h, self.n = ctx, handle, n

    def solve(self, B):
        B = np.asfortranarray(B, dtype=np.float64).reshape(self.n, -1, order="F")
        X = np.zeros_like(B, order="F")
        self.ctx._chk(self.ctx._L.amls_hip_ldlt_factor_solve(self.ctx._h, self.h, B.shape[1],
                                                             dptr(B), dptr(X)))
        return X

    def factor(self):
        LD = np.zeros((self.n, self.n), order="F")
        perm = np.zeros(self.n, dtype=np.int64)
        self.ctx._chk(self.ctx._L.amls_hip_ldlt_factor_get(self.ctx._h, self.h, dptr(LD),
                                                           iptr(perm)))
        return LD, perm

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.amls_hip_ldlt_factor_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tree_stats(t: dict) -> dict:
    """Separator-tree statistics from a tree() dictionary (see
    HipContext.tree_stats)."""
    size, b, lvl = (np.asarray(t[k], dtype=np.float64) for k in ("size", "bsize", "level"))
    levels = []
    for l in range(int(t["maxLevel"]) + 1):
        m = np.asarray(t["level"]) == l
        levels.append({"level": l, "nodes": int(m.sum()),
                       "size_min": int(size[m].min()), "size_max": int(size[m].max()),
                       "size_sum": int(size[m].sum()),
                       "bsize_min": int(b[m].min()), "bsize_max": int(b[m].max()),
                       "bsize_sum": int(b[m].sum())})
    sep = np.asarray(t["level"]) > 0
    return {"nodes": int(len(size)), "levels": levels,
            "separator_total": int(size[sep].sum()),
            "root_size": int(size[np.asarray(t["parent"]) < 0].max()),
            "front_storage": float(((size + b) ** 2).sum()),
            "elim_flops": float((size ** 3 / 3 + size ** 2 * b + size * b ** 2).sum())}


class AmlsSolver:
    """AmlsSolver::solve (src/amls/solver.cpp:100-319) on one GPU."""

    TIMING_KEYS = ("partition", "gauss", "condense", "reduced", "backtransform", "subspace",
                   "total", "p0")

    def __init__(self, device: int = 0):
        self.ctx = HipContext(device)

    def solve(self, pencil: Pencil, opt: AmlsOptions) -> Solution:
        K, M = pencil.K, pencil.M
        n = K.shape[0]
        nev = min(opt.nev, n) if n > 0 else opt.nev
        kp, ki, kx = _csc(K)
        mp, mi, mx = _csc(M)
        cc = None if pencil.coords is None else np.ascontiguousarray(pencil.coords, dtype=np.float64)
        theta = np.zeros(max(nev, 1))
        Z = np.zeros((n, max(nev, 1)), order="F") if opt.recoverVectors else None
        p = C.c_int64()
        timing = np.zeros(8)
        oc = opt.to_c()
        L = self.ctx._L
        st = L.amls_hip_solve(self.ctx._h, n, iptr(kp), iptr(ki), dptr(kx), iptr(mp), iptr(mi),
                              dptr(mx), dptr(cc), C.byref(oc), dptr(theta), dptr(Z), C.byref(p),
                              dptr(timing))
        self.ctx._chk(st)
        self.ctx.n = n
        m = min(nev, p.value) if opt.siSteps == 0 else nev
        return Solution(theta=theta[:m].copy(), Z=None if Z is None else Z[:, :m].copy(order="F"),
                        p=p.value, timing=dict(zip(self.TIMING_KEYS, timing.tolist())))

    def solve_batch(self, pencils, opt: AmlsOptions):
        """Independent pencils in one batched solve; one Solution per pencil
        (timing is that of the whole batch)."""
        cnt = len(pencils)
        ns = [pc.K.shape[0] for pc in pencils]
        nev = min([opt.nev] + ns)
        ntot = int(sum(ns))
        theta = np.zeros(cnt * opt.nev)
        vec = opt.recoverVectors or opt.siSteps > 0
        Z = np.zeros((ntot, nev), order="F") if vec else None
        p = np.zeros(cnt, dtype=np.int64)
        timing = np.zeros(8)
        oc = opt.to_c()
        args = _batch_args(pencils)
        self.ctx._chk(self.ctx._L.amls_hip_solve_batch(self.ctx._h, *args[:9], C.byref(oc),
                                                       dptr(theta), dptr(Z), iptr(p),
                                                       dptr(timing)))
        self.ctx.n = ntot
        tm = dict(zip(self.TIMING_KEYS, timing.tolist()))
        out, r0 = [], 0
        for s_, n_s in enumerate(ns):
            m = min(nev, int(p.min())) if opt.siSteps == 0 else nev
            out.append(Solution(theta=theta[s_ * opt.nev:s_ * opt.nev + m].copy(),
                                Z=None if Z is None else Z[r0:r0 + n_s, :m].copy(order="F"),
                                p=int(p[s_]), timing=tm))
            r0 += n_s
        return out


from .store import ContributionStore  # noqa: E402  (the store only needs _lib)
