"""ColComp blocks and the ColCompOps operator plugin of the reference
(src/linalg/colcomp.hpp:26-75, 124-171) over libamls_hip's device-gathered
GEMM, amls_hip_gemm_gather: a GPU backend of the interface that
GaussianElimination::setOps injects (src/amls/gaussian_elimination.hpp
:122-124).  Dense operands are numpy arrays, Sparse ones scipy sparse
matrices (multiplied as dense blocks)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp


@dataclass
class ColComp:
    """data: the stored columns of a logical rows x full block; cols: their
    ascending logical numbers (int32), None for the identity map (ident)."""
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


class HipColCompOps:
    """The ColCompOps methods the reference's GE and condensation call
    (gemm, subMul, addMul, subTransMul, addTransMul, mul, expandT), each one
    amls_hip_gemm_gather: the stored-column lists travel as index lists and
    the gathers run on the device."""

    def __init__(self, ctx, precision: str = "fp64"):
        self.ctx = ctx
        self.precision = precision

    def _g(self, A, B, C0, alpha, beta, ta=False, tb=False, aRows=None, aCols=None, bRows=None,
           bCols=None, m=None, n=None, k=None):
        return self.ctx.gemm_gather(A, B, C0=C0, alpha=alpha, beta=beta, transA=ta, transB=tb,
                                    aRows=aRows, aCols=aCols, bRows=bRows, bCols=bCols,
                                    m=m, n=n, k=k, precision=self.precision)

    def gemm(self, ta: bool, tb: bool, alpha, A, B, beta, C):
        """C = alpha op(A) op(B) + beta C; with ta the result lives in A's
        stored-row space (colcomp.hpp:132-137)."""
        if isinstance(A, ColComp):
            if not A.ident and len(A.cols) == 0:
                return beta * C
            if not A.ident and not ta:
                return self._g(A.data, B, C, alpha, beta, False, tb, bRows=A.cols)
            A = A.data
        return self._g(A, B, C, alpha, beta, ta, tb)

    def _acc(self, A, B, C, sign, trans):
        if sp.issparse(B):
            B = ColComp(np.asarray(B.todense()), B.shape[1])
        if B.empty() or (isinstance(A, ColComp) and A.empty()):
            return A
        dense_target = not isinstance(A, ColComp)
        data = A if dense_target else A.data
        acols = None if dense_target or A.ident else A.cols
        brows = None if B.ident else B.cols
        if not trans:
            out = self._g(B.data, C, data, sign, 1.0, bRows=brows, bCols=acols)
        else:
            # (B C)^T on the target's stored columns: Cg^T S^T with Cg = C(B.cols, :),
            # S = B.data(A.cols, :)
            out = self._g(C, B.data, data, sign, 1.0, ta=True, tb=True, aRows=brows, bRows=acols)
        return out if dense_target else ColComp(out, A.full, A.cols)

    def subMul(self, A, B, C):
        return self._acc(A, B, C, -1.0, False)

    def addMul(self, A, B, C):
        return self._acc(A, B, C, +1.0, False)

    def subTransMul(self, A, B, C):
        return self._acc(A, B, C, -1.0, True)

    def addTransMul(self, A, B, C):
        return self._acc(A, B, C, +1.0, True)

    def mul(self, B, C):
        if sp.issparse(B):
            B = ColComp(np.asarray(B.todense()), B.shape[1])
        return self.gemm(False, False, 1.0, B, C, 0.0, np.zeros((B.data.shape[0], C.shape[1])))

    def expandT(self, B):
        if sp.issparse(B):
            return np.asarray(B.T.todense())
        out = np.zeros((B.full, B.data.shape[0]))
        if B.ident:
            return B.data.T.copy()
        out[np.asarray(B.cols)] = B.data.T
        return out
