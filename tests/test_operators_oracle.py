"""CPU pins of the oracle's operator restatements: the ColCompOps products
(src/linalg/colcomp.hpp:174-423) against brute-force products of the logical
blocks the compressed storage stands for, and Eigen's pivoted LDL^T
(src/linalg/sparse_direct.cpp:38-63) against reconstruction, the closed
form of a diagonal matrix, and the selection order it must end with."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import amls_oracle as orc


def _cc(rng, rows, full, kind):
    if kind == "ident":
        return orc.ColComp(rng.standard_normal((rows, full)), full)
    if kind == "empty":
        return orc.ColComp(np.zeros((rows, 0)), full, np.zeros(0, dtype=np.int32))
    cols = np.sort(rng.choice(full, size=max(1, full // 3), replace=False)).astype(np.int32)
    return orc.ColComp(rng.standard_normal((rows, cols.size)), full, cols)


KINDS = ["ident", "comp", "empty"]


@pytest.mark.parametrize("akind", KINDS)
@pytest.mark.parametrize("bkind", KINDS + ["sparse"])
@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_colcomp_acc_matches_logical_product(akind, bkind, trans, sign):
    rng = np.random.default_rng(hash((akind, bkind, trans, sign)) % 2**32)
    ops = orc.NumpyColCompOps()
    rowsB, fullB, nC = 9, 11, 7
    C = rng.standard_normal((fullB, nC))
    if bkind == "sparse":
        B = sp.random(rowsB, fullB, density=0.3, random_state=3, format="csc")
        Blog = B.toarray()
    else:
        B = _cc(rng, rowsB, fullB, bkind)
        Blog = B.logical()
    prod = Blog @ C                                  # rowsB x nC
    if not trans:
        A = _cc(rng, rowsB, nC, akind)
        ref = A.logical() + sign * prod
    else:
        A = _cc(rng, nC, rowsB, akind)
        ref = A.logical() + sign * prod.T
    fn = {(False, -1.0): ops.subMul, (False, 1.0): ops.addMul,
          (True, -1.0): ops.subTransMul, (True, 1.0): ops.addTransMul}[(trans, sign)]
    out = fn(A, B, C)
    if A.empty():
        assert out.stored() == 0
        return
    got = out.logical()
    keep = np.arange(A.full) if A.ident else A.cols
    np.testing.assert_allclose(got[:, keep], ref[:, keep], rtol=1e-13, atol=1e-13)
    # columns the target does not store stay zero
    rest = np.setdiff1d(np.arange(A.full), keep)
    assert not np.any(got[:, rest])


@pytest.mark.parametrize("bkind", KINDS + ["sparse"])
def test_colcomp_dense_target_submul(bkind):
    rng = np.random.default_rng(5)
    ops = orc.NumpyColCompOps()
    C = rng.standard_normal((10, 4))
    if bkind == "sparse":
        B = sp.random(6, 10, density=0.4, random_state=1, format="csc")
        Blog = B.toarray()
    else:
        B = _cc(rng, 6, 10, bkind)
        Blog = B.logical()
    A = rng.standard_normal((6, 4))
    np.testing.assert_allclose(ops.subMul(A, B, C), A - Blog @ C, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
def test_colcomp_gemm_contract(kind, ta, tb):
    rng = np.random.default_rng(11)
    ops = orc.NumpyColCompOps()
    A = _cc(rng, 8, 12, kind)
    if not ta:
        if tb and kind == "comp":
            pytest.skip("the reference gathers op(B)'s operand by rows before transposing: "
                        "no logical-block meaning")
        B = rng.standard_normal((5, 12) if tb else (12, 5))
        C0 = rng.standard_normal((8, 5))
        got = ops.gemm(False, tb, 0.5, A, B, 2.0, C0)
        if kind == "empty":
            np.testing.assert_allclose(got, 2.0 * C0)
            return
        ref = 0.5 * (A.logical() @ (B.T if tb else B)) + 2.0 * C0
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)
    else:
        B = rng.standard_normal((5, 8) if tb else (8, 5))
        Bop = B.T if tb else B
        s = A.stored()
        C0 = rng.standard_normal((s, 5))
        got = ops.gemm(True, tb, 1.5, A, B, -1.0, C0)
        if kind == "empty":
            np.testing.assert_allclose(got, -C0)
            return
        # stored-row space: C(j, :) = sum_r data(r, j) op(B)(r, :)
        stored = A.logical() if A.ident else A.logical()[:, A.cols]
        np.testing.assert_allclose(got, 1.5 * stored.T @ Bop - C0, rtol=1e-13, atol=1e-13)


def test_colcomp_expand_and_mul():
    rng = np.random.default_rng(2)
    ops = orc.NumpyColCompOps()
    for kind in KINDS:
        B = _cc(rng, 5, 9, kind)
        np.testing.assert_array_equal(ops.expandT(B), B.logical().T)
        C = rng.standard_normal((9, 3))
        np.testing.assert_allclose(ops.mul(B, C), B.logical() @ C, rtol=1e-13, atol=1e-13)
    S = sp.random(5, 9, density=0.4, random_state=0, format="csc")
    np.testing.assert_array_equal(ops.expandT(S), S.toarray().T)


def test_colcomp_planted_mistake_caught():
    """A dropped row gather in the transposed term (S = B.data instead of
    B.data(A.cols, :)) is rejected by the logical-product check."""
    rng = np.random.default_rng(9)
    B = _cc(rng, 9, 11, "ident")
    A = _cc(rng, 7, 9, "comp")
    C = rng.standard_normal((11, 7))
    ref = (A.logical() - (B.logical() @ C).T)[:, A.cols]
    wrong = A.data - C.T @ B.data[: A.cols.size].T
    assert not np.allclose(wrong, ref)
    np.testing.assert_allclose(orc.NumpyColCompOps().subTransMul(A, B, C).data, ref,
                               rtol=1e-13, atol=1e-13)


def _spd(rng, n, cond=1e3):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(rng.uniform(0, np.log(cond), n))
    return (Q * d) @ Q.T


@pytest.mark.parametrize("n", [1, 2, 7, 40])
def test_eigen_ldlt_reconstructs(n):
    rng = np.random.default_rng(n)
    A = _spd(rng, n)
    perm, L, D = orc.eigen_ldlt(A)
    np.testing.assert_allclose(L @ np.diag(D) @ L.T, A[np.ix_(perm, perm)], rtol=1e-11,
                               atol=1e-11 * np.abs(A).max())
    assert np.all(D > 0)
    np.testing.assert_array_equal(perm, orc.diagonal_selection_order(np.diag(A)))


def test_eigen_ldlt_diagonal_closed_form():
    """A diagonal matrix: D is the diagonal sorted by |d| descending (first
    index wins ties) and L = I."""
    d = np.array([3.0, -7.0, 7.0, 1.0, 0.5, 3.0])
    perm, L, D = orc.eigen_ldlt(np.diag(d))
    np.testing.assert_array_equal(perm, [1, 2, 0, 5, 3, 4])
    np.testing.assert_array_equal(D, d[perm])
    np.testing.assert_array_equal(L, np.eye(6))


def test_eigen_ldlt_order_ignores_updated_diagonal():
    """Left-looking: the pivot order follows the input diagonal, not the
    Schur-complement diagonal a right-looking (LAPACK-style) pivoting
    would use.  Here the two orders differ."""
    A = np.array([[4.0, 3.9, 0.0], [3.9, 4.0, 0.0], [0.0, 0.0, 3.0]])
    perm, L, D = orc.eigen_ldlt(A)
    np.testing.assert_array_equal(perm, [0, 1, 2])      # 4, 4, 3 in input order
    # a right-looking rule would pick 3.0 (> 4 - 3.9^2/4 = 0.1975) second
    assert D[1] == pytest.approx(4.0 - 3.9**2 / 4.0)
    np.testing.assert_allclose(L @ np.diag(D) @ L.T, A, rtol=1e-14)
