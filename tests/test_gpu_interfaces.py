"""The operator-interface exports on the GPU, through the C-ABI, against the
oracle: ColCompOps (src/linalg/colcomp.hpp:124-423) as HipColCompOps over
amls_hip_gemm_gather, ISparseDirectSolver's factor handle with Eigen LDLT
pivoting (src/linalg/sparse_direct.cpp:38-63), and IDenseEigensolver's
solveAll / solveInterval / solveIndex (src/linalg/dense_eigensolver.cpp
:11-133) against LAPACK's ?sygvd / ?sygvx through scipy."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from amls_amd import AmlsError
from amls_amd.colcomp import ColComp as HColComp, HipColCompOps
from oracle import amls_oracle as orc

pytestmark = pytest.mark.gpu

KINDS = ["ident", "comp", "empty"]


def _pair(rng, rows, full, kind):
    if kind == "ident":
        d = rng.standard_normal((rows, full))
        return orc.ColComp(d, full), HColComp(d.copy(), full)
    if kind == "empty":
        z = np.zeros(0, dtype=np.int32)
        return (orc.ColComp(np.zeros((rows, 0)), full, z),
                HColComp(np.zeros((rows, 0)), full, z.copy()))
    cols = np.sort(rng.choice(full, size=max(1, full // 3), replace=False)).astype(np.int32)
    d = rng.standard_normal((rows, cols.size))
    return orc.ColComp(d, full, cols), HColComp(d.copy(), full, cols.copy())


def _data(x):
    return x if isinstance(x, np.ndarray) else x.data


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 2e-5)])
@pytest.mark.parametrize("method", ["subMul", "addMul", "subTransMul", "addTransMul"])
def test_colcomp_ops_match_oracle(gpu_ctx, precision, tol, method):
    rng = np.random.default_rng(["subMul", "addMul", "subTransMul", "addTransMul"].index(method))
    ref_ops = orc.NumpyColCompOps()
    hip_ops = HipColCompOps(gpu_ctx, precision)
    trans = "Trans" in method
    rowsB, fullB, nC = 37, 53, 29
    for akind in KINDS:
        for bkind in KINDS + ["sparse"]:
            C = rng.standard_normal((fullB, nC))
            if bkind == "sparse":
                Bo = Bh = sp.random(rowsB, fullB, density=0.2, random_state=7, format="csc")
            else:
                Bo, Bh = _pair(rng, rowsB, fullB, bkind)
            Ao, Ah = _pair(rng, nC, rowsB, akind) if trans else _pair(rng, rowsB, nC, akind)
            want = getattr(ref_ops, method)(Ao, Bo, C)
            got = getattr(hip_ops, method)(Ah, Bh, C)
            scale = max(1.0, np.abs(_data(want)).max(initial=0.0))
            np.testing.assert_allclose(_data(got), _data(want), rtol=0, atol=tol * scale * 10,
                                       err_msg=f"{method} A={akind} B={bkind}")
    if not trans:                        # dense target
        A = rng.standard_normal((rowsB, nC))
        for bkind in KINDS:
            Bo, Bh = _pair(rng, rowsB, fullB, bkind)
            C = rng.standard_normal((fullB, nC))
            np.testing.assert_allclose(hip_ops.subMul(A.copy(), Bh, C), ref_ops.subMul(A, Bo, C),
                                       rtol=0, atol=tol * 100)


@pytest.mark.parametrize("kind", KINDS)
def test_colcomp_gemm_mul_expand(gpu_ctx, kind):
    rng = np.random.default_rng(3)
    ref_ops, hip_ops = orc.NumpyColCompOps(), HipColCompOps(gpu_ctx)
    Ao, Ah = _pair(rng, 31, 47, kind)
    B = rng.standard_normal((47, 13))
    C0 = rng.standard_normal((31, 13))
    np.testing.assert_allclose(hip_ops.gemm(False, False, 0.7, Ah, B, -0.5, C0),
                               ref_ops.gemm(False, False, 0.7, Ao, B, -0.5, C0), rtol=0, atol=1e-12)
    B2 = rng.standard_normal((31, 9))
    C2 = rng.standard_normal((Ao.stored(), 9))
    np.testing.assert_allclose(hip_ops.gemm(True, False, 1.5, Ah, B2, 1.0, C2),
                               ref_ops.gemm(True, False, 1.5, Ao, B2, 1.0, C2), rtol=0, atol=1e-12)
    np.testing.assert_allclose(hip_ops.mul(Ah, B), ref_ops.mul(Ao, B), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(hip_ops.expandT(Ah), ref_ops.expandT(Ao))


def test_gemm_gather_all_gathers(gpu_ctx):
    """Every combination of row / column gathers and transposes, with
    leading dimensions larger than the gathered blocks."""
    rng = np.random.default_rng(21)
    A = rng.standard_normal((90, 70))
    B = rng.standard_normal((80, 60))
    ran = 0
    for ta in (False, True):
        for tb in (False, True):
            for ga in range(4):
                for gb in range(4):
                    ar = rng.permutation(90)[:40].astype(np.int32) if ga & 1 else None
                    ac = rng.permutation(70)[:30].astype(np.int32) if ga & 2 else None
                    Ag = A[ar if ar is not None else slice(None)][:, ac if ac is not None else slice(None)]
                    opA = Ag.T if ta else Ag
                    kdim = opA.shape[1]
                    # B's gathered op must have kdim rows
                    if tb:
                        bc = rng.permutation(60)[:kdim].astype(np.int32) if kdim <= 60 else None
                        if bc is None and kdim != 60:
                            continue
                        br = rng.permutation(80)[:25].astype(np.int32) if gb & 1 else None
                    else:
                        br = rng.permutation(80)[:kdim].astype(np.int32) if kdim <= 80 else None
                        if br is None and kdim != 80:
                            continue
                        bc = rng.permutation(60)[:20].astype(np.int32) if gb & 2 else None
                    Bg = B[br if br is not None else slice(None)][:, bc if bc is not None else slice(None)]
                    opB = Bg.T if tb else Bg
                    if opB.shape[0] != kdim:
                        continue
                    C0 = rng.standard_normal((opA.shape[0], opB.shape[1]))
                    got = gpu_ctx.gemm_gather(A, B, C0=C0, alpha=-1.25, beta=0.5, transA=ta,
                                              transB=tb, aRows=ar, aCols=ac, bRows=br, bCols=bc)
                    np.testing.assert_allclose(got, -1.25 * opA @ opB + 0.5 * C0, rtol=0,
                                               atol=1e-12, err_msg=f"ta={ta} tb={tb} ga={ga} gb={gb}")
                    ran += 1
    assert ran >= 32


def test_gemm_gather_rejects_bad_index(gpu_ctx):
    A = np.ones((5, 5))
    with pytest.raises(AmlsError, match="index out of range"):
        gpu_ctx.gemm_gather(A, A, aRows=np.array([0, 7], dtype=np.int32), m=2, n=5, k=5)


def _spd(rng, n, cond):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(rng.uniform(0, np.log(cond), n))
    return (Q * d) @ Q.T


@pytest.mark.parametrize("n,cond", [(1, 1.0), (17, 1e3), (130, 1e6), (300, 1e12)])
def test_ldlt_factor_pivot_order_and_solve(gpu_ctx, n, cond):
    """The factor handle: Eigen LDLT's pivot order, L and D against the
    oracle's restatement of Eigen's left-looking factorisation, D > 0, and
    the solve residual, on SPD blocks up to cond 1e12."""
    rng = np.random.default_rng(n)
    A = _spd(rng, n, cond)
    A = 0.5 * (A + A.T)
    f = gpu_ctx.ldlt_factor(sp.csc_matrix(np.tril(A)))
    LD, perm = f.factor()
    perm_o, L_o, D_o = orc.eigen_ldlt(A)
    np.testing.assert_array_equal(perm, perm_o)
    D = np.diag(LD)
    assert np.all(D > 0)
    # entrywise to rounding of the largest pivot (small pivots of an ill-conditioned
    # block carry an absolute error of about n eps max|D|)
    np.testing.assert_allclose(D, D_o, rtol=1e-12, atol=1e-14 * n * D_o.max())
    L = np.tril(LD, -1) + np.eye(n)
    np.testing.assert_allclose(L, L_o, rtol=0, atol=1e-10 * max(1.0, cond ** 0.5))
    B = rng.standard_normal((n, 3))
    X = f.solve(B)
    r = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert r < 1e-14
    f.close()


def test_ldlt_factor_unpivoted_and_errors(gpu_ctx):
    rng = np.random.default_rng(4)
    A = _spd(rng, 40, 1e2)
    f = gpu_ctx.ldlt_factor(sp.csc_matrix(A), pivoting=False)
    LD, perm = f.factor()
    np.testing.assert_array_equal(perm, np.arange(40))
    c = np.linalg.cholesky(A)
    np.testing.assert_allclose(np.diag(LD), np.diag(c) ** 2, rtol=1e-12)
    f.close()
    Ai = A.copy()
    Ai[np.diag_indices(40)] -= 2 * np.linalg.eigvalsh(A).max()       # negative definite
    Ai[0, 0] = 1.0
    with pytest.raises(AmlsError, match="pivot"):
        gpu_ctx.ldlt_factor(sp.csc_matrix(Ai))


def _pencil(rng, n, indefinite=False):
    K = _spd(rng, n, 1e4)
    if indefinite:
        K = K - 0.3 * np.trace(K) / n * np.eye(n)
    M = _spd(rng, n, 1e2)
    return 0.5 * (K + K.T), 0.5 * (M + M.T)


def _check_pairs(K, M, w, V, w_ref):
    np.testing.assert_allclose(w, w_ref, rtol=1e-10, atol=1e-10 * np.abs(w_ref).max(initial=1.0))
    if V is None or V.shape[1] == 0:
        return
    np.testing.assert_allclose(V.T @ M @ V, np.eye(V.shape[1]), rtol=0, atol=1e-10)
    R = K @ V - (M @ V) * w
    assert np.linalg.norm(R) / (np.linalg.norm(K) * max(1, V.shape[1]) ** 0.5) < 1e-10


@pytest.mark.parametrize("n", [1, 8, 150, 700])
@pytest.mark.parametrize("indefinite", [False, True])
def test_dense_eig_all(gpu_ctx, n, indefinite):
    rng = np.random.default_rng(n + indefinite)
    K, M = _pencil(rng, n, indefinite)
    w, V = gpu_ctx.dense_eig(K, M, "all")
    _check_pairs(K, M, w, V, sla.eigh(K, M, driver="gvd", eigvals_only=True))


@pytest.mark.parametrize("indefinite", [False, True])
def test_dense_eig_index_and_value(gpu_ctx, indefinite):
    rng = np.random.default_rng(9)
    n = 260
    K, M = _pencil(rng, n, indefinite)
    ref = sla.eigh(K, M, driver="gvd", eigvals_only=True)
    for il, iu in [(1, 1), (1, 40), (17, 90), (250, 400)]:
        w, V = gpu_ctx.dense_eig(K, M, "index", il=il, iu=iu)
        _check_pairs(K, M, w, V, ref[il - 1:min(iu, n)])
    # interval ends halfway between eigenvalues: an end placed on a computed
    # eigenvalue makes its membership depend on the last bit of two drivers
    mid = 0.5 * (ref[:-1] + ref[1:])
    for lo, hi in [(mid[4], mid[59]), (ref[0] - 1.0, mid[10]), (mid[-1], ref[-1] + 1.0)]:
        w, V = gpu_ctx.dense_eig(K, M, "value", lo=lo, hi=hi)
        want = sla.eigh(K, M, driver="gvx", eigvals_only=True, subset_by_value=(lo, hi))
        _check_pairs(K, M, w, V, want)
    w, _ = gpu_ctx.dense_eig(K, M, "value", lo=1.0, hi=1.0)
    assert w.size == 0
    w, _ = gpu_ctx.dense_eig(K, M, "index", il=5, iu=4)
    assert w.size == 0
    w, V = gpu_ctx.dense_eig(K, M, "index", il=1, iu=12, vectors=False)
    assert V is None
    np.testing.assert_allclose(w, ref[:12], rtol=1e-10)


def test_dense_eig_not_spd_mass(gpu_ctx):
    rng = np.random.default_rng(1)
    K, M = _pencil(rng, 30)
    M[np.diag_indices(30)] -= 2 * np.linalg.eigvalsh(M).max()
    with pytest.raises(AmlsError, match="positive definite"):
        gpu_ctx.dense_eig(K, M, "all")
