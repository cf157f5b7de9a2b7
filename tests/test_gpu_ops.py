"""Operator-granular parity of the HIP kernels (through the C-ABI) against
float64 numpy references of the same operation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (64, 64, 64), (100, 37, 129), (257, 190, 65)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_fp64(gpu_ctx, m, n, k, ta, tb):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    ref = -1.5 * (A.T if ta else A) @ (B.T if tb else B) + 0.5 * C0
    out = gpu_ctx.gemm(A, B, C0, alpha=-1.5, beta=0.5, transA=ta, transB=tb)
    # FP64 MFMA with FP64 accumulation: rounding of a k-term dot product
    assert np.abs(out - ref).max() <= 1e-13 * k * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("beta", [0.0, 1.0])
@pytest.mark.parametrize("ta", [False, True])
def test_gemm_fp64_split_k(gpu_ctx, beta, ta):
    # few tiles and k >= 8192: the product is split over k and the pieces
    # are added into C atomically
    m, n, k = 300, 200, 9000
    rng = np.random.default_rng(7 + int(beta) + 2 * ta)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((k, n))
    C0 = rng.standard_normal((m, n))
    ref = -1.5 * (A.T if ta else A) @ B + beta * C0
    out = gpu_ctx.gemm(A, B, C0, alpha=-1.5, beta=beta, transA=ta)
    assert np.abs(out - ref).max() <= 1e-13 * k * max(1.0, np.abs(ref).max())


def test_gemm_fp32(gpu_ctx):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((200, 150))
    B = rng.standard_normal((150, 130))
    out = gpu_ctx.gemm(A, B, precision="mixed")
    ref = A @ B
    # FP32 operands and accumulation: ~sqrt(k) * 2^-24 relative to |A||B|
    scale = np.abs(A).max() * np.abs(B).max() * 150
    assert np.abs(out - ref).max() <= 2e-6 * scale
    assert np.abs(out - ref).max() > 0.0     # it really ran in FP32


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 5), (64, 64), (200, 70), (517, 130)])
def test_ldlt_solve(gpu_ctx, n, nrhs):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G @ G.T + n * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    LD, X = gpu_ctx.ldlt_solve(A, B)
    L = np.tril(LD, -1) + np.eye(n)
    D = np.diag(LD)
    assert np.all(D > 0)
    assert np.abs(L @ np.diag(D) @ L.T - A).max() <= 1e-12 * np.abs(A).max()
    # solve residual bar of the reference's tests/unit/test_linalg.cpp:57-89
    assert np.linalg.norm(A @ X - B) <= 1e-10 * np.linalg.norm(B)


def test_ldlt_rejects_indefinite(gpu_ctx):
    from amls_amd import AmlsError
    n = 80
    A = np.diag(np.linspace(1, 2, n))
    A[40, 40] = -1.0                        # tests/unit/test_linalg.cpp:69-78
    with pytest.raises(AmlsError) as e:
        gpu_ctx.ldlt_solve(A, np.ones((n, 1)))
    assert e.value.status == 2


@pytest.mark.parametrize("n,k", [(1, 1), (2, 2), (50, 10), (130, 130), (300, 25), (700, 40)])
def test_syevx_top(gpu_ctx, n, k):
    rng = np.random.default_rng(n + k)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    lam, V = gpu_ctx.syevx_top(A, k)
    w, U = np.linalg.eigh(A)
    ref = w[::-1][:k]
    nrm = np.abs(w).max()
    assert np.abs(lam - ref).max() <= 1e-13 * n * nrm
    assert np.abs(V.T @ V - np.eye(k)).max() <= 1e-10
    assert np.linalg.norm(A @ V - V * lam, axis=0).max() <= 1e-11 * n * nrm


def test_syevx_degenerate(gpu_ctx):
    # exactly repeated eigenvalues: vectors must still be orthonormal
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((120, 120)))
    w = np.repeat(np.arange(40, 0, -1.0), 3)
    A = (Q * w) @ Q.T
    A = (A + A.T) / 2
    lam, V = gpu_ctx.syevx_top(A, 12)
    assert np.abs(lam - np.repeat([40.0, 39, 38, 37], 3)).max() <= 1e-11 * 40
    assert np.abs(V.T @ V - np.eye(12)).max() <= 1e-9
    assert np.linalg.norm(A @ V - V * lam, axis=0).max() <= 1e-10 * 40


@pytest.mark.parametrize("scale", [1e-100, 1.0, 1e100])
def test_syevx_graded_spectrum(gpu_ctx, scale):
    # eigenvalues over 14 decades, the whole matrix scaled far from 1: the
    # Sturm counts run on a power-of-two rescaled copy, so the absolute
    # accuracy stays a few ulp of |A| at any scale
    n = 200
    rng = np.random.default_rng(17)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    w = np.logspace(0, -14, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * scale
    A = (Q * w) @ Q.T
    A = (A + A.T) / 2
    lam, V = gpu_ctx.syevx_top(A, n)
    ref = np.sort(np.linalg.eigvalsh(A))[::-1]
    assert np.abs(lam - ref).max() <= 1e-13 * n * scale
    # inverse iteration re-orthogonalises within clusters of 1e-7 |T|, so two
    # vectors just outside one overlap by up to eps |T| / (1e-7 |T|) ~ 2e-9
    # (1e-10 to 1.3e-10 measured here, varying with the order of the
    # reduction's atomic sums)
    assert np.abs(V.T @ V - np.eye(n)).max() <= 2.3e-9
    assert np.linalg.norm(A @ V - V * lam, axis=0).max() <= 1e-11 * n * scale


def test_syevx_decoupled_blocks(gpu_ctx):
    # two blocks with no coupling (zero off-diagonal in the tridiagonal form)
    # and norms 1e6 apart
    rng = np.random.default_rng(23)
    G1, G2 = rng.standard_normal((60, 60)), rng.standard_normal((40, 40))
    A = np.zeros((100, 100))
    A[:60, :60] = (G1 + G1.T) * 1e6
    A[60:, 60:] = G2 + G2.T
    lam, V = gpu_ctx.syevx_top(A, 100)
    ref = np.sort(np.linalg.eigvalsh(A))[::-1]
    nrm = np.abs(ref).max()
    assert np.abs(lam - ref).max() <= 1e-13 * 100 * nrm
    assert np.abs(V.T @ V - np.eye(100)).max() <= 1e-10
    assert np.linalg.norm(A @ V - V * lam, axis=0).max() <= 1e-11 * 100 * nrm


def test_autotune_split_table():
    """The autotuner times the GEMM's split-k choices per shape and installs
    the fastest; long-k products then take the split of the nearest row.
    Products under the tuned table, and under single-row tables forcing 1
    and 32 pieces, match numpy at 1e-12."""
    from amls_amd import HipContext
    ctx = HipContext(0)
    tab = ctx.autotune()
    assert tab.shape == (20, 3)
    assert (tab[:, 0] >= 1).all() and set(tab[:, 1]) == {8192, 32768}
    assert ((tab[:, 2] >= 1) & (tab[:, 2] <= 32)).all()
    rng = np.random.default_rng(8)
    A = rng.standard_normal((9000, 300))
    B = rng.standard_normal((9000, 200))
    ref = A.T @ B
    for table in (tab, [[15, 9000, 1]], [[15, 9000, 32]]):
        ctx.set_tuning(table)
        C = ctx.gemm(A, B, transA=True)
        assert np.abs(C - ref).max() <= 1e-12 * np.abs(ref).max()
    ctx.set_tuning([])
    ctx.close()
