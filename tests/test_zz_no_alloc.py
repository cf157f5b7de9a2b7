"""The allocation-free factor-and-solve entry point (amls_hip_ldlt_solve_noalloc)
against amls_hip_ldlt_solve, through the C-ABI on device buffers: two calls in
a row, the second smaller, give the existing entry point's factor and solution
bit for bit (same kernels, same launches, no atomics); sizes beyond its
scratch are rejected before anything is launched.  The test's own device
buffers come from the HIP runtime the library itself links (the caller
allocates; the entry point does not)."""
import numpy as np
import pytest

from amls_amd import AmlsError, HipContext, _lib
from tests.helpers import DevBuf, hip_runtime

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n), rng.standard_normal((n, 3 + n // 8))


def test_noalloc_ldlt_solve_matches_ldlt_solve():
    L = _lib.lib()
    H = hip_runtime()
    ctx = HipContext(0)
    for n, seed in ((700, 1), (260, 2)):
        A, B = spd(n, seed)
        nrhs = B.shape[1]
        LD_ref, X_ref = ctx.ldlt_solve(A, B)
        dA, dB = DevBuf(H, A), DevBuf(H, B)
        ctx._chk(L.amls_hip_ldlt_solve_noalloc(ctx._h, n, dA.ptr, nrhs, dB.ptr))
        LD, X = dA.get(), dB.get()
        dA.free()
        dB.free()
        assert np.array_equal(LD, LD_ref), n
        assert np.array_equal(X, X_ref), n
        assert np.linalg.norm(A @ X - B) <= 1e-10 * np.linalg.norm(B)
    # beyond the scratch: rejected with a clear error (the buffers are of the
    # stated size, so nothing could run out of bounds even if it were not)
    big = 5000
    dA, dB = DevBuf(H, np.eye(big)), DevBuf(H, np.zeros((big, big)))
    st = L.amls_hip_ldlt_solve_noalloc(ctx._h, big, dA.ptr, big, dB.ptr)
    dA.free()
    dB.free()
    assert st == 1
    assert "scratch" in L.amls_hip_last_error(ctx._h).decode()
    # a non-SPD matrix: the pivot check, as in amls_hip_ldlt_solve
    A, B = spd(200, 3)
    A[5, 5] = -1.0
    dA, dB = DevBuf(H, A), DevBuf(H, B)
    with pytest.raises(AmlsError) as e:
        ctx._chk(L.amls_hip_ldlt_solve_noalloc(ctx._h, 200, dA.ptr, B.shape[1], dB.ptr))
    dA.free()
    dB.free()
    assert e.value.status == 2
    ctx.close()
