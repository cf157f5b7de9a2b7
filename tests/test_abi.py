"""The C-ABI library loads without a GPU and exports every declared symbol;
host-only contexts run the analysis and report errors as status codes."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from amls_amd import AmlsError, HipContext, _lib, cases

HEADER = Path(__file__).resolve().parents[1] / "include" / "amls_hip.h"


def declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"\b(amls_(?:hip_)?[a-z_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == names


def test_options_default_matches_reference():
    o = _lib.AmlsOptionsC()
    assert _lib.lib().amls_options_default(ctypes.byref(o)) == 0
    # src/amls/solver.hpp:18-30
    assert (o.nLevels, o.leafSize, o.thresholdFactor, o.fixedKeep, o.nev,
            o.denseProjectedMax, o.shift, o.twoPass, o.numThreads,
            o.recoverVectors, o.verbosity) == (6, 500, 2.0, 0, 100, 5000, 0.0, 1, 0, 1, 1)


def test_host_context_rejects_gpu_phase():
    K, M = cases.path_graph(30)
    ctx = HipContext(-1)
    ctx.set_problem(K, M, None, 2, 5)
    with pytest.raises(AmlsError) as e:
        ctx.ge()
    assert e.value.status == 6


def test_bad_arguments_are_status_codes():
    ctx = HipContext(-1)
    K, M = cases.path_graph(10)
    bad = K.tocsc().copy()
    bad.indices = bad.indices.copy()
    bad.indices[3] = 99
    with pytest.raises(AmlsError) as e:
        ctx.set_problem(bad, M, None, 2, 5)
    assert e.value.status == 1
    assert "out of range" in str(e.value)


def test_malformed_inputs_are_status_codes():
    """Malformed column pointers, tree parameters, precision, truncation kind
    and subspace options come back as AMLS_E_ARG (1), checked before any
    device work (so a host-only context sees them)."""
    from amls_amd._lib import dptr, iptr
    L = _lib.lib()
    ctx = HipContext(-1)
    K, M = cases.path_graph(10)
    Kc, Mc = K.tocsc(), M.tocsc()
    ki = np.ascontiguousarray(Kc.indices, dtype=np.int64)
    kx = np.ascontiguousarray(Kc.data, dtype=np.float64)
    mp = np.ascontiguousarray(Mc.indptr, dtype=np.int64)
    mi = np.ascontiguousarray(Mc.indices, dtype=np.int64)
    mx = np.ascontiguousarray(Mc.data, dtype=np.float64)
    kp = np.ascontiguousarray(Kc.indptr, dtype=np.int64)

    def setp(colptr, nLevels=2, leafSize=5):
        return L.amls_hip_set_problem(ctx._h, 10, iptr(colptr), iptr(ki), dptr(kx), iptr(mp),
                                      iptr(mi), dptr(mx), dptr(None), nLevels, leafSize, 0.0)

    def err():
        return L.amls_hip_last_error(ctx._h).decode()

    swapped = kp.copy()
    swapped[4], swapped[5] = swapped[5], swapped[4]
    assert setp(swapped) == 1 and "decrease" in err()
    shifted = kp.copy()
    shifted[0] = 1
    assert setp(shifted) == 1 and "start at 0" in err()
    assert setp(kp, nLevels=-1) == 1 and "nLevels" in err()
    assert setp(kp, leafSize=0) == 1 and "leafSize" in err()
    assert setp(kp) == 0
    assert L.amls_hip_ge(ctx._h, 7) == 1 and "precision" in err()
    assert L.amls_hip_project(ctx._h, 2, None) == 1 and "precision" in err()
    assert L.amls_hip_spectral(ctx._h, 5, 0.0, None) == 1 and "kind" in err()
    o = _lib.AmlsOptionsC()
    L.amls_options_default(ctypes.byref(o))
    o.precision = 3
    assert L.amls_hip_solve_numeric(ctx._h, ctypes.byref(o), None, None, None, None) == 1
    o.precision = 1
    o.siSteps = -1
    assert L.amls_hip_solve_numeric(ctx._h, ctypes.byref(o), None, None, None, None) == 1
    o.siSteps = 0
    o.thresholdFactor = 0.0
    assert L.amls_hip_solve_numeric(ctx._h, ctypes.byref(o), None, None, None, None) == 1
    # valid options on a host-only context: the device is what is missing
    o.thresholdFactor = 2.0
    assert L.amls_hip_solve_numeric(ctx._h, ctypes.byref(o), None, None, None, None) == 6
    ctx.close()


def test_set_tuning_validates_rows():
    """A split table (rows of output tiles, k, split) installs on a host-only
    context; an empty table restores the built-in rule; malformed rows are
    AMLS_E_ARG."""
    ctx = HipContext(-1)
    ctx.set_tuning([[300, 26460, 4], [1024, 8192, 2]])
    ctx.set_tuning([])
    for bad in ([[0, 8192, 2]], [[10, 0, 2]], [[10, 8192, 0]], [[10, 8192, 65]]):
        with pytest.raises(AmlsError) as e:
            ctx.set_tuning(bad)
        assert e.value.status == 1
    ctx.close()


def test_operator_sizes_beyond_int32_are_rejected():
    """Sizes the kernels address with 32-bit counts come back as AMLS_E_ARG
    instead of wrapping (checked before the device is needed, so a host-only
    context sees them), as do non-positive subspace-iteration counts."""
    L = _lib.lib()
    ctx = HipContext(-1)
    big = 2 ** 31
    assert L.amls_hip_gemm(ctx._h, 0, 0, 0, big, 4, 4, 1.0, None, big, None, 4, 0.0,
                           None, big) == 1
    assert "32 bits" in L.amls_hip_last_error(ctx._h).decode()
    assert L.amls_hip_gemm(ctx._h, 0, 0, 0, 4, 4, big, 1.0, None, 4, None, big, 0.0,
                           None, 4) == 1
    assert L.amls_hip_ldlt_solve(ctx._h, big, None, 1, None, None, None) == 1
    assert "32 bits" in L.amls_hip_last_error(ctx._h).decode()
    assert L.amls_hip_syevx_top(ctx._h, big, None, 1, None, None) == 1
    assert L.amls_hip_subspace_iter(ctx._h, 2, 10, 0, None, None) == 1
    assert "nev" in L.amls_hip_last_error(ctx._h).decode()
    assert L.amls_hip_subspace_iter(ctx._h, 2, 10, -5, None, None) == 1
    assert L.amls_hip_subspace_iter(ctx._h, 0, 10, 4, None, None) == 1
    ctx.close()


def test_duplicate_entries_are_summed():
    """A CSC with a split (row, col) entry and unsorted rows describes the
    same matrix as its canonical form: the library's analysis sees the same
    pattern, and the wrapper hands over the summed entries."""
    import scipy.sparse as sp
    from amls_amd import _csc
    K, M = cases.grid_graph_2d(6, 5)
    Kc = sp.csc_matrix(K)
    rows, cols = Kc.nonzero()
    vals = np.asarray(Kc[rows, cols]).ravel()
    # split every diagonal entry in two halves, reverse each column's rows
    d = rows == cols
    R = np.concatenate([rows, rows[d]])
    Cc = np.concatenate([cols, cols[d]])
    V = np.concatenate([np.where(d, 0.5 * vals, vals), 0.5 * vals[d]])
    order = np.lexsort((-R, Cc))
    colptr = np.zeros(K.shape[0] + 1, dtype=np.int64)
    np.add.at(colptr, Cc[order] + 1, 1)
    colptr = np.cumsum(colptr)
    ri = np.ascontiguousarray(R[order], dtype=np.int64)
    rv = np.ascontiguousarray(V[order], dtype=np.float64)
    # (data, indices, indptr) keeps the duplicates as given
    Ks = sp.csc_matrix((rv.copy(), ri.copy(), colptr.copy()), shape=K.shape)
    assert Ks.nnz == Kc.nnz + K.shape[0]
    p, i, x = _csc(Ks)
    assert np.array_equal(p, Kc.indptr) and np.array_equal(i, Kc.indices)
    assert np.allclose(x, Kc.data, rtol=0, atol=1e-15)
    # raw arrays with duplicates straight through the C-ABI
    from amls_amd._lib import dptr, iptr
    mp, mi, mx = _csc(M)
    L = _lib.lib()
    trees = []
    for (cp, rr, vv) in ((colptr, ri, rv), _csc(K)):
        ctx = HipContext(-1)
        st = L.amls_hip_set_problem(ctx._h, K.shape[0], iptr(cp), iptr(rr), dptr(vv), iptr(mp),
                                    iptr(mi), dptr(mx), dptr(None), 2, 8, 0.0)
        assert st == 0
        trees.append(ctx.tree())
        ctx.close()
    for key in ("treeId", "offset", "size", "parent", "bsize", "perm"):
        assert np.array_equal(trees[0][key], trees[1][key]), key


def test_sil_mode_entry_points():
    """The local solver switch takes -1, 0 or 1; the per-node paths exist
    only after a spectral call (host-only context: state error)."""
    L = _lib.lib()
    ctx = HipContext(-1)
    for m in (-1, 0, 1):
        ctx.set_sil_mode(m)
    for bad in (2, -2):
        with pytest.raises(AmlsError) as e:
            ctx.set_sil_mode(bad)
        assert e.value.status == 1
    assert L.amls_hip_get_sil(ctx._h, None, None) == 6
    o = _lib.AmlsOptionsC()
    L.amls_options_default(ctypes.byref(o))
    assert o.silMode == 0
    o.silMode = 3
    assert L.amls_hip_solve_numeric(ctx._h, ctypes.byref(o), None, None, None, None) == 1
    ctx.close()


def test_noalloc_size_limit_on_host_context():
    """The allocation-free factor-and-solve checks its scratch limit before
    anything else touches the device: an order whose panel buffer alone
    exceeds the 8 MiB scratch is AMLS_E_ARG; a fitting one needs a device."""
    L = _lib.lib()
    ctx = HipContext(-1)
    assert L.amls_hip_ldlt_solve_noalloc(ctx._h, 20000, None, 1, None) == 1
    assert "scratch" in L.amls_hip_last_error(ctx._h).decode()
    assert L.amls_hip_ldlt_solve_noalloc(ctx._h, 0, None, 1, None) == 1
    buf = np.zeros(16)
    p = buf.ctypes.data
    assert L.amls_hip_ldlt_solve_noalloc(ctx._h, 4, p, 1, p) == 6
    ctx.close()


def test_rr_mode_arguments():
    from amls_amd import AmlsError, HipContext
    c = HipContext(-1)
    with pytest.raises(AmlsError, match="rr mode"):
        c.set_rr_mode(2)
    c.set_rr_mode(1)
    assert c.rr_info() == (0, 0.0)
    c.close()
