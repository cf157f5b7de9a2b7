"""Discreteness of the truncation: pinned per-node counts against the
oracle's pinned run, and the truncation margin (the nearest local eigenvalue
to lambda_cut on either side of the cut) against the oracle's local spectra,
through the C-ABI (amls_hip_spectral_pinned, amls_hip_truncation_margin)."""
import numpy as np
import pytest
import scipy.linalg as sla

from amls_amd import AmlsError, AmlsOptions, HipContext, cases
from oracle import amls_oracle as O
from tests.helpers import tree_of

pytestmark = pytest.mark.gpu


def small_hex(seed=13, dims=(4, 4, 5)):
    rng = np.random.default_rng(seed)
    ne = dims[0] * dims[1] * dims[2]
    return cases.hex_elasticity(1, *dims, kscale=rng.lognormal(0, 0.3, ne),
                                rscale=rng.uniform(0.5, 2.0, ne))


def oracle_ge(pb, levels, leaf):
    tree, _ = tree_of(pb.K, pb.M, pb.coords, levels, leaf)
    perm = tree.perm
    return tree, O.BlockGE(pb.K.tocsr()[perm][:, perm], pb.M.tocsr()[perm][:, perm], tree).run()


def oracle_margin(ge, N, cut):
    """min over nodes of min(|w_k - cut|, |w_k+1 - cut|) / cut from the dense
    local spectra (w ascending, k = #{w <= cut}), and its node."""
    best, node = np.inf, -1
    for p in range(N):
        w = sla.eigh(ge.Khat[p], ge.Mhat[p], eigvals_only=True)
        kk = int(np.sum(w <= cut))
        cand = [abs(w[kk - 1] - cut) / cut] if kk > 0 else []
        if kk < w.size:
            cand.append(abs(w[kk] - cut) / cut)
        if cand and min(cand) < best:
            best, node = min(cand), p
    return best, node


def test_pinned_counts_match_oracle(blas1):
    pb = small_hex()
    tree, ge = oracle_ge(pb, 3, 40)
    base = O.Condensation(ge).run("threshold", 0.6)
    # counts no threshold gives: one more mode on every other node, and one
    # count beyond its node's order (clamped)
    kpin = base.k.copy()
    kpin[::2] += 1
    kpin[0] = tree.size[0] + 5
    want = np.minimum(kpin, tree.size)
    ora = O.Condensation(ge).run("fixed", 0, pinned_k=kpin)
    assert np.array_equal(ora.k, want)
    ctx = HipContext(0)
    ctx.set_sil_mode(-1)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    ctx.ge("fp64")
    ctx.spectral_prepare()
    k = ctx.spectral_pinned(kpin)
    assert np.array_equal(k, want)
    assert ctx.truncation_margin() == (np.inf, -1)
    p = ctx.project("fp64")
    assert p == ora.p == int(want.sum())
    th = ctx.reduced_solve(12, False)
    ref = ora.solve_projected(12, vectors=False)
    assert np.abs(th - ref).max() <= 1e-10 * ref.max()
    with pytest.raises(AmlsError):
        ctx.spectral_pinned(-np.ones(tree.N, dtype=np.int64))
    ctx.close()


def test_truncation_margin_matches_oracle(blas1):
    pb = small_hex()
    tree, ge = oracle_ge(pb, 3, 40)
    ctx = HipContext(0)
    ctx.set_sil_mode(-1)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    ctx.ge("fp64")
    ctx.spectral_prepare()
    for cut in (0.6, 2.5):
        ctx.spectral("threshold", cut)
        m, node = ctx.truncation_margin()
        mo, no = oracle_margin(ge, tree.N, cut)
        assert node == no, (cut, node, no)
        assert abs(m - mo) <= 1e-9 * max(1.0, mo), (cut, m, mo)
    ctx.spectral("fixed", 4)
    assert ctx.truncation_margin() == (np.inf, -1)
    ctx.close()


def test_margin_after_a_solve_is_pass_b(blas1):
    """After a two-pass solve the margin is that of pass B's cut
    (eta times the nev-th pass-A eigenvalue)."""
    pb = small_hex(seed=21, dims=(5, 4, 6))
    tree, ge = oracle_ge(pb, 3, 40)
    ref = O.amls_solve(pb.K, pb.M, tree, nev=10, thresholdFactor=2.0, recoverVectors=False,
                       sil_mode=-1)
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    opt = AmlsOptions(nev=10, nLevels=3, leafSize=40, thresholdFactor=2.0, silMode=-1,
                      recoverVectors=False, verbosity=0)
    th, _, p, _ = ctx.solve_numeric(opt)
    assert p == ref.p
    m, node = ctx.truncation_margin()
    mo, no = oracle_margin(ge, tree.N, ref.lambda_cut)
    assert node == no
    assert abs(m - mo) <= 1e-8 * max(1.0, mo)
    ctx.close()
