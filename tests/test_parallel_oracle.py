"""The node-parallel run of the oracle (the CPU baseline) against the serial
oracle: the same truncation counts, reduced dimension and eigenvalues to
rounding, for plain two-pass AMLS and for the subspace iteration that
follows it."""
import numpy as np

from amls_amd import cases
from oracle import amls_oracle as O
from oracle.parallel import ParallelAmls, ParallelSubspaceIteration, levels_of
from tests.helpers import tree_of


def _case():
    rng = np.random.default_rng(3)
    dims = (5, 5, 6)
    return cases.hex_elasticity(1, *dims, kscale=rng.lognormal(0, 0.3, int(np.prod(dims))))


def test_levels_cover_tree_bottom_up():
    pb = _case()
    tree, _ = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    lv = levels_of(tree)
    assert sorted(p for l in lv for p in l) == list(range(tree.N))
    seen = set()
    for level in lv:
        for p in level:          # every child is in a deeper, earlier level
            kids = [c for c in range(tree.N) if tree.parent[c] == p]
            assert all(c in seen for c in kids)
        seen.update(level)
    assert lv[-1] == [tree.N - 1]


def test_parallel_amls_and_si_match_serial(blas1):
    pb = _case()
    tree, _ = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    nev = 12
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=2.0)
    with ParallelAmls(pb.K, pb.M, tree, 2) as P:
        r = P.solve(nev=nev, thresholdFactor=2.0, recoverVectors=True)
    assert r.p == ref.p and np.array_equal(r.k, ref.k)
    np.testing.assert_allclose(r.theta, ref.theta, rtol=1e-12)
    np.testing.assert_allclose(np.abs(np.einsum("ij,ij->j", r.Z, pb.M @ ref.Z)), 1.0, atol=1e-8)
    # subspace iteration from the whole basis
    ref4 = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=4.0, recoverVectors=False)
    _, Q = ref4.cond.solve_projected(ref4.p)
    z = ref4.cond.backtransform(Q)
    X = np.zeros_like(z)
    X[tree.perm] = z
    t_or, _ = O.subspace_iteration(pb.K, pb.M, ref4.ge, X, 2, nev, tree.perm)
    with ParallelAmls(pb.K, pb.M, tree, 2) as P:
        r4 = P.solve(nev=nev, thresholdFactor=4.0, recoverVectors=False)
    with ParallelSubspaceIteration(r4.ge, P.Kr, P.Mr, r4.p, 2) as S:
        th, Xs = S.run(X, 2, nev, tree.perm)
    np.testing.assert_allclose(th, t_or, rtol=1e-11)
