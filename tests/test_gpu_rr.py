"""The final Rayleigh-Ritz problem of the subspace iteration solved by
Chebyshev-filtered subspace iteration (the library's default for one wide
problem) against the dense reduction of the same q x q matrix and against
the oracle's subspace iteration (papers/Improving_eigenpairs_of_automated_
multilevel_substructuring_with_subspace_iterations.md:384-427), through the
C-ABI."""
import os

import numpy as np
import pytest

from amls_amd import AmlsOptions, HipContext, cases
from oracle import amls_oracle as O
from tests.helpers import tree_of

pytestmark = pytest.mark.gpu


def _solve(ctx, pb, opt, mode):
    ctx.set_rr_mode(mode)
    th, Z, p, _ = ctx.solve_numeric(opt, vectors=True)
    return th.copy(), Z.copy(), p, ctx.rr_info()


def _resid(pb, th, Z):
    Kz = pb.K @ Z
    return np.linalg.norm(Kz - (pb.M @ Z) * th, axis=0) / np.linalg.norm(Kz, axis=0)


def test_filtered_rr_matches_dense_and_oracle_medium(blas1):
    """27-node hexahedra 6^3 (n = 6,084), lowest 24, eta = 10 (p = 681), two
    steps over the whole AMLS basis: the filtered iteration runs (block 80
    of 681), its Ritz residuals meet 1e-11, and the eigenvalues agree with
    the dense reduction at 1e-11 and with the oracle's subspace iteration at
    1e-9."""
    rng = np.random.default_rng(5)
    dims = (6, 6, 6)
    pb = cases.hex_elasticity(2, *dims, kscale=rng.lognormal(0, 0.3, int(np.prod(dims))))
    nev, eta = 24, 10.0
    tree, _ = tree_of(pb.K, pb.M, pb.coords, 4, 200)
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=eta, recoverVectors=False)
    _, Q = ref.cond.solve_projected(ref.p)
    z = ref.cond.backtransform(Q)
    X = np.zeros_like(z)
    X[tree.perm] = z
    t_or, _ = O.subspace_iteration(pb.K, pb.M, ref.ge, X, 2, nev, tree.perm)
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 4, 200)
    opt = AmlsOptions(nev=nev, nLevels=4, leafSize=200, thresholdFactor=eta, precision="fp64",
                      siSteps=2, siVectors=10**6, recoverVectors=False, verbosity=0)
    th_d, Z_d, p, info_d = _solve(ctx, pb, opt, 1)
    th_f, Z_f, _, info_f = _solve(ctx, pb, opt, 0)
    assert p == ref.p
    assert info_d[0] == 0
    assert 1 <= info_f[0] <= 14 and info_f[1] <= 1e-11
    assert np.abs(th_f - th_d).max() <= 1e-11 * th_d.max()
    assert (np.abs(th_f - t_or) / t_or).max() <= 1e-9
    assert np.abs(np.einsum("ij,ij->j", Z_f, pb.M @ Z_f) - 1).max() <= 1e-10
    r_f, r_d = _resid(pb, th_f, Z_f), _resid(pb, th_d, Z_d)
    # the two small-problem solvers leave eigenvector errors of ~1e-11 / gap:
    # the big-space residuals agree to ~1e-9 absolute
    assert np.all(r_f <= r_d * 1.01 + 1e-9)
    ctx.close()


def test_filtered_rr_c2_matches_dense():
    """The benchmark workload (n = 26,460, lowest 100, eta = 4, two steps,
    q = p = 1,499, block 192), mixed precision: filtered and dense final
    Rayleigh-Ritz give the same eigenvalues (within the filter's residual
    rule) and vectors of the same quality, both within 1e-6 of the FP64
    ground-truth fixture."""
    pb = cases.case_hex27_10()
    exact = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "hex27_10_lowest100.txt"))
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 6, 500)
    opt = AmlsOptions(nev=100, nLevels=6, leafSize=500, thresholdFactor=4.0, precision="mixed",
                      siSteps=2, siVectors=10**6, recoverVectors=False, verbosity=0)
    th_d, Z_d, p, info_d = _solve(ctx, pb, opt, 1)
    th_f, Z_f, _, info_f = _solve(ctx, pb, opt, 0)
    assert p == 1499 and info_d[0] == 0
    assert 1 <= info_f[0] <= 14 and info_f[1] <= 1e-11
    # the filter stops at residuals 1e-11 nu_1 of C (nu = 1 / theta): an
    # eigenvalue error of at most 1e-11 nu_1 in nu, 1e-11 theta_j^2 / theta_1
    # in theta (6e-10 at the top of this spectrum); the dense path is exact
    # to rounding
    assert np.all(np.abs(th_f - th_d) <= 1e-11 * th_d ** 2 / th_d[0] + 1e-14)
    assert O.rel_err(th_f, exact).max() <= 1e-6
    r_f, r_d = _resid(pb, th_f, Z_f), _resid(pb, th_d, Z_d)
    # the two small-problem solvers leave eigenvector errors of ~1e-11 / gap:
    # the big-space residuals agree to ~1e-9 absolute
    assert np.all(r_f <= r_d * 1.01 + 1e-9)
    ctx.close()


def test_filtered_rr_not_used_for_narrow_subspace(blas1):
    """q below four block widths (the 8-node 10^3 problem, p = 154 for the
    lowest 20): the dense reduction gives the result."""
    pb = cases.case_hex8_10()
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 300)
    opt = AmlsOptions(nev=20, nLevels=3, leafSize=300, thresholdFactor=4.0, siSteps=2,
                      siVectors=10**6, recoverVectors=False, verbosity=0)
    ctx.set_rr_mode(0)
    ctx.solve_numeric(opt)
    assert ctx.rr_info() == (0, 0.0)
    ctx.close()
