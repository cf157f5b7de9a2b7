"""Batched multi-problem solves (config C4: many independent small pencils in
one forest, every level one launch per kernel class across problems), through
the C-ABI: against single-problem GPU solves of the same pencils, against the
oracle on each problem's tree, and on the 64-problem workload."""
import time

import numpy as np
import pytest

from amls_amd import AmlsOptions, AmlsSolver, HipContext, Pencil, cases
from oracle import amls_oracle as O
from tests.helpers import tree_of

pytestmark = pytest.mark.gpu


def small_batch(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny, nz = (int(v) for v in rng.integers(3, 7, size=3))
        ne = nx * ny * nz
        out.append(cases.hex_elasticity(1, nx, ny, nz, kscale=rng.lognormal(0, 0.5, ne),
                                        rscale=rng.uniform(0.5, 2.0, ne)))
    return out


def resid(pb, th, Z):
    KZ = pb.K @ Z
    return np.linalg.norm(KZ - (pb.M @ Z) * th, axis=0) / np.linalg.norm(KZ, axis=0)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_plain_amls_matches_single_and_oracle(blas1, prec):
    pbs = small_batch(5, 6)
    nev = 8
    opt = AmlsOptions(nev=nev, nLevels=3, leafSize=40, precision=prec, verbosity=0)
    sols = AmlsSolver(0).solve_batch([Pencil(p.K, p.M, p.coords) for p in pbs], opt)
    assert len(sols) == len(pbs)
    for pb, sb in zip(pbs, sols):
        ss = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords), opt)
        assert sb.p == ss.p
        assert np.abs(sb.theta - ss.theta).max() <= 1e-12 * ss.theta.max()
        assert np.abs(np.einsum("ij,ij->j", sb.Z, pb.M @ sb.Z) - 1).max() <= 1e-10
        s = np.sign(np.einsum("ij,ij->j", sb.Z, pb.M @ ss.Z))
        assert np.abs(sb.Z * s - ss.Z).max() <= 1e-8 * np.abs(ss.Z).max()
        if prec == "fp64":
            tree, _ = tree_of(pb.K, pb.M, pb.coords, 3, 40)
            ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, recoverVectors=False)
            assert sb.p == ref.p
            assert np.abs(sb.theta - ref.theta).max() <= 1e-10 * ref.theta.max()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_subspace_iteration(blas1, prec):
    pbs = small_batch(6, 5)
    nev, q = 8, 16
    opt = AmlsOptions(nev=nev, nLevels=3, leafSize=40, precision=prec, siSteps=2,
                      siVectors=q, verbosity=0)
    sols = AmlsSolver(0).solve_batch([Pencil(p.K, p.M, p.coords) for p in pbs], opt)
    plain = AmlsSolver(0).solve_batch([Pencil(p.K, p.M, p.coords) for p in pbs],
                                      AmlsOptions(nev=nev, nLevels=3, leafSize=40, verbosity=0))
    for pb, sb, sp_ in zip(pbs, sols, plain):
        assert sb.theta.shape == (nev,)
        ss = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords), opt)
        assert np.abs(sb.theta - ss.theta).max() <= (1e-10 if prec == "fp64" else 1e-7) * ss.theta.max()
        exact = O.reference_eigs(pb.K, pb.M, nev)
        assert np.all(sb.theta >= exact * (1 - 1e-12))
        assert O.rel_err(sb.theta, exact).max() < O.rel_err(sp_.theta, exact).max()
        assert np.abs(np.einsum("ij,ij->j", sb.Z, pb.M @ sb.Z) - 1).max() <= 1e-10
        assert np.all(resid(pb, sb.theta, sb.Z) <= resid(pb, ss.theta, ss.Z) * 1.1 + 1e-12)


def test_batch_numeric_entry_points_size_outputs_per_problem(blas1):
    """set_problem_batch + solve_numeric return count x nev values and count
    reduced sizes (the output buffers are sized per problem), the numbers
    solve_batch gives."""
    pbs = small_batch(7, 5)
    nev = 6
    opt = AmlsOptions(nev=nev, nLevels=3, leafSize=40, precision="mixed", siSteps=2,
                      siVectors=12, recoverVectors=False, verbosity=0)
    pens = [Pencil(p.K, p.M, p.coords) for p in pbs]
    sols = AmlsSolver(0).solve_batch(pens, opt)
    ctx = HipContext(0)
    ctx.set_problem_batch(pens, 3, 40)
    th, Z, p, _ = ctx.solve_numeric(opt, vectors=True)
    assert th.shape == (len(pbs), nev) and p.shape == (len(pbs),)
    assert Z.shape == (sum(pb.K.shape[0] for pb in pbs), nev)
    rows, _ = ctx.batch_info()
    for s, sb in enumerate(sols):
        assert p[s] == sb.p
        assert np.array_equal(th[s], sb.theta)
        assert np.array_equal(Z[rows[s]:rows[s + 1]], sb.Z)
    ctx.close()


def test_batch64_workload(blas1):
    """The 64-problem workload (8-node hexes, nx, ny, nz ~ U{6..16},
    E ~ lognormal(0, 0.5), rho ~ U[0.5, 2], seed 4321), nev 20 each, mixed
    precision with 2 subspace steps; three problems re-solved alone must agree."""
    pbs = cases.case_batch64()
    assert len(pbs) == 64
    nev = 20
    opt = AmlsOptions(nev=nev, nLevels=6, leafSize=500, thresholdFactor=4.0, precision="mixed",
                      siSteps=2, siVectors=2 * nev, verbosity=0)
    pens = [Pencil(p.K, p.M, p.coords) for p in pbs]
    s = AmlsSolver(0)
    s.solve_batch(pens, opt)                      # warm-up
    t0 = time.perf_counter()
    sols = s.solve_batch(pens, opt)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in (0, 17, 63):
        ss = s.solve(pens[i], opt)
        assert np.abs(sols[i].theta - ss.theta).max() <= 1e-7 * ss.theta.max()
    ts = (time.perf_counter() - t0) / 3 * 64
    print(f"batch64: {tb:.3f} s batched vs ~{ts:.3f} s as 64 single solves "
          f"(n total {sum(p.n for p in pbs)})")
    # two steps on q = 2 nev vectors: the 20th mode's residual is ~1e-2
    for pb, sb in zip(pbs, sols):
        r = resid(pb, sb.theta, sb.Z)
        assert r[:10].max() < 1e-2 and r.max() < 5e-2
        assert np.all(np.diff(sb.theta) >= -1e-12 * sb.theta.max())


def test_batch64_fp64_plain_matches_oracle(blas1):
    """The 64-problem workload in the all-FP64 form: plain batched AMLS on all
    64 problems against the oracle's FP64 AMLS on each problem's own tree
    (the smallest, the largest and two in between), and M-normalised vectors
    everywhere."""
    pbs = cases.case_batch64()
    nev = 20
    opt = AmlsOptions(nev=nev, nLevels=6, leafSize=500, thresholdFactor=4.0, precision="fp64",
                      siSteps=0, verbosity=0)
    sols = AmlsSolver(0).solve_batch([Pencil(p.K, p.M, p.coords) for p in pbs], opt)
    sizes = np.array([p.n for p in pbs])
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 40})
    for i in picks:
        pb = pbs[i]
        tree, _ = tree_of(pb.K, pb.M, pb.coords, 6, 500)
        ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=4.0, recoverVectors=False)
        assert sols[i].p == ref.p, i
        assert (np.abs(sols[i].theta - ref.theta) / ref.theta).max() <= 1e-9, i
    for pb, sb in zip(pbs, sols):
        assert np.abs(np.einsum("ij,ij->j", sb.Z, pb.M @ sb.Z) - 1).max() <= 1e-10
        # plain AMLS at eta 4: the vectors' residuals are large (up to ~0.64
        # measured), which is what the subspace iteration is for
        assert resid(pb, sb.theta, sb.Z).max() < 1.0
