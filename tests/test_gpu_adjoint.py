"""Eigenvalue adjoint (d theta / d element scale) on the GPU, through the
C-ABI: against the oracle's formula on the same vectors, against central
finite differences of GPU solves, through a torch.autograd wrapper, and at
the 27-node 14^3 size (n = 70,644) through size-independent identities:
sum_e s_e d theta/d s_e = x^T K x = theta and sum_e r_e d theta/d r_e =
-theta x^T M x = -theta for M-normalised Rayleigh-Ritz vectors."""
import numpy as np
import pytest

from amls_amd import AmlsOptions, AmlsSolver, Pencil, cases
from oracle import amls_oracle as O

pytestmark = pytest.mark.gpu


def problem(ks, rs, dims):
    return cases.hex_elasticity(1, *dims, kscale=ks, rscale=rs)


def exact_opts(nev):
    # keep every local mode: the reduced problem is the whole pencil, so the
    # eigenpairs are exact to rounding (the FD checks need that)
    return AmlsOptions(nev=nev, nLevels=2, leafSize=40, fixedKeep=100000, verbosity=0)


def params(seed, ne):
    rng = np.random.default_rng(seed)
    return rng.lognormal(0, 0.3, ne), rng.uniform(0.5, 2.0, ne)


def test_adjoint_kernel_matches_oracle(blas1):
    dims = (4, 4, 5)
    ks, rs = params(1, 80)
    pb = problem(ks, rs, dims)
    s = AmlsSolver(0)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords),
                  AmlsOptions(nev=12, nLevels=3, leafSize=40, siSteps=2, verbosity=0))
    gs_dev, gr_dev = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta)      # device vectors
    gs_h, gr_h = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta, sol.Z)  # host vectors
    ref_s, ref_r = O.eig_adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta, sol.Z)
    for g, r in ((gs_dev, ref_s), (gs_h, ref_s), (gr_dev, ref_r), (gr_h, ref_r)):
        assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()


def test_adjoint_matches_finite_differences(blas1):
    """20 random parameters (seed 20), central differences h = 1e-6 relative."""
    dims = (3, 3, 4)
    ne = 36
    ks, rs = params(2, ne)
    nev = 6
    pb = problem(ks, rs, dims)
    s = AmlsSolver(0)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords), exact_opts(nev))
    gs, gr = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta)
    exact = O.reference_eigs(pb.K, pb.M, nev)
    assert np.abs(sol.theta - exact).max() <= 1e-10 * exact.max()
    rng = np.random.default_rng(cases.SEEDS["fd_params"])
    picks = rng.choice(2 * ne, size=20, replace=False)
    h = 1e-6
    for q in picks:
        which, e = divmod(int(q), ne)
        vals = []
        for sgn in (1, -1):
            k2, r2 = ks.copy(), rs.copy()
            (k2 if which == 0 else r2)[e] *= 1 + sgn * h
            p2 = problem(k2, r2, dims)
            vals.append(AmlsSolver(0).solve(Pencil(p2.K, p2.M, p2.coords), exact_opts(nev)).theta)
        base = ks[e] if which == 0 else rs[e]
        fd = (vals[0] - vals[1]) / (2 * h * base)
        g = gs[e] if which == 0 else gr[e]
        assert np.abs(fd - g).max() <= 1e-5 * np.abs(sol.theta).max(), (which, e)


def test_adjoint_torch_autograd(blas1):
    """A torch.autograd.Function whose forward is a GPU solve and whose
    backward is the adjoint kernel passes torch.autograd.gradcheck."""
    torch = pytest.importorskip("torch")
    dims = (3, 3, 3)
    ne = 27
    ks0, rs0 = params(3, ne)
    nev = 4
    sel = np.arange(0, ne, 4)      # a subset of elements carries the parameters

    class Eig(torch.autograd.Function):
        @staticmethod
        def forward(ctx, sk, sr):
            ks, rs = ks0.copy(), rs0.copy()
            ks[sel] = sk.detach().cpu().numpy()
            rs[sel] = sr.detach().cpu().numpy()
            pb = problem(ks, rs, dims)
            s = AmlsSolver(0)
            sol = s.solve(Pencil(pb.K, pb.M, pb.coords), exact_opts(nev))
            gs, gr = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta)
            ctx.save_for_backward(torch.from_numpy(gs[sel].copy()), torch.from_numpy(gr[sel].copy()))
            return torch.from_numpy(sol.theta.copy())

        @staticmethod
        def backward(ctx, gth):
            gs, gr = ctx.saved_tensors
            return gs @ gth, gr @ gth

    sk = torch.tensor(ks0[sel], dtype=torch.float64, requires_grad=True)
    sr = torch.tensor(rs0[sel], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(Eig.apply, (sk, sr), eps=1e-6, atol=1e-7, rtol=1e-4)


def test_adjoint_hex27_14_identities(blas1):
    """The 27-node 14 x 14 x 14 workload (n = 70,644, 2,744 stiffness and
    2,744 density scales, 50 modes): the eigenvalues within 1e-6 of the FP64
    ground truth, and 50 x 5,488 gradients."""
    pb = cases.case_hex27_14()
    assert pb.n == 70644
    s = AmlsSolver(0)
    # three subspace steps: two leave the 50th mode at 7e-6 here (its gap to
    # the first discarded AMLS mode is smaller than on the 10^3 workload)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords),
                  AmlsOptions(nev=50, nLevels=6, leafSize=500, thresholdFactor=4.0,
                              precision="mixed", siSteps=3, siVectors=10 ** 6, verbosity=0))
    gs, gr = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta)
    assert gs.shape == (2744, 50) and gr.shape == (2744, 50)
    th = sol.theta
    # the eigenvalues against the committed FP64 ground truth (shift-invert
    # Lanczos, tol 1e-12; tests/golden/make_fixtures.py)
    import os
    exact = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "hex27_14_lowest50.txt"))
    assert O.rel_err(th, exact).max() <= 1e-6
    assert np.all(th >= exact * (1 - 1e-12))
    assert np.abs((pb.kscale[:, None] * gs).sum(0) - th).max() <= 1e-9 * th.max()
    assert np.abs((pb.rscale[:, None] * gr).sum(0) + th).max() <= 1e-9 * th.max()
    assert np.all(gs >= -1e-14 * th.max()) and np.all(gr <= 1e-14 * th.max())
    ref_s, ref_r = O.eig_adjoint(pb.conn, pb.Ke0, pb.Me0, th, sol.Z)
    assert np.abs(gs - ref_s).max() <= 1e-12 * np.abs(ref_s).max()
    assert np.abs(gr - ref_r).max() <= 1e-12 * np.abs(ref_r).max()


def test_adjoint_with_shift_matches_finite_differences(blas1):
    """A shifted solve (K <- K + shift M) returns theta = lambda + shift; its
    adjoint is d theta / d r_e = -(theta - shift) x^T Me0 x (the shift term
    moves with r_e too), checked by central differences of shifted solves on
    8 parameters and against the oracle's formula with the same shift."""
    dims = (3, 3, 4)
    ne = 36
    ks, rs = params(4, ne)
    nev, shift = 5, 0.23
    pb = problem(ks, rs, dims)
    opts = exact_opts(nev)
    opts.shift = shift
    s = AmlsSolver(0)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords), opts)
    exact = O.reference_eigs(pb.K, pb.M, nev) + shift
    assert np.abs(sol.theta - exact).max() <= 1e-10 * exact.max()
    gs, gr = s.ctx.adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta)
    ref_s, ref_r = O.eig_adjoint(pb.conn, pb.Ke0, pb.Me0, sol.theta, sol.Z, shift=shift)
    assert np.abs(gs - ref_s).max() <= 1e-12 * np.abs(ref_s).max()
    assert np.abs(gr - ref_r).max() <= 1e-12 * np.abs(ref_r).max()
    rng = np.random.default_rng(7)
    h = 1e-6
    for q in rng.choice(2 * ne, size=8, replace=False):
        which, e = divmod(int(q), ne)
        vals = []
        for sgn in (1, -1):
            k2, r2 = ks.copy(), rs.copy()
            (k2 if which == 0 else r2)[e] *= 1 + sgn * h
            p2 = problem(k2, r2, dims)
            vals.append(AmlsSolver(0).solve(Pencil(p2.K, p2.M, p2.coords), opts).theta)
        base = ks[e] if which == 0 else rs[e]
        fd = (vals[0] - vals[1]) / (2 * h * base)
        g = gs[e] if which == 0 else gr[e]
        assert np.abs(fd - g).max() <= 1e-5 * np.abs(sol.theta).max(), (which, e)
