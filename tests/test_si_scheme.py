"""The device's subspace-iteration scheme restated on the CPU, against the
oracle's subspace iteration (Rayleigh-Ritz after every step).  The device
applies K^-1 M through the AMLS factors, scales the columns to unit M-norm
after each step and does one Rayleigh-Ritz at the end; on the whole AMLS
basis the scaled columns turn towards the dominant modes and cond(Y^T M Y)
grows ~1e4 per step, so the projected stiffness stops being numerically
definite after a few steps.  An M-orthonormalisation (Cholesky QR) after
every second step that is not the last keeps it definite, and the final
Ritz values then equal the oracle's, whose Rayleigh-Ritz after every step
spans the same subspace."""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from amls_amd import cases
from oracle import amls_oracle as O
from tests.helpers import tree_of


def setup():
    rng = np.random.default_rng(31)
    dims = (5, 5, 6)
    ne = int(np.prod(dims))
    pb = cases.hex_elasticity(1, *dims, kscale=rng.lognormal(0, 0.3, ne),
                              rscale=rng.uniform(0.5, 2.0, ne))
    tree, _ = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    ref = O.amls_solve(pb.K, pb.M, tree, nev=12, thresholdFactor=4.0, recoverVectors=False)
    _, Qr = ref.cond.solve_projected(ref.p)
    z = ref.cond.backtransform(Qr)
    X = np.zeros_like(z)
    X[tree.perm] = z
    return pb, tree, ref, X


def device_scheme(pb, tree, ref, X, steps, orth):
    perm = tree.perm
    Kr = sp.csr_matrix(pb.K)[perm][:, perm]
    Mr = sp.csr_matrix(pb.M)[perm][:, perm]
    Y = X[perm]
    MY = Mr @ Y
    conds = []
    for it in range(steps):
        Y = O.amls_kinv(ref.ge, MY)
        MY = Mr @ Y
        s = 1.0 / np.sqrt(np.einsum("ij,ij->j", Y, MY))
        Y, MY = Y * s, MY * s
        G = Y.T @ MY
        ev = np.linalg.eigvalsh(0.5 * (G + G.T))
        conds.append(ev.max() / ev.min() if ev.min() > 0 else np.inf)
        if orth and it + 1 < steps and (it + 1) % 2 == 0:
            L = np.linalg.cholesky(0.5 * (G + G.T))
            Y = sla.solve_triangular(L, Y.T, lower=True).T
            MY = Mr @ Y
    Kc = Y.T @ (Kr @ Y)
    Mc = Y.T @ MY
    Kc, Mc = 0.5 * (Kc + Kc.T), 0.5 * (Mc + Mc.T)
    return Kc, Mc, conds


def test_column_scaling_alone_loses_definiteness(blas1):
    pb, tree, ref, X = setup()
    Kc, Mc, conds = device_scheme(pb, tree, ref, X, 6, orth=False)
    assert conds[1] > 1e3 and conds[2] > 1e7
    assert not np.isfinite(conds[5]) or conds[5] > 1e15
    try:
        np.linalg.cholesky(Kc)
        definite = True
    except np.linalg.LinAlgError:
        definite = False
    assert not definite


def test_cholesky_qr_every_second_step_matches_oracle(blas1):
    pb, tree, ref, X = setup()
    t_or, _, hist = O.subspace_iteration(pb.K, pb.M, ref.ge, X, 8, 12, tree.perm, every_step=True)
    for steps in (2, 3, 4, 5, 8):
        Kc, Mc, conds = device_scheme(pb, tree, ref, X, steps, orth=True)
        assert max(conds) < 1e9, (steps, conds)
        np.linalg.cholesky(Kc)
        th = sla.eigh(Kc, Mc, eigvals_only=True, subset_by_index=[0, 11])
        rel = np.abs(th - hist[steps - 1]) / hist[steps - 1]
        assert rel.max() <= 1e-10, (steps, rel.max())
    exact = O.reference_eigs(pb.K, pb.M, 12)
    # more steps, closer to the exact values, always from above
    errs = [O.rel_err(h, exact).max() for h in hist]
    assert all(np.all(h >= exact * (1 - 1e-12)) for h in hist)
    assert errs[7] < errs[1] < errs[0]
