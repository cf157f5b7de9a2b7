"""Phase-by-phase and end-to-end parity of the GPU AMLS path with the CPU
oracle on the same separator tree (libamls_hip's nested dissection)."""
import numpy as np
import pytest
import scipy.linalg as sla

from amls_amd import AmlsOptions, AmlsSolver, HipContext, Pencil, cases
from oracle import amls_oracle as O
from tests.helpers import grid_coords, tree_of

pytestmark = pytest.mark.gpu


def small_hex(seed=11, dims=(4, 4, 5), order=1):
    rng = np.random.default_rng(seed)
    ne = dims[0] * dims[1] * dims[2]
    return cases.hex_elasticity(order, *dims, kscale=rng.lognormal(0, 0.3, ne),
                                rscale=rng.uniform(0.5, 2.0, ne))


def gpu_phases(K, M, coords, nLevels, leafSize, kind, value, want, prec="fp64"):
    ctx = HipContext(0)
    ctx.set_problem(K, M, coords, nLevels, leafSize)
    ctx.ge(prec)
    ctx.spectral_prepare()
    k = ctx.spectral(kind, value)
    p = ctx.project(prec)
    th = ctx.reduced_solve(want, True)
    Z = ctx.backtransform(min(want, p))
    return ctx, k, p, th, Z


def test_full_keep_exactness_grid(blas1):
    """tests/unit/test_amls.cpp:110-150: keeping every local mode reproduces
    the closed-form spectrum of a 10x10 grid, back-transform residual < 1e-10."""
    K, M = cases.grid_graph_2d(10, 10)
    ctx, k, p, th, Z = gpu_phases(K, M, grid_coords(10, 10), 2, 20, "fixed", 0, 10)
    assert p == 100
    ref = cases.grid_graph_2d_spectrum(10, 10)[:10]
    assert np.all(np.abs(th - ref) <= 1e-10 * np.maximum(1.0, np.abs(ref)))
    Kd = K.toarray()
    for j in range(3):
        z = Z[:, j]
        r = Kd @ z - th[j] * (M @ z)
        assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(Kd @ z)


def test_ge_blocks_match_oracle(blas1):
    pb = small_hex()
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 40)
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    ctx.ge("fp64")
    perm = tree.perm
    ge = O.BlockGE(pb.K.tocsr()[perm][:, perm], pb.M.tocsr()[perm][:, perm], tree).run()
    for p in range(tree.N):
        n_p = int(tree.size[p])
        LD, Mh = ctx.node_factor(p, n_p)
        L = np.tril(LD, -1) + np.eye(n_p)
        Kh = L @ np.diag(np.diag(LD)) @ L.T
        Mh = np.tril(Mh) + np.tril(Mh, -1).T
        sk = np.abs(ge.Khat[p]).max()
        sm = np.abs(ge.Mhat[p]).max()
        assert np.abs(Kh - ge.Khat[p]).max() <= 1e-11 * sk, p
        assert np.abs(Mh - ge.Mhat[p]).max() <= 1e-11 * sm, p


def test_local_spectra_match_oracle(blas1):
    pb = small_hex(seed=12)
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 40)
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    ctx.ge("fp64")
    ctx.spectral_prepare()
    kfix = 7
    k = ctx.spectral("fixed", kfix)
    perm = tree.perm
    ge = O.BlockGE(pb.K.tocsr()[perm][:, perm], pb.M.tocsr()[perm][:, perm], tree).run()
    for p in range(tree.N):
        n_p = int(tree.size[p])
        assert k[p] == min(kfix, n_p)
        mu, Z = ctx.local(p, n_p, int(k[p]))
        w, V = sla.eigh(ge.Khat[p], ge.Mhat[p], subset_by_index=[0, k[p] - 1])
        assert np.abs(mu - w).max() <= 1e-10 * w.max(), p
        # M-orthonormal, eigen-residual, and the same subspace as LAPACK's
        Mh, Kh = ge.Mhat[p], ge.Khat[p]
        assert np.abs(Z.T @ Mh @ Z - np.eye(k[p])).max() <= 1e-9
        assert np.linalg.norm(Kh @ Z - Mh @ Z * mu) <= 1e-9 * np.linalg.norm(Kh @ Z)
        C = np.abs(Z.T @ Mh @ V)
        assert np.abs(np.linalg.svd(C, compute_uv=False) - 1).max() <= 1e-8


def test_threshold_counts_and_reduced(blas1):
    pb = small_hex(seed=13)
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 40)
    perm = tree.perm
    ge = O.BlockGE(pb.K.tocsr()[perm][:, perm], pb.M.tocsr()[perm][:, perm], tree).run()
    cut = 0.6
    cond = O.Condensation(ge).run("threshold", cut)
    ctx, k, p, th, Z = gpu_phases(pb.K, pb.M, pb.coords, 3, 40, "threshold", cut, 12)
    assert np.array_equal(k, cond.k)
    assert p == cond.p
    ref = cond.solve_projected(12, vectors=False)
    assert np.abs(th - ref).max() <= 1e-10 * ref.max()
    # M-tilde blocks: same up to the sign of each local mode
    Mt, moff = ctx.reduced(p, tree.N)
    Mt = np.tril(Mt) + np.tril(Mt, -1).T
    Kp, Mp = cond.assemble()
    # map post-order mode layout -> oracle (node-id) layout, aligning signs
    idx = np.zeros(p, dtype=np.int64)
    sgn = np.ones(p)
    for q in range(tree.N):
        kq = int(k[q])
        if kq == 0:
            continue
        _, Zg = ctx.local(q, int(tree.size[q]), kq)
        s = np.sign(np.einsum("ij,ij->j", Zg, ge.Mhat[q] @ cond.Z[q]))
        idx[moff[q]:moff[q] + kq] = cond.projOffset[q] + np.arange(kq)
        sgn[moff[q]:moff[q] + kq] = s
    Mo = Mp[np.ix_(idx, idx)] * np.outer(sgn, sgn)
    assert np.abs(Mt - Mo).max() <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_solve_matches_oracle_hex8(blas1, prec):
    pb = small_hex(seed=21, dims=(5, 5, 6))
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    nev = 12
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev)
    sol = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords),
                              AmlsOptions(nev=nev, nLevels=3, leafSize=60, precision=prec,
                                          verbosity=0))
    assert sol.p == ref.p
    rel = np.abs(sol.theta - ref.theta) / ref.theta
    # FP64: same arithmetic up to rounding.  Mixed: FP32 Schur / congruence /
    # projection GEMMs perturb the reduced pencil at ~1e-7 relative.
    assert rel.max() <= (1e-10 if prec == "fp64" else 5e-5)
    Mz = pb.M @ sol.Z
    assert np.abs(np.einsum("ij,ij->j", sol.Z, Mz) - 1).max() <= 1e-10
    if prec == "fp64":
        # same vectors up to sign
        s = np.sign(np.einsum("ij,ij->j", sol.Z, pb.M @ ref.Z))
        assert np.abs(sol.Z * s - ref.Z).max() <= 1e-7 * np.abs(ref.Z).max()


def test_solve_hex8_10_matches_oracle(blas1):
    """The 8-node 10x10x10 workload (n = 3,630, 3 levels, nev 20)."""
    pb = cases.case_hex8_10()
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 500)
    ref = O.amls_solve(pb.K, pb.M, tree, nev=20)
    sol = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords),
                              AmlsOptions(nev=20, nLevels=3, verbosity=0))
    assert sol.p == ref.p
    assert np.abs(sol.theta - ref.theta).max() <= 1e-10 * ref.theta.max()


@pytest.mark.parametrize("prec,qmode", [("fp64", "ritz"), ("mixed", "ritz"),
                                        ("fp64", "basis"), ("mixed", "basis")])
def test_subspace_iteration_matches_oracle(blas1, prec, qmode):
    """AMLS-SIMb: the GPU M-orthonormalises between steps and does one FP64
    Rayleigh-Ritz at the end, the oracle does Rayleigh-Ritz after every step;
    both iterate the same subspace, so the final Ritz values agree, and they
    approach the FP64 ground truth from above (Courant-Fischer).  'ritz'
    starts from the q lowest AMLS Ritz vectors, 'basis' from the whole AMLS
    subspace (q = p)."""
    pb = small_hex(seed=31, dims=(5, 5, 6))
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    nev, steps = 12, 2
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev)
    q = 30 if qmode == "ritz" else ref.p
    _, Q = ref.cond.solve_projected(q)
    z = ref.cond.backtransform(Q)
    X = np.zeros_like(z)
    X[tree.perm] = z
    t_ref, _ = O.subspace_iteration(pb.K, pb.M, ref.ge, X, steps, nev, tree.perm)
    sol = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords),
                              AmlsOptions(nev=nev, nLevels=3, leafSize=60, precision=prec,
                                          siSteps=steps, siVectors=q if qmode == "ritz" else 10**6,
                                          verbosity=0))
    exact = O.reference_eigs(pb.K, pb.M, nev)
    assert sol.theta.shape == (nev,)
    assert np.all(sol.theta >= exact * (1 - 1e-12))
    rel = np.abs(sol.theta - t_ref) / t_ref
    assert rel.max() <= (1e-9 if prec == "fp64" else 1e-7)
    assert O.rel_err(sol.theta, exact).max() < O.rel_err(ref.theta, exact).max()
    assert np.abs(np.einsum("ij,ij->j", sol.Z, pb.M @ sol.Z) - 1).max() <= 1e-10

    def resid(th, Z):
        Kz = pb.K @ Z
        return np.linalg.norm(Kz - (pb.M @ Z) * th, axis=0) / np.linalg.norm(Kz, axis=0)
    t_or, Z_or = O.subspace_iteration(pb.K, pb.M, ref.ge, X, steps, nev, tree.perm)
    # FP64: the same iteration; mixed: the FP32-perturbed factors steer the
    # iteration slightly differently (Rayleigh-Ritz stays exact FP64)
    slack = 1e-3 if prec == "fp64" else 0.1
    assert np.all(resid(sol.theta, sol.Z) <= resid(t_or, Z_or) * (1 + slack) + 1e-12)


@pytest.mark.parametrize("si", [0, 2])
def test_shift_matches_oracle(blas1, si):
    """K <- K + shift M (src/amls/solver.cpp:114-122): the eigenvalues of the
    shifted pencil are returned, the vectors are unchanged."""
    pb = small_hex(seed=41, dims=(4, 4, 5))
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 40)
    nev, shift = 8, 0.37
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, shift=shift)
    sol = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords),
                              AmlsOptions(nev=nev, nLevels=3, leafSize=40, shift=shift,
                                          siSteps=si, verbosity=0))
    if si == 0:
        assert np.abs(sol.theta - ref.theta).max() <= 1e-10 * ref.theta.max()
    else:
        exact = O.reference_eigs(pb.K, pb.M, nev) + shift
        assert np.all(sol.theta >= exact * (1 - 1e-12))
        assert O.rel_err(sol.theta, exact).max() < O.rel_err(ref.theta, exact).max()
    if si > 0:
        Ks = pb.K + shift * pb.M
        r = Ks @ sol.Z - (pb.M @ sol.Z) * sol.theta
        rel = np.linalg.norm(r, axis=0) / np.linalg.norm(Ks @ sol.Z, axis=0)
        assert rel.max() <= 0.05


def test_ge_graph_replay_matches_eager():
    """GE runs eagerly the first time for a problem and precision, captures
    its launch sequence as a HIP graph the second time and replays it after
    that.  The captured run and every replay, with other device work and
    synchronisations between replays, give the node factors of the eager run
    (kernel statistics on) bit for bit; a new precision captures again, and
    so does a new problem."""
    pb = small_hex(seed=5, dims=(4, 5, 5))
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 60)
    tr = ctx.tree()
    nodes = sorted({0, len(tr["size"]) // 2, len(tr["size"]) - 1})
    rng = np.random.default_rng(3)
    A = rng.standard_normal((70, 50))
    B = rng.standard_normal((50, 40))

    def factors():
        return [ctx.node_factor(p, int(tr["size"][p])) for p in nodes]

    for prec in ("mixed", "fp64"):
        ctx.profile(True)
        ctx.ge(prec)
        eager = factors()
        ctx.profile(False)
        for rep in range(4):          # rep 0 captures (the eager run set the key), 1..3 replay
            ctx.ge(prec)
            for (ld, mh), (ld0, mh0) in zip(factors(), eager):
                assert np.array_equal(ld, ld0) and np.array_equal(mh, mh0), (prec, rep)
            C = ctx.gemm(A, B)        # eager work between replays
            assert np.abs(C - A @ B).max() <= 1e-12 * np.abs(A @ B).max()
    # a new problem (other values, same sizes) must not replay the old graph
    pb2 = small_hex(seed=6, dims=(4, 5, 5))
    ctx.set_problem(pb2.K, pb2.M, pb2.coords, 3, 60)
    ctx.profile(True)
    ctx.ge("fp64")
    eager2 = factors()
    ctx.profile(False)
    ctx.ge("fp64")
    for (ld, mh), (ld0, mh0) in zip(factors(), eager2):
        assert np.array_equal(ld, ld0) and np.array_equal(mh, mh0)
    assert not np.array_equal(eager2[0][0], eager[0][0])
    ctx.close()


@pytest.mark.parametrize("q", [0, 20])
def test_subspace_graph_replay_matches_eager(blas1, q):
    """The subspace iteration's steps and the set-up of its final
    Rayleigh-Ritz problem run eagerly the first time for a key, are captured
    as a HIP graph the second time and replayed after that.  Every captured
    and replayed solve, with other device work between them, gives the eager
    (profiled) solve's eigenvalues and vectors to rounding (the split-k Gram
    products add in any order); changing the width or the problem builds
    again.  q = 0 starts from the whole AMLS basis, q = 20 from recovered Ritz
    vectors."""
    pb = small_hex(seed=8, dims=(4, 5, 6))
    ctx = HipContext(0)
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    opt = AmlsOptions(nev=8, nLevels=3, leafSize=40, thresholdFactor=3.0, siSteps=3,
                      siVectors=q, precision="mixed", verbosity=0)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((60, 40))
    B = rng.standard_normal((40, 30))
    ctx.profile(True)
    th0, Z0, p0, _ = ctx.solve_numeric(opt, vectors=True)
    ctx.profile(False)
    for rep in range(4):              # rep 0 captures, 1..3 replay
        th, Z, p, _ = ctx.solve_numeric(opt, vectors=True)
        assert p == p0
        assert np.abs(th - th0).max() <= 1e-11 * th0.max(), rep
        s = np.sign(np.einsum("ij,ij->j", Z, pb.M @ Z0))
        assert np.abs(Z * s - Z0).max() <= 1e-8 * np.abs(Z0).max(), rep
        C = ctx.gemm(A, B)            # eager work between replays
        assert np.abs(C - A @ B).max() <= 1e-12 * np.abs(A @ B).max()
    # another width, then a new problem of the same sizes: built again
    th1 = ctx.solve_numeric(AmlsOptions(**{**opt.__dict__, "nev": 6}))[0]
    assert np.abs(th1 - th0[:6]).max() <= 1e-6 * th0.max()
    pb2 = small_hex(seed=9, dims=(4, 5, 6))
    ctx.set_problem(pb2.K, pb2.M, pb2.coords, 3, 40)
    ref = AmlsSolver(0).solve(Pencil(pb2.K, pb2.M, pb2.coords), opt)
    for rep in range(3):
        th2 = ctx.solve_numeric(opt)[0]
        assert np.abs(th2 - ref.theta).max() <= 1e-11 * ref.theta.max(), rep
    ctx.close()


def test_c2_subspace_iteration_matches_oracle():
    """The benchmark workload at full size (27-node hexahedra, 10^3 elements,
    n = 26,460, lowest 100, eta = 4): plain FP64 AMLS matches the oracle's
    eigenvalues within 1e-9 and its reduced size p exactly; with two
    subspace-iteration steps over the whole AMLS basis the FP64 and
    mixed-precision GPU eigenvalues agree with
    the oracle's FP64 AMLS + subspace iteration on the same tree within 1e-7
    relative (about 1e-10 measured), and both reach 1e-6 against the
    committed FP64 ground-truth fixture; three and four steps match the
    oracle's after as many steps."""
    import os
    pb = cases.case_hex27_10()
    tree, _ = tree_of(pb.K, pb.M, pb.coords, 6, 500)
    nev = 100
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=4.0, recoverVectors=False)
    _, Q = ref.cond.solve_projected(ref.p)
    z = ref.cond.backtransform(Q)
    X = np.zeros_like(z)
    X[tree.perm] = z
    _, _, hist = O.subspace_iteration(pb.K, pb.M, ref.ge, X, 4, nev, tree.perm, every_step=True)
    exact = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "hex27_10_lowest100.txt"))
    solver = AmlsSolver(0)
    # plain AMLS in FP64 (the drop-in form): the oracle's eigenvalues and p
    plain = solver.solve(Pencil(pb.K, pb.M, pb.coords),
                         AmlsOptions(nev=nev, nLevels=6, leafSize=500, thresholdFactor=4.0,
                                     precision="fp64", siSteps=0, verbosity=0))
    assert plain.p == ref.p
    assert (np.abs(plain.theta - ref.theta) / ref.theta).max() <= 1e-9
    mz = np.einsum("ij,ij->j", plain.Z, pb.M @ plain.Z)
    assert np.abs(mz - 1).max() <= 1e-10
    for prec in ("fp64", "mixed"):
        sol = solver.solve(Pencil(pb.K, pb.M, pb.coords),
                           AmlsOptions(nev=nev, nLevels=6, leafSize=500, thresholdFactor=4.0,
                                       precision=prec, siSteps=2, siVectors=10**6, verbosity=0))
        assert sol.p == ref.p
        assert (np.abs(sol.theta - hist[1]) / hist[1]).max() <= 1e-7, prec
        assert O.rel_err(sol.theta, exact).max() <= 1e-6, prec
    # three and four steps: the basis is M-orthonormalised after the second
    # step (scaling alone makes the projected stiffness indefinite here)
    for steps in (3, 4):
        sol = solver.solve(Pencil(pb.K, pb.M, pb.coords),
                           AmlsOptions(nev=nev, nLevels=6, leafSize=500, thresholdFactor=4.0,
                                       precision="mixed", siSteps=steps, siVectors=10**6,
                                       verbosity=0))
        t_or = hist[steps - 1]
        assert (np.abs(sol.theta - t_or) / t_or).max() <= 1e-7, steps
        assert O.rel_err(sol.theta, exact).max() <= O.rel_err(hist[1], exact).max(), steps


def test_context_reuse_smaller_then_larger_problem(blas1):
    """One AmlsSolver (one context) solves a small problem with subspace
    iteration and then a larger one with the same nev: the output block is
    sized by n x nev, so the second solve matches a fresh context's result."""
    nev = 6
    opt = AmlsOptions(nev=nev, nLevels=3, leafSize=40, siSteps=2, siVectors=16, verbosity=0)
    small = small_hex(seed=51, dims=(3, 3, 4))
    large = small_hex(seed=52, dims=(5, 5, 6))
    s = AmlsSolver(0)
    s.solve(Pencil(small.K, small.M, small.coords), opt)
    got = s.solve(Pencil(large.K, large.M, large.coords), opt)
    ref = AmlsSolver(0).solve(Pencil(large.K, large.M, large.coords), opt)
    assert got.Z.shape == (large.n, nev)
    assert np.abs(got.theta - ref.theta).max() <= 1e-12 * ref.theta.max()
    sg = np.sign(np.einsum("ij,ij->j", got.Z, large.M @ ref.Z))
    assert np.abs(got.Z * sg - ref.Z).max() <= 1e-9 * np.abs(ref.Z).max()


def test_duplicate_entries_summed_on_device_path(blas1):
    """Raw CSC arrays with every diagonal entry split in two and rows in
    descending order, straight through amls_hip_solve: the same eigenvalues
    as the canonical matrices (the library sums duplicates before building
    the front assembly maps and the sparse operators)."""
    import ctypes as C
    import scipy.sparse as sp
    from amls_amd import _csc, _lib
    from amls_amd._lib import dptr, iptr
    pb = small_hex(seed=53, dims=(4, 4, 4))
    nev = 6

    def split(A):
        A = sp.csc_matrix(A)
        r, c = A.nonzero()
        v = np.asarray(A[r, c]).ravel()
        d = r == c
        R = np.concatenate([r, r[d]])
        Cc = np.concatenate([c, c[d]])
        V = np.concatenate([np.where(d, 0.5 * v, v), 0.5 * v[d]])
        o = np.lexsort((-R, Cc))
        cp = np.zeros(A.shape[0] + 1, dtype=np.int64)
        np.add.at(cp, Cc[o] + 1, 1)
        return (np.cumsum(cp), np.ascontiguousarray(R[o], dtype=np.int64),
                np.ascontiguousarray(V[o], dtype=np.float64))

    opt = AmlsOptions(nev=nev, nLevels=3, leafSize=40, siSteps=1, siVectors=12, verbosity=0)
    ref = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords), opt)
    s = AmlsSolver(0)
    (kp, ki, kx), (mp, mi, mx) = split(pb.K), split(pb.M)
    th = np.zeros(nev)
    p = C.c_int64()
    oc = opt.to_c()
    cc = np.ascontiguousarray(pb.coords, dtype=np.float64)
    st = _lib.lib().amls_hip_solve(s.ctx._h, pb.n, iptr(kp), iptr(ki), dptr(kx), iptr(mp),
                                   iptr(mi), dptr(mx), dptr(cc), C.byref(oc), dptr(th), None,
                                   C.byref(p), None)
    s.ctx._chk(st)
    assert p.value == ref.p
    assert np.abs(th - ref.theta).max() <= 1e-12 * ref.theta.max()


@pytest.mark.parametrize("steps", [3, 4, 6])
def test_subspace_iteration_many_steps(blas1, steps):
    """More steps over the whole AMLS basis (q = p): the device
    M-orthonormalises after every second step, and the Ritz values follow the
    oracle's (Rayleigh-Ritz after every step) at 1e-9."""
    pb = small_hex(seed=31, dims=(5, 5, 6))
    tree, t = tree_of(pb.K, pb.M, pb.coords, 3, 60)
    nev = 12
    ref = O.amls_solve(pb.K, pb.M, tree, nev=nev, thresholdFactor=4.0, recoverVectors=False)
    _, Q = ref.cond.solve_projected(ref.p)
    z = ref.cond.backtransform(Q)
    X = np.zeros_like(z)
    X[tree.perm] = z
    t_or, _ = O.subspace_iteration(pb.K, pb.M, ref.ge, X, steps, nev, tree.perm)
    sol = AmlsSolver(0).solve(Pencil(pb.K, pb.M, pb.coords),
                              AmlsOptions(nev=nev, nLevels=3, leafSize=60, thresholdFactor=4.0,
                                          siSteps=steps, siVectors=10**6, verbosity=0))
    assert np.abs(sol.theta - t_or).max() <= 1e-9 * t_or.max()
    assert np.abs(np.einsum("ij,ij->j", sol.Z, pb.M @ sol.Z) - 1).max() <= 1e-10


def test_indefinite_pencil_is_not_spd_end_to_end():
    """A negated diagonal entry of K (tests/unit/test_linalg.cpp:69-78 in the
    reference) makes the node holding it indefinite: amls_hip_solve returns
    AMLS_E_NOT_SPD naming the node, in both precisions, and the context
    solves a valid pencil afterwards."""
    from amls_amd import AmlsError
    pb = small_hex(seed=61, dims=(4, 4, 5))
    K = pb.K.tolil()
    K[7, 7] = -K[7, 7]
    K = K.tocsc()
    s = AmlsSolver(0)
    for prec in ("fp64", "mixed"):
        with pytest.raises(AmlsError) as e:
            s.solve(Pencil(K, pb.M, pb.coords),
                    AmlsOptions(nev=6, nLevels=3, leafSize=40, precision=prec, verbosity=0))
        assert e.value.status == 2 and "node" in str(e.value)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords), AmlsOptions(nev=6, nLevels=3, leafSize=40,
                                                             verbosity=0))
    exact = O.reference_eigs(pb.K, pb.M, 6)
    assert np.all(sol.theta >= exact * (1 - 1e-12))


def test_empty_or_small_reduced_problem_is_noconv():
    """A cut below every local eigenvalue empties the reduced problem
    (src/amls/solver.cpp:148-149 throws there): AMLS_E_NOCONV.  A reduced
    problem smaller than nev returns its p eigenvalues for plain AMLS, as the
    reference's solveProjected does, and is AMLS_E_NOCONV when the subspace
    iteration needs nev vectors."""
    from amls_amd import AmlsError
    pb = small_hex(seed=62, dims=(4, 4, 5))
    s = AmlsSolver(0)
    with pytest.raises(AmlsError) as e:
        s.solve(Pencil(pb.K, pb.M, pb.coords),
                AmlsOptions(nev=6, nLevels=3, leafSize=40, thresholdFactor=1e-6, verbosity=0))
    assert e.value.status == 3 and "empty" in str(e.value)
    sol = s.solve(Pencil(pb.K, pb.M, pb.coords),
                  AmlsOptions(nev=200, nLevels=3, leafSize=40, thresholdFactor=0.7, verbosity=0))
    assert sol.p < 200 and sol.theta.size == sol.p
    with pytest.raises(AmlsError) as e:
        s.solve(Pencil(pb.K, pb.M, pb.coords),
                AmlsOptions(nev=200, nLevels=3, leafSize=40, thresholdFactor=0.7, siSteps=2,
                            verbosity=0))
    assert e.value.status == 3
