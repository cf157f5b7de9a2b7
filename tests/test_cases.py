"""Known answers for the synthetic pencils (generator = input side)."""
import numpy as np
import pytest
import scipy.sparse as sp

from amls_amd import cases


@pytest.mark.parametrize("order", [1, 2])
def test_element_rigid_modes_and_mass(order):
    Ke, Me = cases.hex_element_matrices(order)
    nd = Ke.shape[0]
    w = np.linalg.eigvalsh(Ke)
    # six rigid-body modes of an unconstrained element, no more
    assert np.sum(np.abs(w) < 1e-10 * w.max()) == 6
    assert w[6] > 1e-6 * w.max()
    for c in range(3):
        t = np.zeros(nd)
        t[c::3] = 1.0
        assert np.abs(Ke @ t).max() < 1e-12          # translation
        assert abs(t @ Me @ t - 1.0) < 1e-13          # rho * V
    # linear rotation field (x, y) -> (-y, x) is stress free
    p1 = order + 1
    g = np.linspace(0, 1, p1)
    xx, yy = np.meshgrid(g, g, indexing="xy")
    X = np.tile(xx.ravel(), p1)
    Y = np.tile(yy.ravel(), p1)
    r = np.zeros(nd)
    r[0::3] = -Y
    r[1::3] = X
    assert np.abs(Ke @ r).max() < 1e-12


@pytest.mark.parametrize("fn,n,nnz", [(cases.case_hex8_10, 3630, 242172)])
def test_workload_sizes(fn, n, nnz):
    pb = fn()
    assert pb.n == n
    assert pb.K.nnz == nnz and pb.M.nnz == nnz
    assert np.array_equal(pb.K.indices, pb.M.indices)
    assert abs(pb.K - pb.K.T).max() == 0.0
    assert abs(pb.M - pb.M.T).max() == 0.0


def test_hex27_sizes_small():
    # 27-node hex on 2x2x2 elements: nodes 5^3, clamped face removes 25
    pb = cases.hex_elasticity(2, 2, 2, 2)
    assert pb.n == 3 * (125 - 25)
    # K SPD after clamping, M SPD
    w = np.linalg.eigvalsh(pb.K.toarray())
    assert w.min() > 0
    assert np.linalg.eigvalsh(pb.M.toarray()).min() > 0
    # total mass of the free dofs' rows equals volume minus clamped share
    one = np.zeros(pb.n)
    one[0::3] = 1
    assert 0 < one @ pb.M @ one < 8.0


def test_scales_enter_linearly():
    ks = np.full(8, 2.0)
    a = cases.hex_elasticity(1, 2, 2, 2)
    b = cases.hex_elasticity(1, 2, 2, 2, kscale=ks, rscale=ks * 1.5)
    assert np.allclose(b.K.toarray(), 2.0 * a.K.toarray(), rtol=0, atol=1e-14)
    assert np.allclose(b.M.toarray(), 3.0 * a.M.toarray(), rtol=0, atol=1e-14)


def test_closed_forms_match_dense():
    for (K, M), spec in [(cases.path_graph(50), cases.path_graph_spectrum(50)),
                         (cases.path_graph(40, True), cases.path_graph_spectrum(40, True)),
                         (cases.grid_graph_2d(7, 5), cases.grid_graph_2d_spectrum(7, 5)),
                         (cases.grid_graph_3d(4, 3, 5), cases.grid_graph_3d_spectrum(4, 3, 5))]:
        import scipy.linalg as sla
        w = sla.eigh(K.toarray(), M.toarray(), eigvals_only=True)
        assert np.abs(w - spec).max() <= 1e-12 * max(1.0, spec.max())
