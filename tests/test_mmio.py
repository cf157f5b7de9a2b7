"""MatrixMarket pencil I/O (src/io/matrix_market.cpp:48-115): exact round
trip of a generated pencil, the symmetric / general / pattern readers, and
the reference writer's size line (full nonzero count) accepted."""
import numpy as np
import scipy.sparse as sp

from amls_amd import cases
from amls_amd.mmio import read_mtx, write_mtx


def test_round_trip_exact(tmp_path):
    pb = cases.hex_elasticity(2, 2, 2, 2, kscale=np.random.default_rng(0).lognormal(0, 0.3, 8))
    for A, name in ((pb.K, "K.mtx"), (pb.M, "M.mtx")):
        p = tmp_path / name
        write_mtx(p, A)
        B = read_mtx(p)
        assert B.shape == A.shape and B.nnz == A.nnz          # explicit zeros kept
        assert (abs(B - A)).max() == 0.0


def test_readers(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n2 2 3\n1 1 2.5\n2 1 -1\n1 2 -1\n")
    assert np.array_equal(read_mtx(p).toarray(), [[2.5, -1], [-1, 0]])
    p = tmp_path / "s.mtx"
    # size line with the full count, as the reference's writer emits it
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 2 3\n1 1 4\n2 1 1\n2 2 3\n")
    assert np.array_equal(read_mtx(p).toarray(), [[4, 1], [1, 3]])
    p = tmp_path / "p.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n2 2 2\n1 1\n2 1\n")
    assert np.array_equal(read_mtx(p).toarray(), [[1, 1], [1, 0]])
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n1 1 2\n")
    assert read_mtx(p)[0, 0] == 3.0                             # duplicates summed


def test_closed_form_through_files(tmp_path):
    K, M = cases.path_graph(50, True)
    write_mtx(tmp_path / "K.mtx", K)
    write_mtx(tmp_path / "M.mtx", M)
    import scipy.linalg as sla
    w = sla.eigh(read_mtx(tmp_path / "K.mtx").toarray(), read_mtx(tmp_path / "M.mtx").toarray(),
                 eigvals_only=True)
    ref = cases.path_graph_spectrum(50, True)
    assert np.abs(w - ref).max() <= 1e-12 * ref.max()
