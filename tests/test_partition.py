"""Nested dissection and symbolic analysis: the structural properties the
reference's tests pin (tests/unit/test_partition.cpp:16-100) plus the
multifrontal invariants the GPU path relies on."""
import numpy as np
import pytest
import scipy.sparse as sp

from amls_amd import HipContext, cases
from tests.helpers import grid_coords


def analyse(K, M, coords, nl, leaf):
    c = HipContext(-1)
    c.set_problem(K, M, coords, nl, leaf)
    t = c.tree()
    c.close()
    return t


CASES = [
    ("path", lambda: (*cases.path_graph(300), None), 4, 20),
    ("grid_bfs", lambda: (*cases.grid_graph_2d(30, 20), None), 4, 40),
    ("grid_plane", lambda: (*cases.grid_graph_2d(30, 20), grid_coords(30, 20)), 4, 40),
    ("hex8", lambda: (lambda p: (p.K, p.M, p.coords))(cases.hex_elasticity(1, 6, 5, 4)), 3, 60),
    ("hex27", lambda: (lambda p: (p.K, p.M, p.coords))(cases.hex_elasticity(2, 3, 3, 3)), 3, 60),
]


@pytest.mark.parametrize("name,make,nl,leaf", CASES)
def test_tree_structure(name, make, nl, leaf):
    K, M, coords = make()
    n = K.shape[0]
    t = analyse(K, M, coords, nl, leaf)
    N = len(t["size"])
    perm = t["perm"]
    # perm[new] = old is a permutation; nodes are contiguous in post-order
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert t["offset"][0] == 0
    assert np.array_equal(t["offset"][1:], np.cumsum(t["size"])[:-1])
    # <= 2^nLevels leaves, binary tree, parents after children
    par = t["parent"]
    assert (par == -1).sum() == 1 and par[-1] == -1
    kids = np.bincount(par[par >= 0], minlength=N)
    assert set(kids.tolist()) <= {0, 2}
    assert (kids == 0).sum() <= 2 ** nl
    assert np.all(par[:-1] > np.arange(N - 1))
    # levels: leaves 0, parent = max(child) + 1
    for p in range(N):
        ch = np.flatnonzero(par == p)
        assert t["level"][p] == (0 if ch.size == 0 else t["level"][ch].max() + 1)
    # vertex-separator property on the permuted graph
    A = (abs(K) + abs(M)).tocoo()
    owner = np.repeat(np.arange(N), t["size"])
    iperm = np.empty(n, dtype=np.int64)
    iperm[perm] = np.arange(n)
    ro, co = owner[iperm[A.row]], owner[iperm[A.col]]
    anc = [set() for _ in range(N)]
    for p in range(N):
        q = par[p]
        while q >= 0:
            anc[p].add(q)
            q = par[q]
    for a, b in set(zip(ro.tolist(), co.tolist())):
        assert a == b or a in anc[b] or b in anc[a]


@pytest.mark.parametrize("name,make,nl,leaf", CASES)
def test_boundary_sets_brute_force(name, make, nl, leaf):
    """b_p = |B_p|, B_p = indices after node p that couple to p after the
    elimination of its subtree: checked by symbolic elimination."""
    K, M, coords = make()
    n = K.shape[0]
    t = analyse(K, M, coords, nl, leaf)
    perm = t["perm"]
    A = (abs(K) + abs(M)).tocsr()[perm][:, perm]
    adj = [set(A.indices[A.indptr[i]:A.indptr[i + 1]].tolist()) - {i} for i in range(n)]
    N = len(t["size"])
    B = [None] * N
    for p in range(N):
        lo, hi = t["offset"][p], t["offset"][p] + t["size"][p]
        s = set()
        for j in range(lo, hi):
            s |= {x for x in adj[j] if x >= hi}
        for c in np.flatnonzero(t["parent"] == p):
            s |= {x for x in B[c] if x >= hi}
        B[p] = s
        assert len(s) == t["bsize"][p]


def test_plane_separators_on_hex27():
    pb = cases.case_hex27_10()
    t = analyse(pb.K, pb.M, pb.coords, 6, 500)
    # root separator = one 21 x 20 node plane, 3 dofs per node
    assert t["size"][-1] == 1260
    assert t["maxLevel"] == 6


def test_batch_forest_equals_single_trees():
    """A batch of independent pencils is partitioned problem by problem: the
    forest's nodes, sizes, boundaries and permutation are the single trees',
    shifted by each problem's row and node offsets."""
    from amls_amd import Pencil
    rng = np.random.default_rng(2)
    pbs = [cases.hex_elasticity(1, *(int(v) for v in rng.integers(3, 7, 3))) for _ in range(4)]
    c = HipContext(-1)
    c.set_problem_batch([Pencil(p.K, p.M, p.coords) for p in pbs], 3, 40)
    t = c.tree()
    rows, nodes = c.batch_info()
    c.close()
    assert rows[-1] == sum(p.n for p in pbs) and len(t["size"]) == nodes[-1]
    for i, pb in enumerate(pbs):
        s = analyse(pb.K, pb.M, pb.coords, 3, 40)
        a, b = nodes[i], nodes[i + 1]
        assert np.array_equal(t["size"][a:b], s["size"])
        assert np.array_equal(t["offset"][a:b] - rows[i], s["offset"])
        assert np.array_equal(t["bsize"][a:b], s["bsize"])
        assert np.array_equal(t["level"][a:b], s["level"])
        par = t["parent"][a:b]
        assert np.array_equal(np.where(par >= 0, par - a, -1), s["parent"])
        assert np.array_equal(t["perm"][rows[i]:rows[i + 1]] - rows[i], s["perm"])


@pytest.mark.parametrize("coords", [False, True])
def test_one_sided_pattern_is_symmetrised(coords):
    """The graph of |K| + |M| joins i and j when either (i, j) or (j, i) is
    stored: a pencil given by its lower triangles only (and one with a single
    extra one-sided entry) is partitioned like the full symmetric one.  The
    first takes the library's general symmetrising path, the full pencil its
    merge path."""
    pb = cases.hex_elasticity(1, 5, 4, 3)
    xyz = pb.coords if coords else None
    full = analyse(pb.K, pb.M, xyz, 3, 40)
    lower = analyse(sp.tril(pb.K).tocsc(), sp.tril(pb.M).tocsc(), xyz, 3, 40)
    for key in ("treeId", "offset", "size", "parent", "bsize", "perm"):
        assert np.array_equal(full[key], lower[key]), key
    # one entry (i, j) whose mirror (j, i) is absent from both matrices
    K = sp.lil_matrix(pb.K)
    i, j = 0, pb.n - 1
    assert pb.K[i, j] == 0 and pb.M[i, j] == 0
    K[i, j] = 1e-3
    one = analyse(K.tocsc(), pb.M, xyz, 3, 40)
    Ks = sp.lil_matrix(K)
    Ks[j, i] = 1e-3
    both = analyse(Ks.tocsc(), pb.M, xyz, 3, 40)
    for key in ("treeId", "offset", "size", "parent", "bsize", "perm"):
        assert np.array_equal(one[key], both[key]), key
