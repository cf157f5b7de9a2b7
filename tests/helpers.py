"""Shared helpers for the test modules."""
import numpy as np


def tree_of(K, M, coords=None, nLevels=6, leafSize=500):
    """Separator tree from the library's host-side analysis (no GPU)."""
    from amls_amd import HipContext
    from oracle.amls_oracle import Tree
    c = HipContext(-1)
    c.set_problem(K, M, coords, nLevels, leafSize)
    t = c.tree()
    c.close()
    return Tree.from_dict(t), t


def grid_coords(nx, ny):
    j, i = np.divmod(np.arange(nx * ny), nx)
    return np.stack([i, j, np.zeros_like(i)], 1).astype(float)
