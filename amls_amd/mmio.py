"""MatrixMarket pencil files, the exchange format of the reference's
`amls_solve --K K.mtx --M M.mtx` (src/io/matrix_market.cpp:48-115), so a
pencil generated here can be solved by a build of the reference and the
results compared.

Reading accepts `coordinate` `real` / `double` / `integer` / `pattern` with
`symmetric` or `general` symmetry (1-based indices, duplicates summed,
explicit zeros kept), as the reference's reader does.  Writing emits the
lower triangle of a symmetric matrix with 17 significant digits, so values
round-trip exactly.  The reference's writer puts the full nonzero count in
the size line; this one writes the number of entry lines that follow, and
both readers take the entries to the end of the file either way.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = ["read_mtx", "write_mtx"]


def read_mtx(path) -> sp.csc_matrix:
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0] != "%%MatrixMarket":
            raise ValueError(f"{path}: missing %%MatrixMarket banner")
        obj, fmt, field, sym = (b.lower() for b in banner[1:5])
        if obj != "matrix" or fmt != "coordinate":
            raise ValueError(f"{path}: only coordinate matrices are supported")
        if sym not in ("symmetric", "general"):
            raise ValueError(f"{path}: only symmetric / general matrices are supported")
        if field not in ("real", "double", "integer", "pattern"):
            raise ValueError(f"{path}: only real / integer / pattern fields are supported")
        line = f.readline()
        while line and (not line.strip() or line.startswith("%")):
            line = f.readline()
        if not line:
            raise ValueError(f"{path}: missing size line")
        nrows, ncols, _ = (int(v) for v in line.split()[:3])
        rows, cols, vals = [], [], []
        for line in f:
            if not line.strip() or line.startswith("%"):
                continue
            parts = line.split()
            rows.append(int(parts[0]) - 1)
            cols.append(int(parts[1]) - 1)
            vals.append(1.0 if field == "pattern" else float(parts[2]))
    r = np.asarray(rows, dtype=np.int64)
    c = np.asarray(cols, dtype=np.int64)
    v = np.asarray(vals, dtype=np.float64)
    if sym == "symmetric":
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    A = sp.coo_matrix((v, (r, c)), shape=(nrows, ncols)).tocsc()
    A.sum_duplicates()
    return A


def write_mtx(path, A) -> None:
    """Lower triangle of the symmetric matrix A (entries as stored, explicit
    zeros included)."""
    A = sp.coo_matrix(A)
    keep = A.row >= A.col
    r, c, v = A.row[keep], A.col[keep], A.data[keep]
    order = np.lexsort((r, c))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.shape[0]} {A.shape[1]} {len(v)}\n")
        f.writelines(f"{int(r[k]) + 1} {int(c[k]) + 1} {float(v[k]):.17g}\n" for k in order)
