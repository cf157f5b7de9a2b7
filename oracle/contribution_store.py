"""Oracle of the out-of-core contribution store: a numpy restatement of the
reference's ContributionStore (src/amls/contribution_store.hpp:23-124).

TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing in amls_amd/.

The reference keeps, per target block, one append-only file
<dir>/t<target>.bin (path_, :117-119) of records written by append
(:34-49):

    int64 rows, int64 cols, then rows*cols doubles in Eigen's default
    column-major order, native byte order (little-endian on x86-64 and the
    MI355X hosts);

readSum (:53-92) reads the records in file order, the first one becomes the
output and every later one is added to it (out.noalias() += m), stopping at
the first record it cannot read whole (a truncated tail is ignored); it
returns false when the file cannot be opened, and deletes the file when
consume is set.  diskBytes (:106) counts rows*cols*8 per append since the
last reset.

Pinned by tests/test_store.py: the byte layout of a record against a
known-answer byte string packed field by field, the truncated-tail rule on a
hand-cut file, and the sums against numpy's own element-wise sums.
"""
from __future__ import annotations

import os
from typing import List, Optional, Tuple

import numpy as np


def record_path(directory: str, target: int) -> str:
    """path_ (:117-119)."""
    return os.path.join(directory, f"t{int(target)}.bin")


def record_bytes(m: np.ndarray) -> bytes:
    """One record as append writes it (:39-45)."""
    a = np.asarray(m, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    hdr = np.array([a.shape[0], a.shape[1]], dtype="<i8").tobytes()
    return hdr + np.asfortranarray(a).astype("<f8").tobytes(order="F")


def append(directory: str, target: int, m: np.ndarray) -> int:
    """append (:34-49): the record added to the target's file; returns the
    payload bytes it adds to diskBytes."""
    os.makedirs(directory, exist_ok=True)
    with open(record_path(directory, target), "ab") as f:
        f.write(record_bytes(m))
    a = np.asarray(m)
    return int(a.size) * 8


def records(directory: str, target: int) -> Optional[List[np.ndarray]]:
    """The complete records of a target in file order (None: no file); the
    reading loop of readSum (:62-78)."""
    p = record_path(directory, target)
    if not os.path.exists(p):
        return None
    data = open(p, "rb").read()
    out, off = [], 0
    while off + 16 <= len(data):
        r, c = np.frombuffer(data, dtype="<i8", count=2, offset=off)
        need = int(r) * int(c) * 8
        if off + 16 + need > len(data):
            break
        m = np.frombuffer(data, dtype="<f8", count=int(r) * int(c), offset=off + 16)
        out.append(m.reshape(int(r), int(c), order="F").copy())
        off += 16 + need
    return out


def read_sum(directory: str, target: int, consume: bool = False
             ) -> Tuple[bool, Optional[np.ndarray]]:
    """readSum (:53-92): (found, first record + every later one, in order)."""
    recs = records(directory, target)
    if recs is None:
        return False, None
    out = None
    for m in recs:
        out = m.copy() if out is None else out + m
    if consume:
        os.remove(record_path(directory, target))
    return True, out
