"""ContributionStore (src/amls/contribution_store.hpp:23-124) over the
library's store (include/amls_hip.h, amls_hip_store_*).

    st = ContributionStore("/tmp/amls_ooc")            # host blocks
    st.append(7, block)                                 # append(target, m)
    found, total = st.read_sum(7, consume=True)         # readSum(target, out, consume)

With `device=d` the blocks are device memory of GPU d, passed as addresses
(`append_device`, `read_sum_device`); the records are staged through pinned
host buffers and summed on the GPU.  Files are the reference's: one
<dir>/t<target>.bin per target, records "int64 rows, int64 cols,
rows*cols doubles" (column-major).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import AmlsError

__all__ = ["ContributionStore"]


def _col_major(m: np.ndarray) -> Tuple[np.ndarray, int]:
    """(array, ld) of a column-major view, copying only when needed."""
    a = np.asarray(m)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2:
        raise ValueError("a block is 2-D")
    if a.dtype == np.float64 and a.strides[0] == 8 and (a.shape[1] <= 1 or a.strides[1] % 8 == 0) \
            and (a.shape[1] <= 1 or a.strides[1] >= 8 * a.shape[0]):
        return a, (a.strides[1] // 8 if a.shape[1] > 1 else max(a.shape[0], 1))
    a = np.asfortranarray(a, dtype=np.float64)
    return a, max(a.shape[0], 1)


class ContributionStore:
    """Append-only per-target store of signed dense contributions."""

    def __init__(self, directory: str, device: int = -1, staging_bytes: int = 0):
        self._L = _lib.lib()
        h = C.c_void_p()
        st = self._L.amls_hip_store_open(str(directory).encode(), int(device), int(staging_bytes),
                                         C.byref(h))
        if st != 0:
            raise AmlsError(st, f"amls_hip_store_open({directory!r}, device={device}) failed")
        self._h = h
        self.directory = str(directory)
        self.device = int(device)

    def _chk(self, st: int):
        if st != 0:
            msg = self._L.amls_hip_store_last_error(self._h)
            raise AmlsError(st, msg.decode() if msg else "")

    def close(self):
        """Deletes the store's files (as the reference's destructor) and frees it."""
        if getattr(self, "_h", None):
            self._L.amls_hip_store_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- host blocks (device < 0) -------------------------------------------
    def append(self, target: int, m) -> None:
        if self.device >= 0:
            raise AmlsError(1, "a device store takes device blocks (append_device)")
        a, ld = _col_major(m)
        self._chk(self._L.amls_hip_store_append(self._h, int(target), a.ctypes.data,
                                                a.shape[0], a.shape[1], ld))

    def read_sum(self, target: int, consume: bool = False) -> Tuple[bool, Optional[np.ndarray]]:
        """(found, sum of the target's records in append order)."""
        if self.device >= 0:
            raise AmlsError(1, "a device store takes device blocks (read_sum_device)")
        rows, cols, recs = self.shape(target)
        if recs < 0:
            return False, None
        out = np.zeros((rows, cols), order="F")
        found = C.c_int32()
        self._chk(self._L.amls_hip_store_read_sum(self._h, int(target), out.ctypes.data, rows,
                                                  cols, max(rows, 1), int(bool(consume)),
                                                  C.byref(found)))
        return bool(found.value), (out if recs > 0 else None)

    # ---- device blocks (device >= 0): addresses on that GPU ------------------
    def append_device(self, target: int, ptr: int, rows: int, cols: int, ld: int) -> None:
        self._chk(self._L.amls_hip_store_append(self._h, int(target), C.c_void_p(int(ptr)),
                                                int(rows), int(cols), int(ld)))

    def read_sum_device(self, target: int, ptr: int, rows: int, cols: int, ld: int,
                        consume: bool = False) -> bool:
        found = C.c_int32()
        self._chk(self._L.amls_hip_store_read_sum(self._h, int(target), C.c_void_p(int(ptr)),
                                                  int(rows), int(cols), int(ld),
                                                  int(bool(consume)), C.byref(found)))
        return bool(found.value)

    # ---- bookkeeping ---------------------------------------------------------
    def shape(self, target: int) -> Tuple[int, int, int]:
        """(rows, cols) of the first complete record and the record count
        (-1: the target has no file)."""
        r, c, n = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._L.amls_hip_store_shape(self._h, int(target), C.byref(r), C.byref(c),
                                               C.byref(n)))
        return r.value, c.value, n.value

    def remove(self, target: int) -> None:
        self._chk(self._L.amls_hip_store_remove(self._h, int(target)))

    def reset(self) -> None:
        self._chk(self._L.amls_hip_store_reset(self._h))

    def disk_bytes(self) -> float:
        b = C.c_double()
        self._chk(self._L.amls_hip_store_disk_bytes(self._h, C.byref(b)))
        return b.value
