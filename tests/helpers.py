"""Shared helpers for the test modules."""
import ctypes as C

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


H2D, D2H = 1, 2


def hip_runtime():
    """The libamdhip64 the library was loaded with (not another copy that a
    framework may bundle)."""
    from amls_amd import _lib
    _lib.lib()
    paths = []
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64.so" in p and p not in paths:
                paths.append(p)
    rocm = [p for p in paths if "site-packages" not in p and "dist-packages" not in p]
    assert rocm, paths
    H = C.CDLL(rocm[0])
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipDeviceSynchronize.argtypes = []
    return H


class DevBuf:
    def __init__(self, H, arr):
        self.H = H
        self.shape = arr.shape
        self.bytes = arr.nbytes
        self.ptr = C.c_void_p()
        assert H.hipMalloc(C.byref(self.ptr), self.bytes) == 0
        host = np.asfortranarray(arr, dtype=np.float64)
        assert H.hipMemcpy(self.ptr, host.ctypes.data, self.bytes, H2D) == 0

    def get(self):
        out = np.zeros(self.shape, order="F")
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, self.bytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.ptr)
