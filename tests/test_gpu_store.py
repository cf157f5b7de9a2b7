"""The out-of-core contribution store on device blocks: files written from
HBM equal the host store's files of the same blocks byte for byte, and the
GPU's sums of the records equal the host's bit for bit, through pieces of
whole columns and pieces of one column (src/amls/contribution_store.hpp
:23-124).  Device blocks are allocated with the HIP runtime the library
links."""
import os

import numpy as np
import pytest

from amls_amd.store import ContributionStore
from tests.helpers import DevBuf, hip_runtime

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols,ld,staging", [
    (300, 17, 300, 8 * 1000),      # pieces of three whole columns
    (2500, 3, 2600, 8 * 1000),     # pieces of 1,000 rows of one column
    (64, 64, 80, 0),               # one piece (default staging)
])
def test_device_store_matches_host_store(tmp_path, rows, cols, ld, staging):
    H = hip_runtime()
    rng = np.random.default_rng(rows + cols)
    dev = ContributionStore(str(tmp_path / "dev"), device=0, staging_bytes=staging)
    host = ContributionStore(str(tmp_path / "host"))
    for k in range(5):
        full = np.asfortranarray(rng.standard_normal((ld, cols)) * 10.0 ** rng.integers(-6, 6))
        d = DevBuf(H, full)                       # rows x cols block, leading dimension ld
        dev.append_device(11, d.ptr.value, rows, cols, ld)
        d.free()
        host.append(11, full[:rows, :])
    assert (open(tmp_path / "dev" / "t11.bin", "rb").read()
            == open(tmp_path / "host" / "t11.bin", "rb").read())
    found_h, want = host.read_sum(11)
    ldo = rows + 7
    out = DevBuf(H, np.full((ldo, cols), 3.0))
    assert dev.read_sum_device(11, out.ptr.value, rows, cols, ldo, consume=True)
    got = out.get()
    assert found_h
    assert np.ascontiguousarray(got[:rows]).tobytes() == np.ascontiguousarray(want).tobytes()
    assert np.all(got[rows:] == 3.0)                                     # padding untouched
    assert not os.path.exists(tmp_path / "dev" / "t11.bin")
    assert not dev.read_sum_device(11, out.ptr.value, rows, cols, ldo)
    assert dev.disk_bytes() == 5 * rows * cols * 8
    out.free()
    dev.close()
    host.close()


def test_device_store_reads_host_written_records(tmp_path):
    """Files are shared: records appended from the host sum on the GPU."""
    H = hip_runtime()
    rng = np.random.default_rng(5)
    d = str(tmp_path)
    host = ContributionStore(d)
    blocks = [rng.standard_normal((33, 9)) for _ in range(4)]
    for b in blocks:
        host.append(3, b)
    dev = ContributionStore(d, device=0, staging_bytes=8 * 100)
    out = DevBuf(H, np.zeros((33, 9)))
    assert dev.read_sum_device(3, out.ptr.value, 33, 9, 33)
    want = ((blocks[0] + blocks[1]) + blocks[2]) + blocks[3]
    assert out.get().tobytes(order="F") == np.asfortranarray(want).tobytes(order="F")   # bit for bit
    out.free()
    dev.close()
    host.close()
