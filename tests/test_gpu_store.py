"""The out-of-core contribution store on device blocks: files written from
HBM equal the host store's files of the same blocks byte for byte, and the
GPU's sums of the records equal the host's bit for bit, through pieces of
whole columns and pieces of one column (src/amls/contribution_store.hpp
:23-124)."""
import os

import numpy as np
import pytest
import torch

from amls_amd.store import ContributionStore

pytestmark = pytest.mark.gpu


def device_block(rng, rows, cols, ld):
    """A column-major rows x cols block with leading dimension ld in HBM:
    (tensor holding it, the same block on the host)."""
    host = np.asfortranarray(rng.standard_normal((ld, cols)) * 10.0 ** rng.integers(-6, 6))
    t = torch.from_numpy(host.T.copy()).to("cuda")        # row c of t = column c
    return t, host[:rows, :]


@pytest.mark.parametrize("rows,cols,ld,staging", [
    (300, 17, 300, 8 * 1000),      # pieces of three whole columns
    (2500, 3, 2600, 8 * 1000),     # pieces of 1,000 rows of one column
    (64, 64, 80, 0),               # one piece (default staging)
])
def test_device_store_matches_host_store(tmp_path, rows, cols, ld, staging):
    rng = np.random.default_rng(rows + cols)
    dev = ContributionStore(str(tmp_path / "dev"), device=0, staging_bytes=staging)
    host = ContributionStore(str(tmp_path / "host"))
    keep = []
    for k in range(5):
        t, h = device_block(rng, rows, cols, ld)
        keep.append(t)
        torch.cuda.synchronize()
        dev.append_device(11, t.data_ptr(), rows, cols, ld)
        host.append(11, h)
    assert (open(tmp_path / "dev" / "t11.bin", "rb").read()
            == open(tmp_path / "host" / "t11.bin", "rb").read())
    found_h, want = host.read_sum(11)
    ldo = rows + 7
    out = torch.full((cols, ldo), 3.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert dev.read_sum_device(11, out.data_ptr(), rows, cols, ldo, consume=True)
    got = out.cpu().numpy().T
    assert found_h
    assert got[:rows].tobytes(order="F") == want.tobytes(order="F")     # bit for bit
    assert np.all(got[rows:] == 3.0)                                     # padding untouched
    assert not os.path.exists(tmp_path / "dev" / "t11.bin")
    assert not dev.read_sum_device(11, out.data_ptr(), rows, cols, ldo)
    assert dev.disk_bytes() == 5 * rows * cols * 8
    dev.close()
    host.close()


def test_device_store_reads_host_written_records(tmp_path):
    """Files are shared: records appended from the host sum on the GPU."""
    rng = np.random.default_rng(5)
    d = str(tmp_path)
    host = ContributionStore(d)
    blocks = [rng.standard_normal((33, 9)) for _ in range(4)]
    for b in blocks:
        host.append(3, b)
    dev = ContributionStore(d, device=0, staging_bytes=8 * 100)
    out = torch.zeros((9, 33), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert dev.read_sum_device(3, out.data_ptr(), 33, 9, 33)
    want = ((blocks[0] + blocks[1]) + blocks[2]) + blocks[3]
    assert out.cpu().numpy().T.tobytes(order="F") == np.asfortranarray(want).tobytes(order="F")
    dev.close()
    host.close()
