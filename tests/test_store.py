"""The out-of-core contribution store on host blocks (no GPU): the oracle
pinned by known-answer bytes and the truncated-tail rule, then the library
against the oracle byte for byte and bit for bit
(src/amls/contribution_store.hpp:23-124)."""
import ctypes
import os
import struct
import threading

import numpy as np
import pytest

from amls_amd import AmlsError, _lib
from amls_amd.store import ContributionStore
from oracle import contribution_store as O


def test_oracle_record_layout_known_answer(tmp_path):
    # a 2 x 3 block: int64 rows, int64 cols, then the columns one after another
    m = np.array([[1.0, 3.0, 5.0], [2.0, -4.0, 0.5]])
    want = struct.pack("<qq", 2, 3) + struct.pack("<6d", 1.0, 2.0, 3.0, -4.0, 5.0, 0.5)
    assert O.record_bytes(m) == want
    d = str(tmp_path)
    assert O.append(d, 7, m) == 48
    O.append(d, 7, 2 * m)
    assert open(os.path.join(d, "t7.bin"), "rb").read() == want + O.record_bytes(2 * m)
    found, s = O.read_sum(d, 7)
    assert found and np.array_equal(s, 3 * m)
    assert O.read_sum(d, 8) == (False, None)


def test_oracle_ignores_truncated_tail(tmp_path):
    d = str(tmp_path)
    a = np.arange(6.0).reshape(2, 3)
    O.append(d, 1, a)
    O.append(d, 1, a)
    p = os.path.join(d, "t1.bin")
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-8])            # the second record loses its last value
    found, s = O.read_sum(d, 1)
    assert found and np.array_equal(s, a)
    open(p, "wb").write(raw[:60])            # only a header and part of the first record
    found, s = O.read_sum(d, 1)
    assert found and s is None


def _blocks(rng, shape, count):
    return [rng.standard_normal(shape) * 10.0 ** rng.integers(-8, 8) for _ in range(count)]


def test_files_and_sums_match_oracle(tmp_path):
    rng = np.random.default_rng(3)
    lib_dir, ora_dir = str(tmp_path / "lib"), str(tmp_path / "ora")
    st = ContributionStore(lib_dir, staging_bytes=8 * 40)   # 40-element pieces
    want_bytes = 0
    for target, shape in [(0, (13, 7)), (5, (1, 1)), (-3, (90, 2)), (12, (3, 100))]:
        for b in _blocks(rng, shape, 4):
            st.append(target, b)
            want_bytes += O.append(ora_dir, target, b)
        assert (open(os.path.join(lib_dir, f"t{target}.bin"), "rb").read()
                == open(os.path.join(ora_dir, f"t{target}.bin"), "rb").read())
        found, s = st.read_sum(target)
        f2, s2 = O.read_sum(ora_dir, target)
        assert found and f2
        assert s.tobytes(order="F") == s2.tobytes(order="F")       # the same additions, bit for bit
        assert st.shape(target) == (shape[0], shape[1], 4)
    assert st.disk_bytes() == want_bytes
    st.close()
    assert not any(n.endswith(".bin") for n in os.listdir(lib_dir))   # close resets


def test_padded_and_c_ordered_blocks(tmp_path):
    rng = np.random.default_rng(4)
    big = np.asfortranarray(rng.standard_normal((20, 9)))
    view = big[3:14, 1:8]                       # column-major with leading dimension 20
    c_ordered = np.ascontiguousarray(rng.standard_normal((11, 7)))
    with ContributionStore(str(tmp_path / "s"), staging_bytes=64) as st:
        st.append(1, view)
        st.append(1, c_ordered)
        O.append(str(tmp_path / "o"), 1, view)
        O.append(str(tmp_path / "o"), 1, c_ordered)
        assert (open(tmp_path / "s" / "t1.bin", "rb").read()
                == open(tmp_path / "o" / "t1.bin", "rb").read())
        _, s = st.read_sum(1)
        assert np.array_equal(s, view + c_ordered)


def test_consume_remove_reset_and_missing(tmp_path):
    d = str(tmp_path)
    st = ContributionStore(d)
    a = np.ones((4, 4))
    assert st.read_sum(9) == (False, None)
    assert st.shape(9) == (0, 0, -1)
    st.append(9, a)
    st.append(10, a)
    st.append(11, a)
    found, s = st.read_sum(9, consume=True)
    assert found and np.array_equal(s, a)
    assert not os.path.exists(os.path.join(d, "t9.bin"))
    assert st.read_sum(9) == (False, None)
    st.remove(10)
    assert not os.path.exists(os.path.join(d, "t10.bin"))
    assert st.disk_bytes() == 3 * 128            # consumed / removed bytes still count
    st.reset()
    assert st.disk_bytes() == 0.0
    assert not os.path.exists(os.path.join(d, "t11.bin"))
    st.close()


def test_shape_mismatch_and_truncated_tail(tmp_path):
    d = str(tmp_path)
    st = ContributionStore(d)
    st.append(2, np.ones((3, 3)))
    st.append(2, np.ones((2, 3)))
    with pytest.raises(AmlsError, match="2 x 3"):
        st.read_sum(2)
    st.remove(2)
    a = np.arange(12.0).reshape(3, 4)
    st.append(4, a)
    st.append(4, a)
    p = os.path.join(d, "t4.bin")
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-1])
    assert st.shape(4) == (3, 4, 1)
    found, s = st.read_sum(4)
    _, s2 = O.read_sum(d, 4)
    assert found and np.array_equal(s, a) and np.array_equal(s2, a)
    st.close()


def test_concurrent_appends_stay_whole(tmp_path):
    d = str(tmp_path)
    st = ContributionStore(d, staging_bytes=8 * 16)
    blocks = {t: [np.full((17, 5), float(t * 100 + k)) for k in range(25)] for t in range(4)}

    def work(t, ks):
        for k in ks:
            st.append(t, blocks[t][k])

    # two threads per target, interleaving their appends
    th = [threading.Thread(target=work, args=(t, range(h, 25, 2))) for t in range(4) for h in (0, 1)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for t in range(4):
        recs = O.records(d, t)
        assert len(recs) == 25
        assert sorted(float(r[0, 0]) for r in recs) == [float(t * 100 + k) for k in range(25)]
        assert all(np.all(r == r[0, 0]) for r in recs)        # no record torn by another
        found, s = st.read_sum(t)
        assert found and np.array_equal(s, sum(blocks[t]))   # integers: any order is exact
    st.close()


def test_bad_arguments(tmp_path):
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.amls_hip_store_open(None, -1, 0, ctypes.byref(h)) == 1
    assert L.amls_hip_store_open(str(tmp_path).encode(), -1, -5, ctypes.byref(h)) == 1
    assert L.amls_hip_store_append(None, 0, None, 1, 1, 1) == 1
    assert L.amls_hip_store_close(None) == 1
    st = ContributionStore(str(tmp_path))
    with pytest.raises(AmlsError, match="ld < rows"):
        st._chk(L.amls_hip_store_append(st._h, 0, ctypes.c_void_p(8), 5, 3, 4))
    with pytest.raises(AmlsError, match="null block"):
        st._chk(L.amls_hip_store_append(st._h, 0, None, 5, 3, 5))
    with pytest.raises(AmlsError, match="ld < rows"):
        st.append_device(0, 8, 2, 2, 1)       # checked before any access
    st.close()
