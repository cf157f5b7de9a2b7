"""The C-ABI library loads without a GPU and exports every declared symbol;
host-only contexts run the analysis and report errors as status codes."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from amls_amd import AmlsError, HipContext, _lib, cases

HEADER = Path(__file__).resolve().parents[1] / "include" / "amls_hip.h"


def declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"\b(amls_(?:hip_)?[a-z_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == names


def test_options_default_matches_reference():
    o = _lib.AmlsOptionsC()
    assert _lib.lib().amls_options_default(ctypes.byref(o)) == 0
    # src/amls/solver.hpp:18-30
    assert (o.nLevels, o.leafSize, o.thresholdFactor, o.fixedKeep, o.nev,
            o.denseProjectedMax, o.shift, o.twoPass, o.numThreads,
            o.recoverVectors, o.verbosity) == (6, 500, 2.0, 0, 100, 5000, 0.0, 1, 0, 1, 1)


def test_host_context_rejects_gpu_phase():
    K, M = cases.path_graph(30)
    ctx = HipContext(-1)
    ctx.set_problem(K, M, None, 2, 5)
    with pytest.raises(AmlsError) as e:
        ctx.ge()
    assert e.value.status == 6


def test_bad_arguments_are_status_codes():
    ctx = HipContext(-1)
    K, M = cases.path_graph(10)
    bad = K.tocsc().copy()
    bad.indices = bad.indices.copy()
    bad.indices[3] = 99
    with pytest.raises(AmlsError) as e:
        ctx.set_problem(bad, M, None, 2, 5)
    assert e.value.status == 1
    assert "out of range" in str(e.value)
