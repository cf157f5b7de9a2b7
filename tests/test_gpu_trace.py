"""The unified trace on the GPU: host phase spans and every GPU launch of two
numeric solves on one timeline (src/core/trace.hpp's Chrome-trace format),
each launch inside its solve's host span and tagged with its phase."""
import json

import numpy as np
import pytest

from amls_amd import AmlsOptions, HipContext, cases

pytestmark = pytest.mark.gpu


def test_trace_host_and_gpu_spans(tmp_path):
    rng = np.random.default_rng(3)
    dims = (5, 5, 6)
    ne = dims[0] * dims[1] * dims[2]
    pb = cases.hex_elasticity(1, *dims, kscale=rng.lognormal(0, 0.3, ne))
    path = tmp_path / "trace.json"
    ctx = HipContext(0)
    ctx.trace_begin(str(path))
    ctx.set_problem(pb.K, pb.M, pb.coords, 3, 60)
    o = AmlsOptions(nev=10, nLevels=3, leafSize=60, thresholdFactor=3.0, precision="mixed",
                    siSteps=2, siVectors=40, verbosity=0)
    th1 = ctx.solve_numeric(o)[0]
    th2 = ctx.solve_numeric(o)[0]
    n = ctx.trace_end()
    ctx.close()
    assert np.array_equal(th1, th2)
    ev = [e for e in json.loads(path.read_text())["traceEvents"] if e["ph"] == "X"]
    assert len(ev) == n
    host = [e for e in ev if e["tid"] == 0]
    gpu = [e for e in ev if e["tid"] in (1, 2)]
    names = [e["name"] for e in host]
    for nm in ("ANALYSIS", "TABLES", "UPLOAD"):
        assert names.count(nm) == 1, nm
    for nm in ("GE", "CONDENSE", "SUBSPACE", "SOLVE"):
        assert names.count(nm) == 2, nm
    solves = sorted((e["ts"], e["ts"] + e["dur"]) for e in host if e["name"] == "SOLVE")
    assert solves[0][1] <= solves[1][0]
    classes = {e["name"] for e in gpu}
    assert {"gemm_f32", "gemm_f64", "ldlt_panel", "spmm"} <= classes
    phases = {e["args"]["phase"] for e in gpu}
    assert {"ge", "spectral", "select", "project", "si"} <= phases
    # every launch sits inside one of the two solves (the GPU clock is
    # placed on the host's through one event: allow 2 ms)
    def gap(e, s):
        mid = e["ts"] + 0.5 * e["dur"]
        return 0.0 if s[0] <= mid <= s[1] else min(abs(mid - s[0]), abs(mid - s[1]))

    per = [0, 0]
    for e in gpu:
        assert e["dur"] >= 0
        g = [gap(e, s) for s in solves]
        assert min(g) <= 2000, e
        per[int(np.argmin(g))] += 1
    # both solves trace the same launch sequence
    assert per[0] == per[1] and per[0] > 50
