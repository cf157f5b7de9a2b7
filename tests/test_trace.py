"""The context's Chrome trace on the host side: a host-only context records
the set-up analysis span and writes a well-formed trace file (the format of
src/core/trace.hpp:143-181); the GPU launch spans are covered by the GPU
tests."""
import json
import os
import subprocess
import sys

from amls_amd import HipContext, cases


def test_host_trace_file(tmp_path):
    pb = cases.hex_elasticity(1, 4, 4, 4)
    path = tmp_path / "t.json"
    c = HipContext(-1)
    c.trace_begin(str(path))
    c.set_problem(pb.K, pb.M, pb.coords, 3, 40)
    n = c.trace_end()
    c.close()
    assert n == 1
    ev = json.loads(path.read_text())["traceEvents"]
    xs = [e for e in ev if e["ph"] == "X"]
    assert [e["name"] for e in xs] == ["ANALYSIS"]
    assert xs[0]["tid"] == 0 and xs[0]["cat"] == "setup" and xs[0]["dur"] >= 0
    names = {e["args"]["name"] for e in ev if e["ph"] == "M"}
    assert names == {"host", "gpu stream 1", "gpu stream 2"}


def test_trace_end_without_begin_is_state_error():
    from amls_amd import AmlsError
    c = HipContext(-1)
    try:
        c.trace_end()
        raise AssertionError("expected AmlsError")
    except AmlsError as e:
        assert e.status == 6
    c.close()


def test_env_variable_writes_on_destroy(tmp_path):
    path = tmp_path / "env.json"
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from amls_amd import HipContext, cases\n"
            "pb = cases.hex_elasticity(1, 3, 3, 3)\n"
            "c = HipContext(-1); c.set_problem(pb.K, pb.M, pb.coords, 2, 30); c.close()\n"
            % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    env = dict(os.environ, AMLS_TRACE=str(path))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "[trace] 1 events" in r.stderr
    ev = json.loads(path.read_text())["traceEvents"]
    assert any(e.get("name") == "ANALYSIS" for e in ev)
