"""The command-line drivers end to end on the GPU: amls_solve on a closed-
form case writes "j theta" lines that match the spectrum, and amls_verify's
quick closed-form suite writes both reports with every case passing the
reference's 4 % rule."""
import subprocess
import sys

import numpy as np
import pytest

from amls_amd import apps, cases

pytestmark = pytest.mark.gpu


def test_solve_driver_path_case(tmp_path, capsys):
    out = tmp_path / "theta.txt"
    rc = apps.solve_main(["--case", "path:2000", "--nev", "12", "--nlevels", "4", "--eta", "4",
                          "--si-steps", "2", "--out", str(out)])
    assert rc == 0
    lines = out.read_text().splitlines()
    assert len(lines) == 12 and lines[0].startswith("1 ")
    th = np.array([float(l.split()[1]) for l in lines])
    ref = cases.path_graph_spectrum(2000)[:12]
    assert (np.abs(th - ref) / ref).max() <= 1e-5
    text = capsys.readouterr().out
    assert "p=" in text and "relerr=" in text and "# residual:" in text


def test_verify_driver_quick_closed(tmp_path):
    """The quick closed-form suite at the default nev (50): one row per case,
    each row's pass flag is the 4 % rule on its own error, the exit code says
    whether all passed, and the path cases pass."""
    r = subprocess.run([sys.executable, "-m", "amls_amd.verify", "--suite", "closed", "--quick",
                        "--outdir", str(tmp_path)], capture_output=True, text=True, timeout=300)
    csv = (tmp_path / "verify_results.csv").read_text().splitlines()
    assert csv[0].startswith("case,eta,nLevels,p,maxRelErr")
    rows = [l.split(",") for l in csv[1:]]
    assert [x[0] for x in rows] == ["path2000", "path2000-consM", "grid64x64", "path500"]
    flags = [x[-1] == "1" for x in rows]
    for x, f in zip(rows, flags):
        assert f == (0 <= float(x[4]) <= 0.04)
    assert flags[0] and flags[1] and flags[3]
    assert r.returncode == (0 if all(flags) else 1), r.stdout + r.stderr
    md = (tmp_path / "verify_report.md").read_text()
    assert f"通过 {sum(flags)}/4。" in md
