"""The command-line drivers' host side: amls_solve's argument handling and
amls_verify's report formats (src/apps/amls_solve.cpp:35-86,
src/apps/amls_verify.cpp:128-158) and error metrics
(src/reference/error_metrics.cpp:6-57)."""
import numpy as np
import pytest

from amls_amd import apps, cases


def test_parse_solve_arguments():
    a = apps.parse_solve(["--case", "grid:8,6", "--nev", "7", "--nlevels", "3", "--eta", "3.5",
                          "--shift", "0.25", "--fixed-keep", "4", "--leaf-size", "40",
                          "--no-recover", "--quiet", "--out", "t.txt", "--si-steps", "2"])
    o = a["opt"]
    assert a["case"] == ("grid", [8, 6]) and a["out"] == "t.txt"
    assert (o.nev, o.nLevels, o.thresholdFactor, o.shift, o.fixedKeep, o.leafSize) == \
        (7, 3, 3.5, 0.25, 4, 40)
    assert not o.recoverVectors and o.verbosity == 0 and o.siSteps == 2
    for bad in (["--nev"], ["--case", "path"], ["--bogus"], []):
        with pytest.raises(apps.UsageError):
            apps.parse_solve(bad)


def test_solve_main_usage_exit_code(capsys):
    assert apps.solve_main(["--bogus"]) == 2
    assert "usage" in capsys.readouterr().err


def test_load_case_spectra():
    K, M, ref = apps.load_case("path", [50], False)
    assert K.shape == (50, 50) and np.allclose(ref, cases.path_graph_spectrum(50))
    K, M, ref = apps.load_case("grid", [5, 4], False)
    assert K.shape == (20, 20) and len(ref) == 20


def test_error_metrics():
    th = np.array([1.01, 4.0, 9.5])
    ref = np.array([1.0, 4.0, 9.0, 16.0])
    e = apps.eigenvalue_rel_err(th, ref)
    assert np.allclose(e, [0.01, 0.0, 0.5 / 9.0])
    f = apps.frequency_rel_err(th, ref)
    assert np.allclose(f, np.abs(np.sqrt(th) - np.sqrt(ref[:3])) / np.sqrt(ref[:3]))
    # zero reference: absolute error
    assert apps.eigenvalue_rel_err(np.array([0.5]), np.array([0.0]))[0] == 0.5
    K, M = cases.path_graph(30)
    w = cases.path_graph_spectrum(30)
    import scipy.linalg as sla
    lam, V = sla.eigh(K.toarray(), M.toarray())
    assert apps.residual(K, M, lam[0], V[:, 0]) < 1e-13
    assert np.allclose(apps.reference_eigenvalues(K, M, 5), w[:5], rtol=1e-12)


def test_report_formats(tmp_path):
    rows = [apps.Row("path2000", 2.0, 4, p=48, maxRelErr=0.0123456789, maxFreqErr=0.006,
                     phase1=0.5, phase2=0.01, back=0.02, partition=0.003, passed=True),
            apps.Row("grid64x64", 1.2, 4, p=30, maxRelErr=0.08, maxFreqErr=0.04, passed=False)]
    apps.write_report(rows, str(tmp_path))
    csv = (tmp_path / "verify_results.csv").read_text().splitlines()
    assert csv[0] == ("case,eta,nLevels,p,maxRelErr,maxFreqErr,phase1_s,phase2_s,back_s,"
                      "partition_s,pass")
    assert csv[1] == "path2000,2,4,48,0.0123457,0.006,0.5,0.01,0.02,0.003,1"
    assert csv[2].startswith("grid64x64,1.2,4,30,0.08,0.04,") and csv[2].endswith(",0")
    md = (tmp_path / "verify_report.md").read_text()
    assert "| 算例 | η | 层数 | p | max ε_λ | max ε_f | Phase1(s) | 通过 |" in md
    assert "| path2000 | 2 | 4 | 48 | 0.0123457 | 0.006 | 0.5 | ✅ |" in md
    assert md.rstrip().endswith("通过 1/2。")
