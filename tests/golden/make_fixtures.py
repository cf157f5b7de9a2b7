"""Write the FP64 ground-truth eigenvalue fixtures used by tests and bench.py.

    python tests/golden/make_fixtures.py [--only hex8_10|hex27_10|hex27_14]

Ground truth follows ReferenceSolver::smallestEigenvalues
(src/reference/reference_solver.cpp:17-95): dense generalized eigh for
n <= 5000, else shift-invert Lanczos at sigma = 0 with ncv = min(3 nev + 20, n)
and tol = 1e-12 (ARPACK through scipy in place of Spectra).
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from amls_amd import cases  # noqa: E402
from oracle.amls_oracle import reference_eigs  # noqa: E402

FIXTURES = {
    "hex8_10": (cases.case_hex8_10, 20),
    "hex27_10": (cases.case_hex27_10, 100),
    "hex27_14": (cases.case_hex27_14, 50),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    for name, (make, nev) in FIXTURES.items():
        if a.only and a.only != name:
            continue
        pb = make()
        w = reference_eigs(pb.K, pb.M, nev)
        path = os.path.join(HERE, f"{name}_lowest{nev}.txt")
        with open(path, "w") as f:
            f.write(f"# {pb.meta['name']} n={pb.n} seed={pb.meta.get('seed', '-')} "
                    f"lowest {nev} eigenvalues of K x = lambda M x (FP64 ground truth)\n")
            for v in w:
                f.write(f"{v:.17e}\n")
        print(path, w[:3], w[-1])


if __name__ == "__main__":
    main()
