"""Host-CPU timings of the 26,460-dof hexahedral workload -- test
infrastructure, run by bench.py as a child process before it touches the
GPU (the oracle is the checker and the CPU baseline, never the product).

Three timings, each from the pencil in memory (the library's host-side
nested dissection and symbolic analysis included) to the eigenvalues out,
on every CPU this process may use (at most 16):
  1. plain AMLS as the reference runs it: two-pass truncation at its default
     eta = 2, the fused pipeline per pass, node-level parallelism with BLAS
     single-threaded (src/amls/solver.cpp:100-319, 136-143, 199-212), back-
     transformation and M-normalisation of the vectors;
  2. the GPU value's algorithm: the same AMLS at eta = 4, then two subspace-
     iteration steps from the whole AMLS basis with a Rayleigh-Ritz after
     each (their block products split by columns over the same workers);
  3. the FP64 ground truth: shift-invert Lanczos at sigma = 0, tol 1e-12
     (src/reference/reference_solver.cpp:17-95; ARPACK through scipy), on
     one core (its sparse factorisation and recurrences are serial).

    python -m oracle.cpu_baseline [--legs 1,2,3]
prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--legs", default="1,2,3")
    ap.add_argument("--workers", type=int, default=0)
    args = ap.parse_args()
    legs = {int(x) for x in args.legs.split(",") if x}
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    from threadpoolctl import threadpool_limits

    from amls_amd import HipContext, cases
    from oracle import amls_oracle as O
    from oracle.parallel import ParallelAmls, ParallelSubspaceIteration, host_workers

    W = args.workers or host_workers()
    pb = cases.case_hex27_10()
    ref = np.loadtxt(os.path.join(root, "tests", "golden", "hex27_10_lowest100.txt"))
    nev = 100

    def tree():
        t0 = time.perf_counter()
        c = HipContext(-1)
        c.set_problem(pb.K, pb.M, pb.coords, 6, 500)
        tr = O.Tree.from_dict(c.tree())
        c.close()
        return tr, time.perf_counter() - t0

    out = {"cores": W, "host_cpus": W, "host_cpus_visible": os.cpu_count(), "kind": "port"}
    if 1 in legs:
        with threadpool_limits(1):
            t0 = time.perf_counter()
            tr, t_part = tree()
            with ParallelAmls(pb.K, pb.M, tr, W) as P:
                r = P.solve(nev=nev, thresholdFactor=2.0, recoverVectors=True)
            t = time.perf_counter() - t0
        out["plain_amls_eta2"] = {"value": round(t, 3), "unit": "s/solve", "cores": W, "p": int(r.p),
                                  "partition_s": round(t_part, 3),
                                  "phases_s": {k: round(v, 3) for k, v in r.timing.items()},
                                  "max_rel_err": float(O.rel_err(r.theta, ref).max())}
    if 2 in legs:
        with threadpool_limits(1):
            t0 = time.perf_counter()
            tr, _ = tree()
            P = ParallelAmls(pb.K, pb.M, tr, W)
            r = P.solve(nev=nev, thresholdFactor=4.0, recoverVectors=False)
            P.close()
            t1 = time.perf_counter()
            _, Q = r.cond.solve_projected(r.p)
            z = r.cond.backtransform(Q)
            X = np.zeros_like(z)
            X[tr.perm] = z
            t2 = time.perf_counter()
            with ParallelSubspaceIteration(r.ge, P.Kr, P.Mr, r.p, W) as S:
                th, _ = S.run(X, 2, nev, tr.perm)
            t = time.perf_counter() - t0
        out["amls_si_eta4"] = {"value": round(t, 3), "unit": "s/solve", "cores": W, "p": int(r.p),
                               "si_steps": 2, "amls_s": round(t1 - t0, 3),
                               "basis_s": round(t2 - t1, 3),
                               "max_rel_err": float(O.rel_err(th, ref).max())}
    if 3 in legs:
        # one thread: the sparse LU and the Lanczos recurrences are serial,
        # and raising the BLAS thread count after the worker pools forked
        # crashes scipy's OpenBLAS
        with threadpool_limits(1):
            t0 = time.perf_counter()
            w = O.reference_eigs(pb.K, pb.M, nev)
            t = time.perf_counter() - t0
        out["lanczos"] = {"value": round(t, 3), "unit": "s/solve", "cores": 1,
                          "max_rel_err": float(O.rel_err(w, ref).max())}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
