"""Runs the benchmark workload's dominant kernel in isolation, for counter
collection: `reps` launches of the CG-operator SpMV of the selected
workload, then `reps` launches of a plain device copy of known size (the
calibration for FETCH_SIZE / WRITE_SIZE). Prints the algorithmic bytes of
both so a profiler summary can be scaled.

Usage (under rocprofv3):
  rocprofv3 --pmc FETCH_SIZE -d OUT -- python gpu-conjugate-gradient-solver_amd/kernel_probe.py
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))
import gpusolver as gs  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="hetero_elasticity_96")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--precision", type=int, default=gs.PREC_DOUBLE)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--graph-iters", type=int, default=0)
    a = ap.parse_args()
    import bench
    A, b, hk, sk, desc = bench.build_workload(a.config)
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, **hk)
    s = gs.Solver(a.precision, graph_iters=a.graph_iters, **sk)
    s.setup(H)
    s.solve(b, tol=1e-8, maxit=3)  # leaves valid vectors in p
    ms, by = s.time_kernel(0, a.reps)
    # Calibration copy: n doubles read and written once.
    n = a.copy_mib * (1 << 20) // 8
    be = gs.Backend(np.float64)
    x = be.create_vector(np.ones(n))
    y = be.create_vector(n)
    for _ in range(a.reps):
        e = be.create_encoder()
        be.encode_copy(e, x, y)
        be.submit_and_wait(e)
    print(json.dumps({"config": a.config, "spmv_ms": ms, "spmv_algorithmic_bytes": by,
                      "copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "reps": a.reps}))


if __name__ == "__main__":
    main()
