"""Runs the benchmark workload's hot kernels in isolation, for counter
collection: `reps` launches of the CG-operator SpMV (FP64), `reps`
level-0 and `reps` level-1 symmetric smoothing sweeps, `reps` level-0
restrictions, `reps` level-0 prolongations and `reps` level-0 residuals in
the V-cycle precision of the headline (mixed) mode, then `reps` launches of
a plain device copy of known size (the calibration for FETCH_SIZE /
WRITE_SIZE). Prints the algorithmic bytes of each so a profiler summary can
be scaled.

With --headline-solve it instead runs one solve of the benchmark's headline
mode (mixed precision with the learned weights, every kernel its own
dispatch) after one untimed solve, then the same copies, so that a counter
run attributes HBM bytes to every kernel of a solve.

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


def copies(a) -> int:
    """`reps` launches of the plugin's k_copy<double> over copy_mib MiB;
    returns the element count."""
    n = a.copy_mib * (1 << 20) // 8
    be = gs.Backend(np.float64)
    x = be.create_vector(np.ones(n))
    y = be.create_vector(n)
    for _ in range(a.reps):
        e = be.create_encoder()
        be.encode_copy(e, x, y)
        be.submit_and_wait(e)
    return n


def headline_solve(a, H, b, sk, bench) -> None:
    mk = dict(sk)
    mk.update(coarse_direct=1, direct_truncate=1)
    m = gs.Solver(gs.PREC_MIXED, graph_iters=0, **mk)
    m.setup(H)
    w = bench.learned_weights(a.config, H)
    if w is not None:
        m.set_level_weights(omega=w[0], sweeps=w[1])
    m.solve(b, tol=1e-8, maxit=2000)
    x, info = m.solve(b, tol=1e-8, maxit=2000)
    n = copies(a)
    print(json.dumps({"config": a.config, "mode": "mixed_learned" if w is not None else "mixed",
                      "iterations": info.iterations, "solve_ms": info.solve_ms,
                      "copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": a.reps}))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="hetero_elasticity_96")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--precision", type=int, default=gs.PREC_DOUBLE)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--graph-iters", type=int, default=0)
    ap.add_argument("--headline-solve", action="store_true")
    a = ap.parse_args()
    import bench
    A, b, hk, sk, desc = bench.build_workload(a.config)
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, **hk)
    if a.headline_solve:
        headline_solve(a, H, b, sk, bench)
        return
    s = gs.Solver(a.precision, graph_iters=a.graph_iters, **sk)
    s.setup(H)
    s.solve(b, tol=1e-8, maxit=3)  # leaves valid vectors in p
    ms, by = s.time_kernel(0, a.reps)
    # Level-0 sweep of the mixed-precision solver (FP32 V-cycle).
    mk = dict(sk)
    mk.update(coarse_direct=1, direct_truncate=1)
    m = gs.Solver(gs.PREC_MIXED, graph_iters=a.graph_iters, **mk)
    m.setup(H)
    m.solve(b, tol=1e-8, maxit=3)
    sw_ms, sw_by = m.time_kernel(3, a.reps)
    sw1_ms, sw1_by = m.time_kernel(4, a.reps)
    r_ms, r_by = m.time_kernel(5, a.reps)
    p_ms, p_by = m.time_kernel(6, a.reps)
    res_ms, res_by = m.time_kernel(7, a.reps)
    # Calibration copy: n doubles read and written once.
    n = copies(a)
    print(json.dumps({"config": a.config, "spmv_ms": ms, "spmv_algorithmic_bytes": by,
                      "sweep_ms": sw_ms, "sweep_algorithmic_bytes": sw_by,
                      "sweep1_ms": sw1_ms, "sweep1_algorithmic_bytes": sw1_by,
                      "restrict_ms": r_ms, "restrict_algorithmic_bytes": r_by,
                      "prolong_ms": p_ms, "prolong_algorithmic_bytes": p_by,
                      "residual_ms": res_ms, "residual_algorithmic_bytes": res_by,
                      "copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "reps": a.reps}))


if __name__ == "__main__":
    main()
