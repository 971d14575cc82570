"""Benchmark: AMG-PCG solve time to a 1e-8 relative residual on one MI355X.

One "step" is one full preconditioned-CG solve (x0 = 0, tol 1e-8 on the
recursive relative residual, FP64 true residual verified afterwards) of the
selected workload, with the matrix hierarchy, right-hand side and all
vectors already resident in HBM. The headline mode is the FP32 V-cycle
inside FP64 CG with learned per-level smoother weights ("mixed+learned");
plain FP64 AMG-PCG on the same hierarchy is timed beside it. The dominant
kernel (the CG operator's SpMV) is timed with HIP events on the solver's
own stream and reported against the 8 TB/s HBM roofline. The CPU oracle
(a C++/OpenMP restatement of the reference's CPU path) is timed on a
bounded sample as a baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gpu-conjugate-gradient-solver_amd"))

METRIC = "AMG-PCG solve-time to 1e-8 relres, mixed+learned vs plain; SpMV % of HBM peak"
HBM_PEAK_GBS = 8000.0


def build_workload(name: str):
    """Returns (A, b, hierarchy kwargs, solver kwargs, description)."""
    import gpusolver as gs
    if name == "poisson2d_256":
        A = gs.poisson2d(256, 256)
        b = gs.uniform_pm1(1, A.rows)
        return A, b, dict(block_size=1, min_dofs=1000), dict(smoother=gs.SGS, coarse_steps=2,
                                                              omega=1.0), \
            "2D 5-point Poisson 256x256 (65 536 unknowns), SA(1, 0.1), MG-SGS 1/1"
    if name == "poisson3d_128":
        A = gs.poisson3d(128)
        b = gs.uniform_pm1(1, A.rows)
        return A, b, dict(block_size=1, min_dofs=1000), dict(smoother=gs.SGS, coarse_steps=2,
                                                              omega=1.0), \
            "3D 7-point Poisson 128^3 (2 097 152 unknowns), SA(1, 0.1), MG-SGS 1/1"
    if name == "varcoef3d_160":
        A = gs.varcoef3d(160, seed=200, sigma=8.0, scale=2.0)
        b = gs.uniform_pm1(1, A.rows)
        return A, b, dict(block_size=1, min_dofs=1000), dict(smoother=gs.SGS, coarse_steps=2,
                                                              omega=1.0), \
            "3D FV -div(k grad u) 160^3, k = exp(2 g), SA(1, 0.1), MG-SGS 1/1"
    if name in ("elasticity_64", "hetero_elasticity_96"):
        nv = 64 if name == "elasticity_64" else 96
        hs = 0.0 if nv == 64 else 6.0
        A, b, _ = gs.elasticity(nv, young=1.0, poisson=0.3, hetero_sigma=hs, hetero_scale=1.5,
                                seed=1)
        desc = (f"P1-tet elasticity on a {nv}^3-vertex Kuhn cube ({A.rows} DOF, 3x3 blocks"
                f"{', E = exp(1.5 g)' if hs else ''}), SA(3, 0.1) minDofs 5000, MG block-GS 1/1 w=0.7")
        return A, b, dict(block_size=3, min_dofs=5000), dict(smoother=gs.BLOCK_GS, coarse_steps=0,
                                                              omega=0.7, cg_block3=1), desc
    raise SystemExit(f"unknown config {name}")


def measured_traffic(config: str, kernel_id: str):
    """HBM bytes per launch of kernel `kernel_id` (cg_spmv, l0_sweep,
    l1_sweep, l0_restrict, l0_prolong, l0_residual) from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the factor the same
    counter reads for a copy of known size), or None when none was taken
    for this workload and kernel."""
    path = REPO / "profiles" / f"pmc_{config}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic|"):
            _, name, val = line.split("|")
            if name.strip() == kernel_id:
                return float(val)
    return None


def learned_weights(config: str, H):
    """Per-level (omega, sweeps) from the trained smoother model (GPU
    inference), or None when no model has been trained for this smoother."""
    import learned
    return learned.weights_for(config, H)


def progress(msg: str) -> None:
    """One progress line per finished phase on stderr (stdout carries only
    the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def progress_mode(modes: dict, name: str) -> None:
    m = modes[name]
    progress(f"{name}: {m['device_ms']:.2f} ms, {m['iterations']} iterations")


def copy_bandwidth(mib: int = 512, reps: int = 10) -> float:
    """GB/s of a plain device copy (one FP64 element per thread, the kernel
    the counter calibration uses), `reps` copies of `mib` MiB back to back
    on one stream, host-timed around launch and completion: a practical
    streaming ceiling next to the published peak."""
    import gpusolver as gs
    n = mib * (1 << 20) // 8
    be = gs.Backend(np.float64)
    x = be.create_vector(np.ones(n))
    y = be.create_vector(n)
    e = be.create_encoder()
    be.encode_copy(e, x, y)
    be.submit_and_wait(e)
    e = be.create_encoder()
    t0 = time.perf_counter()
    for _ in range(reps):
        be.encode_copy(e, x, y)
    be.submit_and_wait(e)
    t = time.perf_counter() - t0
    return 2.0 * n * 8 * reps / t / 1e9


def time_solves(s, b, steps, warmup, tol=1e-8, maxit=2000):
    infos = []
    for _ in range(warmup):
        s.solve(b, tol=tol, maxit=maxit)
    t0 = time.perf_counter()
    for _ in range(steps):
        x, info = s.solve(b, tol=tol, maxit=maxit)
        infos.append(info)
    t1 = time.perf_counter()
    return x, infos, (t1 - t0) * 1e3 / steps


def cpu_baseline(H, b, kinds, omega, coarse_steps):
    """The oracle (a C++/OpenMP restatement of the reference's CPU solve)
    runs the plain FP64 method on the same hierarchy to 1e-8: its wall time
    is the CPU baseline and its solution the FP64 reference every GPU mode's
    error is reported against."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_api as orc
    L = H.levels
    Al = [H.get("A", k) for k in range(L)]
    hv = orc.HierView([(a.ptr, a.col, a.val) for a in Al],
                      [(p.ptr, p.col, p.val) for p in (H.get("P", k) for k in range(L - 1))],
                      [(r.ptr, r.col, r.val) for r in (H.get("R", k) for k in range(L - 1))])
    orc.lib.orc_set_parallel_dot(1)
    threads = orc.lib.orc_num_threads()
    t0 = time.perf_counter()
    x, its, rr, _ = orc.pcg(hv, b, tol=1e-8, maxit=2000, kinds=kinds, omegas=omega,
                            coarse_steps=coarse_steps)
    t = time.perf_counter() - t0
    return x, {
        "value": round(t * 1e3, 1),
        "unit": "ms per solve",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full FP64 solve of the same workload (same hierarchy, smoother and "
                   f"tolerance) by the oracle on {threads} OpenMP threads: {its} iterations, "
                   f"relres^2 {rr:.3g}; includes its smoother setup (block inverses, colouring)"),
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="hetero_elasticity_96")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-iters", type=int, default=8,
                    help="PCG iterations per captured HIP graph (0: eager launches of the same kernels)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "the solve runs on one GPU; --gpus above 1 is not supported "
                                   "(a multi-GPU run would shard rows by contiguous colour "
                                   "blocks and exchange halos, which is not built)"}))
        sys.exit(2)
    import gpusolver as gs

    t0 = time.perf_counter()
    A, b, hk, sk, desc = build_workload(a.config)
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, **hk)
    t_hier = time.perf_counter() - t0
    progress(f"hierarchy built: {H.levels} levels, {t_hier:.1f} s")

    modes = {}
    # Plain FP64 AMG-PCG, the reference's method in FP64.
    t0 = time.perf_counter()
    plain = gs.Solver(gs.PREC_DOUBLE, graph_iters=a.graph_iters, **sk)
    plain.setup(H)
    t_setup_plain = time.perf_counter() - t0
    xp, ip, ms_plain = time_solves(plain, b, a.steps, a.warmup)
    modes["plain_fp64"] = dict(ms=ms_plain, iterations=ip[-1].iterations,
                               true_relres=ip[-1].true_relres,
                               device_ms=float(np.mean([i.solve_ms for i in ip])))
    progress_mode(modes, "plain_fp64")
    mixed_kw = dict(sk)
    # FP64 dense direct solve at the first level within max_direct unknowns
    # (the inverse is built on the device); deeper levels are not used.
    mixed_kw["coarse_direct"] = 1
    mixed_kw["direct_truncate"] = 1
    # FP64 throughout with the FP64 direct coarsest solve: separates the gain of
    # the direct coarsest solve from that of the FP32 V-cycle.
    fd = gs.Solver(gs.PREC_DOUBLE, graph_iters=a.graph_iters, **mixed_kw)
    fd.setup(H)
    xd, idd, ms_fd = time_solves(fd, b, a.steps, a.warmup)
    modes["fp64_direct"] = dict(ms=ms_fd, iterations=idd[-1].iterations,
                                true_relres=idd[-1].true_relres,
                                device_ms=float(np.mean([i.solve_ms for i in idd])))
    progress_mode(modes, "fp64_direct")
    del fd
    # Mixed: FP32 V-cycle in FP64 CG, FP64 dense direct coarse solve.
    mixed = gs.Solver(gs.PREC_MIXED, graph_iters=a.graph_iters, **mixed_kw)
    mixed.setup(H)
    xm, im, ms_mixed = time_solves(mixed, b, a.steps, a.warmup)
    modes["mixed"] = dict(ms=ms_mixed, iterations=im[-1].iterations,
                          true_relres=im[-1].true_relres,
                          device_ms=float(np.mean([i.solve_ms for i in im])))
    progress_mode(modes, "mixed")
    head = "mixed"
    ml = None
    lw = learned_weights(a.config, H)
    if lw is not None:
        ml = gs.Solver(gs.PREC_MIXED, graph_iters=a.graph_iters, **mixed_kw)
        ml.setup(H)
        ml.set_level_weights(omega=lw[0], sweeps=lw[1])
        ml.set_fallback(plain, lag=1.5)
        xl, il, ms_l = time_solves(ml, b, a.steps, a.warmup)
        modes["mixed_learned"] = dict(ms=ms_l, iterations=il[-1].iterations,
                                      true_relres=il[-1].true_relres,
                                      device_ms=float(np.mean([i.solve_ms for i in il])),
                                      fell_back=ml.fell_back(),
                                      omega=[round(float(w), 3) for w in lw[0]],
                                      sweeps=[int(v) for v in lw[1]])
        progress_mode(modes, "mixed_learned")
        head = "mixed_learned"
    hm = modes[head]
    sols = {"plain_fp64": xp, "fp64_direct": xd, "mixed": xm}
    if head == "mixed_learned":
        sols["mixed_learned"] = xl
    # Error of the headline solution against the plain FP64 GPU solve (both
    # verified by their FP64 true residual).
    err_vs_plain = float(np.linalg.norm(sols[head] - xp) / np.linalg.norm(xp))

    # Hot kernels, HIP events on each solver's own stream: the plain mode's
    # CG SpMV (FP64), and the headline solver's level-0 and level-1
    # symmetric Gauss-Seidel sweeps (one colour launch per colour and
    # direction). The roofline entry is the headline kernel with the largest
    # time per PCG iteration (sweep time x sweeps per V-cycle).
    hs = ml if head == "mixed_learned" else mixed
    hdesc = hs.describe()
    spmv_ms, spmv_bytes = plain.time_kernel(0, 50)
    cands = [{"id": "cg_spmv", "kernel": "CG operator SpMV (FP64, plain mode)", "ms": spmv_ms,
              "bytes": spmv_bytes, "launches": 1, "per_iteration_ms": spmv_ms}]
    for lvl in (0, 1):
        if lvl >= len(hdesc["levels"]) or hdesc["levels"][lvl]["direct"]:
            continue
        lv = hdesc["levels"][lvl]
        sw_ms, sw_bytes = hs.time_kernel(3 + lvl, 20)
        cands.append({"id": f"l{lvl}_sweep",
                      "kernel": f"level-{lvl} Gauss-Seidel colour launches ("
                                f"{'lane groups of %d' % lv['group'] if lv['group'] else 'thread per row'}"
                                f", FP32, headline mode)",
                      "ms": sw_ms, "bytes": sw_bytes, "launches": 2 * lv["colors"] - 1,
                      "per_iteration_ms": sw_ms * (lv["pre"] + lv["post"])})
    if len(hdesc["levels"]) > 1:
        # Level-0 transfers of the headline solver, once each per V-cycle.
        l0 = hdesc["levels"][0]
        unit = "block row" if l0["transfer_b3"] else "row"
        store = "3x3 blocks" if l0["transfer_b3"] else "scalar"
        for kid, tid, what, grp in (("l0_restrict", 5, "restriction R0 (level 0 -> 1)", l0["r_group"]),
                                    ("l0_prolong", 6, "prolongation P0 (level 1 -> 0)", l0["p_group"])):
            t_ms, t_bytes = hs.time_kernel(tid, 20)
            cands.append({"id": kid,
                          "kernel": f"level-0 {what} ({store}, "
                                    f"{'lane groups of %d' % grp if grp else 'thread per %s, SELL-C-sigma' % unit}"
                                    f", FP32, headline mode)",
                          "ms": t_ms, "bytes": t_bytes, "launches": 1, "per_iteration_ms": t_ms})
    if len(hdesc["levels"]) > 1 and not hdesc["levels"][0]["direct"]:
        # Level-0 residual rhs - A x of the V-cycle (the headline mode's
        # operator SpMV, once per V-cycle).
        r_ms, r_bytes = hs.time_kernel(7, 20)
        grp = hdesc["levels"][0]["group"]
        cands.append({"id": "l0_residual",
                      "kernel": f"level-0 residual SpMV rhs - A x ("
                                f"{'lane groups of %d' % grp if grp else 'thread per row'}"
                                f", FP32, headline mode)",
                      "ms": r_ms, "bytes": r_bytes, "launches": 1, "per_iteration_ms": r_ms})
    dom = max(cands[1:] or cands, key=lambda c: c["per_iteration_ms"])
    copy_gbs = copy_bandwidth()
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    desc_plain = plain.describe()

    cpu = None
    if not a.no_cpu:
        progress("kernels timed; CPU baseline running")
        xo, cpu = cpu_baseline(H, b, sk["smoother"], sk["omega"], sk["coarse_steps"])
        no = float(np.linalg.norm(xo))
        for name, xs in sols.items():
            modes[name]["rel_err_vs_cpu_oracle"] = float(np.linalg.norm(xs - xo) / no)

    # Inputs resident in HBM: the device (HIP-event) time of the solve loop;
    # the host time per step additionally holds the b / x transfers, the
    # ordering gather / scatter and the FP64 true-residual check.
    value = hm["device_ms"]
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(hm["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 CG / f32 V-cycle" if head != "plain_fp64" else "f64",
        "data": "synthetic (seeded generators; no external data)",
        "config": {"workload": a.config, "description": desc, "n": int(A.rows), "nnz": int(A.nnz),
                   "levels": H.levels, "tol": 1e-8, "headline_mode": head},
        "modes": {k: {kk: (float(f"{vv:.6g}") if isinstance(vv, float) else vv)
                      for kk, vv in v.items()}
                  for k, v in modes.items()},
        "speedup_vs_plain": round(modes["plain_fp64"]["device_ms"] / value, 3),
        "rel_err_vs_plain_fp64": err_vs_plain,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(a.config, dom["id"]),
                     "kernel": dom["kernel"], "launches": dom["launches"],
                     "kernel_ms": round(dom["ms"] / dom["launches"], 5),
                     "algorithmic_bytes": dom["bytes"] / dom["launches"]},
        "kernels": [{"kernel": c["kernel"], "achieved_GBs": round(c["bytes"] / c["ms"] / 1e6, 1),
                     "frac": round(c["bytes"] / c["ms"] / 1e6 / HBM_PEAK_GBS, 4),
                     "avg_launch_ms": round(c["ms"] / c["launches"], 5),
                     "launches": c["launches"],
                     "per_iteration_ms": round(c["per_iteration_ms"], 4),
                     "traffic": measured_traffic(a.config, c["id"])} for c in cands],
        "cpu_baseline": cpu,
        # The metric's "SpMV % of HBM peak": the FP64 CG-operator SpMV.
        "spmv": {"kernel": cands[0]["kernel"],
                 "achieved_GBs": round(cands[0]["bytes"] / cands[0]["ms"] / 1e6, 1),
                 "frac": round(cands[0]["bytes"] / cands[0]["ms"] / 1e6 / HBM_PEAK_GBS, 4),
                 "traffic": measured_traffic(a.config, "cg_spmv")},
        "hbm_copy": {"GBs": round(copy_gbs, 1), "frac": round(copy_gbs / HBM_PEAK_GBS, 4),
                     "kernel": "k_copy<double>, 10 x 512 MiB, host-timed"},
        "setup_s": {"hierarchy": round(t_hier, 2), "plain_upload": round(t_setup_plain, 2)},
        "levels": desc_plain["levels"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
