"""Benchmark: AMG-PCG solve time to a 1e-8 relative residual on one MI355X.

One "step" is one full preconditioned-CG solve (x0 = 0, tol 1e-8 on the
recursive relative residual, FP64 true residual verified afterwards) of the
selected workload, with the matrix hierarchy, right-hand side and all
vectors already resident in HBM. The headline mode is the FP32 V-cycle
inside FP64 CG with learned per-level smoother weights ("mixed+learned");
plain FP64 AMG-PCG on the same hierarchy is timed beside it. The hot
kernels are timed with HIP events on each solver's own stream and reported
against the 8 TB/s HBM roofline. The CPU oracle (a C++/OpenMP restatement
of the reference's CPU path) is timed on the same workload as a baseline,
in FP64 and in FP32 (the reference's own precision).

The other workloads run after the headline one (fewer steps) and go into
the "configs" object: plain FP64, mixed and mixed + learned device times,
iterations, FP64 true residuals, the CG SpMV's fraction of the HBM peak,
kernel launches per PCG iteration and each mode's error against an FP64
oracle solve. stderr carries one comparison table per workload, as the
reference's drivers print (examples/multigridPoisson/mainAMGtest.cc:786-834).

Counter collection (rocprofv3) runs the hot kernels without a timed region:
  --probe kernels        the headline workload's hot kernels, `--reps` times
                         each, then a 512 MiB device copy (the calibration)
  --probe headline-solve one headline-mode solve, every kernel its own
                         dispatch, then the same copies

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
                       [--configs all|none|a,b] [--no-cpu] [--probe MODE]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gpu-conjugate-gradient-solver_amd"))

METRIC = "AMG-PCG solve-time to 1e-8 relres, mixed+learned vs plain; SpMV % of HBM peak"
HBM_PEAK_GBS = 8000.0
HEADLINE = "hetero_elasticity_96"
OTHER_CONFIGS = ["poisson2d_256", "poisson3d_128", "elasticity_64", "varcoef3d_160"]


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
    l1_sweep, l0_restrict, l0_prolong, l0_residual) from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by the
    factor the same counter reads for a copy of known size), or None when
    none was taken for this workload and kernel."""
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


def learned_smoothers(config: str) -> dict:
    """Smoother options of the configuration's learned mode: those of the
    trained kind it uses (for the elasticity workloads the reference's
    hybrid block smoother -- block Gauss-Seidel on the finest level, block
    Jacobi below -- when that model is stored); {} without a model."""
    import learned
    kind = learned.config_kind(config)
    return learned.solver_options(kind) if kind else {}


def progress(msg: str) -> None:
    """One progress line per finished phase on stderr (stdout carries only
    the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_time(s, which: int, reps: int):
    """(ms per launch, algorithmic bytes) of one of the solver's kernels:
    the median of three batches of `reps` back-to-back launches timed with
    HIP events, so one disturbed batch does not set the figure."""
    runs = [s.time_kernel(which, reps) for _ in range(3)]
    return float(np.median([r[0] for r in runs])), runs[0][1]


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
    x = None
    for _ in range(warmup):
        s.solve(b, tol=tol, maxit=maxit)
    t0 = time.perf_counter()
    for _ in range(steps):
        x, info = s.solve(b, tol=tol, maxit=maxit)
        infos.append(info)
    t1 = time.perf_counter()
    return x, infos, (t1 - t0) * 1e3 / steps


def hier_view(H, levels=None):
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_api as orc
    L = H.levels if levels is None else levels
    Al = [H.get("A", k) for k in range(L)]
    return orc.HierView([(a.ptr, a.col, a.val) for a in Al],
                        [(p.ptr, p.col, p.val) for p in (H.get("P", k) for k in range(L - 1))],
                        [(r.ptr, r.col, r.val) for r in (H.get("R", k) for k in range(L - 1))])


def host_cpu() -> dict:
    """What the CPU baseline ran on: logical CPUs of the host, the CPUs this
    process may use, the CPU model and the OpenMP thread setting."""
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "host_logical_cpus": os.cpu_count(), "model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def oracle_solve(H, b, kinds, omega, coarse_steps, fp32=False, maxit=2000):
    """One plain-method solve by the oracle (a C++/OpenMP restatement of
    the reference's CPU solve) on the same hierarchy to 1e-8: FP64 (the
    parity reference) or FP32 throughout (the reference's own precision).
    Returns x, iterations, relres^2 and the wall time in ms (smoother setup
    included, as the reference's solver builds it per solve)."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_api as orc
    hv = hier_view(H)
    orc.lib.orc_set_parallel_dot(1)
    t0 = time.perf_counter()
    x, its, rr, _ = orc.pcg(hv, b, tol=1e-8, maxit=maxit, kinds=kinds, omegas=omega,
                            coarse_steps=coarse_steps, cg_double=not fp32, mg_double=not fp32)
    return x, its, rr, (time.perf_counter() - t0) * 1e3


def cpu_baseline(H, b, kinds, omega, coarse_steps):
    """FP64 and FP32 oracle solves of the headline workload on the host's
    cores; the FP64 solution is the reference every GPU mode's error is
    reported against."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_api as orc
    threads = orc.lib.orc_num_threads()
    xo, its, rr, ms = oracle_solve(H, b, kinds, omega, coarse_steps)
    progress(f"CPU oracle FP64: {ms:.0f} ms, {its} iterations")
    # FP32 throughout may stall above the FP64 target: capped at 1.5 x the
    # FP64 iteration count (reported as not converged then).
    x32, its32, rr32, ms32 = oracle_solve(H, b, kinds, omega, coarse_steps, fp32=True,
                                          maxit=int(1.5 * its) + 10)
    progress(f"CPU oracle FP32: {ms32:.0f} ms, {its32} iterations")
    cpu = {
        "value": round(ms, 1),
        "unit": "ms per solve",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full FP64 solve of the same workload (same hierarchy, smoother and "
                   f"tolerance) by the oracle on {threads} OpenMP threads: {its} iterations, "
                   f"relres^2 {rr:.3g}; includes its smoother setup (block inverses, colouring)"),
        "host": host_cpu(),
        "fp32_as_reference": {"ms": round(ms32, 1), "iterations": its32,
                              "relres2": float(f"{rr32:.4g}"),
                              "converged": bool(rr32 < 1e-16)},
    }
    return xo, cpu


def make_modes(gs, sk, graph_iters, config):
    """(name, precision, extra options, learned weights?) of each mode. When
    the learned mode runs another smoother family (the hybrid block
    smoother), "mixed_hybrid" runs that family with the plain weights, so
    the weights' own gain is visible."""
    mixed_kw = dict(sk)
    # FP64 dense direct solve at the first level within max_direct unknowns
    # (the inverse is built on the device); deeper levels are not used.
    mixed_kw["coarse_direct"] = 1
    mixed_kw["direct_truncate"] = 1
    learned_kw = dict(mixed_kw)
    learned_kw.update(learned_smoothers(config))
    modes = [("plain_fp64", gs.PREC_DOUBLE, dict(sk), False),
             ("fp64_direct", gs.PREC_DOUBLE, mixed_kw, False),
             ("mixed", gs.PREC_MIXED, mixed_kw, False)]
    if learned_kw != mixed_kw:
        modes.append(("mixed_hybrid", gs.PREC_MIXED, learned_kw, False))
    modes.append(("mixed_learned", gs.PREC_MIXED, learned_kw, True))
    return modes


def run_modes(config, H, b, sk, steps, warmup, graph_iters, names):
    """Sets up and times each named mode; returns (modes, solvers,
    solutions). The plain FP64 solver serves as the learned mode's
    fallback."""
    import gpusolver as gs
    modes, solvers, sols = {}, {}, {}
    for name, prec, kw, use_learned in make_modes(gs, sk, graph_iters, config):
        if name not in names:
            continue
        lw = None
        if use_learned:
            lw = learned_weights(config, H)
            if lw is None:
                continue
        t0 = time.perf_counter()
        s = gs.Solver(prec, graph_iters=graph_iters, **kw)
        s.setup(H)
        if lw is not None:
            s.set_level_weights(omega=lw[0], sweeps=lw[1])
            s.set_fallback(solvers["plain_fp64"], lag=1.5)
        setup_ms = (time.perf_counter() - t0) * 1e3
        x, infos, host_ms = time_solves(s, b, steps, warmup)
        d = s.describe()
        m = dict(ms=host_ms, device_ms=float(np.mean([i.solve_ms for i in infos])),
                 iterations=infos[-1].iterations, true_relres=infos[-1].true_relres,
                 launched_iterations=int(s.result("launched_iterations")),
                 setup_ms=setup_ms, launches_per_iteration=d["launches_per_iteration"],
                 levels_used=len(d["levels"]))
        if lw is not None:
            m.update(fell_back=s.fell_back(), omega=[round(float(w), 3) for w in lw[0]],
                     sweeps=[int(v) for v in lw[1]])
        modes[name] = m
        solvers[name] = s
        sols[name] = x
        progress(f"{config} {name}: {m['device_ms']:.2f} ms, {m['iterations']} iterations, "
                 f"{m['launches_per_iteration']:.0f} launches per iteration")
    return modes, solvers, sols


def comparison_table(config, modes, t_hier_ms):
    """The reference drivers' comparison table (examples/multigridPoisson/
    mainAMGtest.cc:786-834), on stderr: device solve time of the resident
    solve, host setup time of each solver."""
    w = 96
    lines = ["=" * w, f"PRECONDITIONER COMPARISON: {config}".center(w), "=" * w,
             f"{'Mode':<28}{'Iterations':>12}{'Setup (ms)':>14}{'Solve (ms)':>14}"
             f"{'Total (ms)':>14}{'True residual':>14}", "-" * w]
    for name, m in modes.items():
        lines.append(f"{name:<28}{m['iterations']:>12}{m['setup_ms']:>14.3f}"
                     f"{m['device_ms']:>14.3f}{m['setup_ms'] + m['device_ms']:>14.3f}"
                     f"{m['true_relres']:>14.3e}")
    lines += ["=" * w, f"Shared AMG preprocessing: hierarchy construction = {t_hier_ms:.1f} ms"]
    print("\n".join(lines), file=sys.stderr, flush=True)


def spmv_entry(s, reps=50):
    ms, by = kernel_time(s, 0, reps)
    gbs = by / ms / 1e6
    return {"avg_launch_ms": round(ms, 5), "algorithmic_bytes": by, "achieved_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def other_config(name, steps, warmup, graph_iters, with_oracle):
    """One non-headline workload: plain FP64, mixed and mixed + learned,
    the plain solver's CG SpMV against the roofline, and each mode's error
    against an FP64 oracle solve."""
    import gpusolver as gs
    t0 = time.perf_counter()
    A, b, hk, sk, desc = build_workload(name)
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, device=True, **hk)
    t_hier = (time.perf_counter() - t0) * 1e3
    modes, solvers, sols = run_modes(name, H, b, sk, steps, warmup, graph_iters,
                                     ("plain_fp64", "mixed", "mixed_hybrid", "mixed_learned"))
    out = {"description": desc, "n": int(A.rows), "nnz": int(A.nnz), "levels": H.levels,
           "hierarchy_ms": round(t_hier, 1),
           "cg_spmv": spmv_entry(solvers["plain_fp64"]),
           "modes": modes}
    head = "mixed_learned" if "mixed_learned" in modes else "mixed"
    out["speedup_vs_plain"] = round(modes["plain_fp64"]["device_ms"] / modes[head]["device_ms"], 3)
    if with_oracle:
        xo, its, rr, ms = oracle_solve(H, b, sk["smoother"], sk["omega"], sk["coarse_steps"])
        no = float(np.linalg.norm(xo))
        for k, x in sols.items():
            modes[k]["rel_err_vs_cpu_oracle"] = float(np.linalg.norm(x - xo) / no)
        out["oracle_fp64"] = {"ms": round(ms, 1), "iterations": its}
    for m in modes.values():
        for k, v in list(m.items()):
            if isinstance(v, float):
                m[k] = float(f"{v:.6g}")
    comparison_table(name, modes, t_hier)
    return out


def probe(a) -> None:
    """Counter-collection runs (no timed region): see the module docstring."""
    import gpusolver as gs
    A, b, hk, sk, desc = build_workload(a.config)
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, **hk)

    def copies() -> int:
        n = 512 * (1 << 20) // 8
        be = gs.Backend(np.float64)
        x = be.create_vector(np.ones(n))
        y = be.create_vector(n)
        for _ in range(a.reps):
            e = be.create_encoder()
            be.encode_copy(e, x, y)
            be.submit_and_wait(e)
        return n

    mk = dict(sk)
    mk.update(coarse_direct=1, direct_truncate=1)
    mk.update(learned_smoothers(a.config))
    if a.probe == "headline-solve":
        m = gs.Solver(gs.PREC_MIXED, graph_iters=0, **mk)
        m.setup(H)
        w = learned_weights(a.config, H)
        if w is not None:
            m.set_level_weights(omega=w[0], sweeps=w[1])
        m.solve(b, tol=1e-8, maxit=2000)
        x, info = m.solve(b, tol=1e-8, maxit=2000)
        n = copies()
        print(json.dumps({"config": a.config, "mode": "mixed_learned" if w is not None else "mixed",
                          "iterations": info.iterations, "solve_ms": info.solve_ms,
                          "copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                          "reps": a.reps}))
        return
    s = gs.Solver(gs.PREC_DOUBLE, graph_iters=0, **sk)
    s.setup(H)
    s.solve(b, tol=1e-8, maxit=3)  # leaves valid vectors in p
    res = {"config": a.config}
    res["spmv_ms"], res["spmv_algorithmic_bytes"] = s.time_kernel(0, a.reps)
    # Hot kernels of the mixed-precision solver (FP32 V-cycle).
    m = gs.Solver(gs.PREC_MIXED, graph_iters=0, **mk)
    m.setup(H)
    m.solve(b, tol=1e-8, maxit=3)
    for key, which in (("sweep", 3), ("sweep1", 4), ("restrict", 5), ("prolong", 6),
                       ("residual", 7)):
        try:
            res[f"{key}_ms"], res[f"{key}_algorithmic_bytes"] = m.time_kernel(which, a.reps)
        except gs.GpuSolverError:
            pass  # not present in this hierarchy
    n = copies()
    res.update(copy_bytes_read=n * 8, copy_bytes_written=n * 8, reps=a.reps)
    print(json.dumps(res))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=HEADLINE)
    ap.add_argument("--configs", default="all",
                    help="other workloads for the configs object: all, none, or a comma list")
    ap.add_argument("--config-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-iters", type=int, default=8,
                    help="PCG iterations per captured HIP graph (0: eager launches of the same kernels)")
    ap.add_argument("--probe", choices=["kernels", "headline-solve"])
    ap.add_argument("--reps", type=int, default=5, help="launches per kernel in --probe runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "the solve runs on one GPU; --gpus above 1 is not supported "
                                   "(a multi-GPU run would shard rows by contiguous colour "
                                   "blocks and exchange halos, which is not built)"}))
        sys.exit(2)
    if a.probe:
        probe(a)
        return
    import gpusolver as gs

    A, b, hk, sk, desc = build_workload(a.config)
    # The hierarchy with the products of each coarsening step on the GPU
    # (the one the solvers use), and the host-only build timed beside it
    # (the same hierarchy bit for bit).
    t0 = time.perf_counter()
    H = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, device=True, **hk)
    t_hier = time.perf_counter() - t0
    t0 = time.perf_counter()
    Hh = gs.Hierarchy.smoothed_aggregation(A, eps_strong=0.1, max_levels=10, **hk)
    t_hier_host = time.perf_counter() - t0
    del Hh
    progress(f"hierarchy built: {H.levels} levels, {t_hier:.1f} s with GPU products, "
             f"{t_hier_host:.1f} s on the host")

    modes, solvers, sols = run_modes(a.config, H, b, sk, a.steps, a.warmup, a.graph_iters,
                                     ("plain_fp64", "fp64_direct", "mixed", "mixed_hybrid",
                                      "mixed_learned"))
    plain = solvers["plain_fp64"]
    head = "mixed_learned" if "mixed_learned" in modes else "mixed"
    hm = modes[head]
    xp = sols["plain_fp64"]
    # Error of the headline solution against the plain FP64 GPU solve (both
    # verified by their FP64 true residual).
    err_vs_plain = float(np.linalg.norm(sols[head] - xp) / np.linalg.norm(xp))

    # Hot kernels, HIP events on each solver's own stream: the plain mode's
    # CG SpMV (FP64), and the headline solver's level-0 and level-1
    # symmetric Gauss-Seidel sweeps (one colour launch per colour and
    # direction), transfers and residual. The roofline entry is the
    # headline kernel with the largest time per PCG iteration.
    hs = solvers[head]
    hdesc = hs.describe()
    spmv_ms, spmv_bytes = kernel_time(plain, 0, 50)
    cands = [{"id": "cg_spmv", "kernel": "CG operator SpMV (FP64, plain mode)", "ms": spmv_ms,
              "bytes": spmv_bytes, "launches": 1, "per_iteration_ms": spmv_ms}]
    for lvl in (0, 1):
        if lvl >= len(hdesc["levels"]) or hdesc["levels"][lvl]["direct"]:
            continue
        lv = hdesc["levels"][lvl]
        if lv["pre"] + lv["post"] == 0:
            continue  # an unsmoothed level
        sw_ms, sw_bytes = kernel_time(hs, 3 + lvl, 20)
        store = 'lane groups of %d' % lv['group'] if lv['group'] else 'thread per row'
        if lv["kind"] in (1, 3):  # Gauss-Seidel: one launch per colour and direction
            cands.append({"id": f"l{lvl}_sweep",
                          "kernel": f"level-{lvl} Gauss-Seidel colour launches ({store}, FP32, "
                                    f"headline mode)",
                          "ms": sw_ms, "bytes": sw_bytes, "launches": 2 * lv["colors"] - 1,
                          "per_iteration_ms": sw_ms * (lv["pre"] + lv["post"])})
        else:  # Jacobi: the sweep and its copy back; the first sweep from 0 reads no matrix
            cands.append({"id": f"l{lvl}_jacobi",
                          "kernel": f"level-{lvl} Jacobi sweep + copy ({store}, FP32, "
                                    f"headline mode)",
                          "ms": sw_ms, "bytes": sw_bytes, "launches": 2,
                          "per_iteration_ms": sw_ms * (lv["pre"] + lv["post"] - 1)})
    if len(hdesc["levels"]) > 1:
        # Level-0 transfers of the headline solver, once each per V-cycle.
        l0 = hdesc["levels"][0]
        unit = "block row" if l0["transfer_b3"] else "row"
        store = "3x3 blocks" if l0["transfer_b3"] else "scalar"
        for kid, tid, what, grp in (("l0_restrict", 5, "restriction R0 (level 0 -> 1)", l0["r_group"]),
                                    ("l0_prolong", 6, "prolongation P0 (level 1 -> 0)", l0["p_group"])):
            t_ms, t_bytes = kernel_time(hs, tid, 20)
            cands.append({"id": kid,
                          "kernel": f"level-0 {what} ({store}, "
                                    f"{'lane groups of %d' % grp if grp else 'thread per %s, SELL-C-sigma' % unit}"
                                    f", FP32, headline mode)",
                          "ms": t_ms, "bytes": t_bytes, "launches": 1, "per_iteration_ms": t_ms})
    if len(hdesc["levels"]) > 1 and not hdesc["levels"][0]["direct"]:
        # Level-0 residual rhs - A x of the V-cycle (the headline mode's
        # operator SpMV, once per V-cycle).
        r_ms, r_bytes = kernel_time(hs, 7, 20)
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
    progress("headline kernels timed")
    comparison_table(a.config, modes, t_hier * 1e3)

    cpu = None
    if not a.no_cpu:
        xo, cpu = cpu_baseline(H, b, sk["smoother"], sk["omega"], sk["coarse_steps"])
        no = float(np.linalg.norm(xo))
        for name, xs in sols.items():
            modes[name]["rel_err_vs_cpu_oracle"] = float(np.linalg.norm(xs - xo) / no)
    setup_s = {"hierarchy": round(t_hier, 2), "hierarchy_host_only": round(t_hier_host, 2),
               "plain_upload": round(modes["plain_fp64"]["setup_ms"] / 1e3, 2),
               "headline_upload": round(hm["setup_ms"] / 1e3, 2),
               "plain_phases_ms": {k: round(v, 1) for k, v in desc_plain["setup_ms"].items()},
               "headline_phases_ms": {k: round(v, 1) for k, v in hdesc["setup_ms"].items()}}
    del solvers, sols, hs, plain

    configs = {}
    names = OTHER_CONFIGS if a.configs == "all" else \
        ([] if a.configs == "none" else [c for c in a.configs.split(",") if c])
    for name in names:
        if name == a.config:
            continue
        configs[name] = other_config(name, a.config_steps, 1, a.graph_iters, not a.no_cpu)

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
        "setup_s": setup_s,
        "levels": desc_plain["levels"],
        "configs": configs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
