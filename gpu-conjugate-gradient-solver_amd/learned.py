"""Learned per-level smoother weights for the mixed-precision V-cycle.

A 5 -> 64 -> 64 -> 2 MLP (HIP forward / backward / Adam, gs_mlp_* in
include/gpusolver.h) maps five per-level features of an AMG hierarchy
(mean row nnz, diagonal dominance, coarsening ratio, coefficient contrast,
colour count) to that level's smoother damping w and its pre = post sweep
count. Labels come from the solver itself: for each seeded training
problem a coordinate search over (w, sweeps) per level keeps the setting
with the lowest measured device time (HIP events, best of two) of a
mixed-precision solve to 1e-8; a candidate that does not converge or hits
a CG breakdown is rejected. The preconditioner stays a fixed
symmetric operator within a solve (pre = post sweeps of a symmetric
smoother, w in (0, 2)), so CG remains valid; the plain weights stay the
fallback (gs_solver_set_fallback).

Usage on a GPU box:
  python learned.py label --kind sgs --first 0 --count 100 --size 64 --out labels_sgs.jsonl
  python learned.py fit --kind sgs --labels labels_sgs.jsonl --out learned_weights.json
  python learned.py evaluate --kind sgs --size 160 --weights learned_weights.json
(label appends one JSON line per seed, so a long label run can be split
over several invocations; `train` runs label and fit in one go.)
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import gpusolver as gs  # noqa: E402

MODEL_PATH = HERE / "learned_weights.json"
# Training / evaluation problem sizes: varcoef3d grid edge, elasticity
# vertices per edge.
DEFAULT_SIZE = {"sgs": 64, "bgs": 48, "hbgs": 48, "hsgs": 128}
HIDDEN = 64

# Search grids per smoother kind; the plain weight is the reference's
# (MG-GS w = 1, examples/multigridPoisson/mainAMGtest.cc:697-698; block GS
# w = 0.7, examples/elasticity/cantileverBeam.cc:572-573).
# coarse_sweeps: the sweep counts tried on levels below the finest (0 turns a
# level's smoothing off: its restricted residual goes straight down and the
# prolongated correction straight back, which keeps the V-cycle symmetric;
# for block GS this is the reference's own 0-sweep smoother).
# "hbgs" is the reference's hybrid block smoother
# (MultigridHybridBlockPreconditioner, src/Preconditioners/
# MultigridHybridBlockPreconditioner.cc:9-31, used by
# examples/elasticity/cantileverBeam.cc:589): block Gauss-Seidel on the
# finest level and block Jacobi below it, each level's (w, sweeps) learned;
# its `reference` kind is the all-block-Gauss-Seidel method the benchmark's
# plain mode runs. "hsgs" is the same pairing for scalar levels, built the
# way that preconditioner builds its MultigridSmoother (a fine and a coarse
# smoother type, include/GPUSolver/Smoothers/MultigridSmoother.h:59-65):
# symmetric Gauss-Seidel on the finest level, damped Jacobi below.
# coarse_omegas: the damping grid of the levels below the finest when it
# differs from the finest level's.
KINDS = {
    "sgs": dict(smoother=gs.SGS, plain_omega=1.0, omegas=[0.7, 0.85, 1.0, 1.15, 1.3, 1.45, 1.6],
                sweeps=[1, 2], coarse_sweeps=[0, 1, 2], block=1, coarse_steps=2),
    "bgs": dict(smoother=gs.BLOCK_GS, plain_omega=0.7,
                omegas=[0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2], sweeps=[1, 2],
                coarse_sweeps=[0, 1, 2], block=3, coarse_steps=0),
    "hbgs": dict(smoother=gs.BLOCK_GS, coarse_smoother=gs.BLOCK_JACOBI, plain_omega=0.7,
                 omegas=[0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2], sweeps=[1, 2],
                 coarse_sweeps=[0, 1, 2], block=3, coarse_steps=0, reference="bgs"),
    "hsgs": dict(smoother=gs.SGS, coarse_smoother=gs.DAMPED_JACOBI, plain_omega=1.0,
                 omegas=[0.7, 0.85, 1.0, 1.15, 1.3, 1.45, 1.6],
                 coarse_omegas=[0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1], sweeps=[1, 2],
                 coarse_sweeps=[0, 1, 2], block=1, coarse_steps=2, reference="sgs"),
}


def solver_options(kind: str) -> dict:
    """The smoother options of a kind's solver (the finest level's smoother
    and, for the hybrid, the coarse levels' one)."""
    return solver_options_of(KINDS[kind])


def config_kind(config: str, path: Path = MODEL_PATH):
    """The trained kind a benchmark configuration's learned mode uses: the
    hybrid smoother (Jacobi below the finest level) when its model is
    stored, else the all-Gauss-Seidel one -- block smoothers for the
    elasticity configurations, scalar ones for the others; None without a
    stored model."""
    if not Path(path).exists():
        return None
    d = json.loads(Path(path).read_text())
    for kind in (("hbgs", "bgs") if "elasticity" in config else ("hsgs", "sgs")):
        if kind in d:
            return kind
    return None


def training_problem(kind: str, seed: int, size: int):
    """Seeded training / evaluation problems, at a reduced size with the
    correlation length scaled to keep the same physical field."""
    if KINDS[kind]["block"] == 1:
        A = gs.varcoef3d(size, seed=seed, sigma=8.0 * size / 160.0, scale=2.0)
        b = gs.uniform_pm1(seed + 1_000_000, A.rows)
        H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                              min_dofs=1000)
        return A, b, H
    A, b, _ = gs.elasticity(size, young=1.0, poisson=0.3, hetero_sigma=6.0 * size / 96.0,
                            hetero_scale=1.5, seed=seed)
    min_dofs = max(300, int(5000 * (size / 96.0) ** 3))
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=min_dofs)
    return A, b, H


def normalize(F: np.ndarray) -> np.ndarray:
    F = np.asarray(F, np.float64)
    return np.stack([np.log2(np.maximum(F[:, 0], 1.0)) / 6.0,
                     np.log(np.maximum(F[:, 1], 1e-3)),
                     F[:, 2],
                     F[:, 3] / 4.0,
                     np.log2(np.maximum(F[:, 4], 1.0)) / 4.0], axis=1)


def encode_targets(omega, sweeps) -> np.ndarray:
    return np.stack([(np.asarray(omega) - 1.0) / 0.5, (np.asarray(sweeps) - 1.5) / 0.5], axis=1)


def decode(Y: np.ndarray, min_sweeps: int = 1):
    """(w, sweeps) per level; the finest level keeps at least one sweep."""
    omega = np.clip(1.0 + 0.5 * Y[:, 0], 0.2, 1.9)
    sweeps = np.clip(np.rint(1.5 + 0.5 * Y[:, 1]), min_sweeps, 3).astype(np.int32)
    if sweeps.size:
        sweeps[0] = max(int(sweeps[0]), 1)
    return omega, sweeps


def level_nnz(H) -> np.ndarray:
    return np.array([H.get("A", k).nnz for k in range(H.levels)], np.float64)


def work_per_iteration(nnz: np.ndarray, sweeps) -> float:
    """Modelled bytes-proportional work of one PCG iteration: the CG SpMV
    plus, per smoothed level, 2 x sweeps forward/backward passes for pre and
    post smoothing and one residual pass."""
    w = nnz[0]
    for k in range(len(nnz) - 1):
        w += (4.0 * sweeps[k] + 1.0) * nnz[k]
    return w


def solver_options_of(spec) -> dict:
    """solver_options for a kind's spec."""
    o = dict(smoother=spec["smoother"])
    if "coarse_smoother" in spec:
        o["coarse_smoother"] = spec["coarse_smoother"]
    return o


def make_solver(H, spec, omega, sweeps):
    """The mixed-precision solver of the headline configuration: FP32
    V-cycle, FP64 CG, FP64 direct solve at the first level within
    max_direct unknowns."""
    s = gs.Solver(gs.PREC_MIXED, coarse_direct=1, direct_truncate=1,
                  coarse_steps=spec["coarse_steps"], omega=spec["plain_omega"], graph_iters=8,
                  cg_block3=1 if spec["block"] == 3 else 0, **solver_options_of(spec))
    s.setup(H)
    s.set_level_weights(omega=omega, sweeps=sweeps)
    return s


# A candidate replaces the current setting only when it is at least 2%
# faster, so run-to-run timing noise does not pick the label.
ACCEPT = 0.98


def sweep_labels(H, b, spec, maxit=400, passes=3):
    """Coordinate search over (w, sweeps) per smoothed level, repeated until
    a pass changes nothing (at most `passes`): a level's best w depends on
    the sweeps chosen below it (with level 1 unsmoothed, level 0 wants a
    larger w than with level 1 smoothed)."""
    L = H.levels
    omega = [spec["plain_omega"]] * L
    sweeps = [1] * L
    s = make_solver(H, spec, omega, sweeps)
    used = len(s.describe()["levels"])  # levels in use; the last is solved directly

    def cost(om, sw):
        s.set_level_weights(omega=om, sweeps=sw)
        best_ms = np.inf
        for _ in range(2):
            x, info = s.solve(b, tol=1e-8, maxit=maxit)
            if not info.converged or info.breakdown:
                return np.inf, info.iterations
            best_ms = min(best_ms, info.solve_ms)
        return best_ms, info.iterations

    best, it0 = cost(omega, sweeps)
    plain_cost = best
    for _ in range(passes):
        changed = False
        for k in range(used - 1):  # the last level in use is solved directly
            omegas = spec["omegas"] if k == 0 else spec.get("coarse_omegas", spec["omegas"])
            for om in omegas:
                for sw in (spec["sweeps"] if k == 0 else spec["coarse_sweeps"]):
                    if om == omega[k] and sw == sweeps[k]:
                        continue
                    if sw == 0 and om != spec["plain_omega"]:
                        continue  # an unsmoothed level has no damping
                    o2, s2 = list(omega), list(sweeps)
                    o2[k], s2[k] = om, sw
                    c, _ = cost(o2, s2)
                    if c < ACCEPT * best:
                        best, omega, sweeps = c, o2, s2
                        changed = True
        if not changed:
            break
    return np.array(omega[:used - 1]), np.array(sweeps[:used - 1]), plain_cost, best


def load_model(kind: str, path: Path = MODEL_PATH):
    d = json.loads(Path(path).read_text())
    m = d[kind]
    net = gs.MLP(5, m["hidden"], 2, seed=0)
    net.set_params(np.array(m["params"], np.float64))
    return net


def feature_range(rows):
    """Per-feature (min, max) of the labelled levels a model is fit on."""
    F = np.concatenate([np.array(r["features"], np.float64) for r in rows])
    return F.min(axis=0).tolist(), F.max(axis=0).tolist()


def predict(H, kind: str, path: Path = MODEL_PATH):
    """Per-level (omega, sweeps) for hierarchy H from the trained model,
    inference on the GPU. A model that stores its training range sees each
    feature clipped to it: outside the labelled problems (e.g. a uniform
    coefficient, contrast 0, for a model trained on heterogeneous ones) the
    network answers as at the nearest training edge instead of
    extrapolating."""
    spec = KINDS[kind]
    net = load_model(kind, path)
    F = np.asarray(H.features(spec["block"]), np.float64)
    m = json.loads(Path(path).read_text())[kind]
    if "feature_min" in m:
        F = np.clip(F, np.asarray(m["feature_min"]), np.asarray(m["feature_max"]))
    omega, sweeps = decode(net.forward(normalize(F)), min(spec["coarse_sweeps"]))
    return omega, sweeps


def weights_for(config: str, H):
    kind = config_kind(config)
    if kind is None:
        return None
    return predict(H, kind)


def label(kind: str, seeds, size: int, out: Path) -> None:
    """Appends one JSON line per seed: raw level features (smoothed levels),
    the searched (w, sweeps) labels and the plain / best solve times."""
    spec = KINDS[kind]
    t0 = time.time()
    with open(out, "a") as f:
        for seed in seeds:
            A, b, H = training_problem(kind, seed, size)
            om, sw, pc, bc = sweep_labels(H, b, spec)
            F = H.features(spec["block"])[:len(om)]
            f.write(json.dumps({"kind": kind, "seed": int(seed), "size": size,
                                "levels": H.levels, "features": F.tolist(),
                                "omega": om.tolist(), "sweeps": sw.tolist(),
                                "plain_ms": pc, "best_ms": bc}) + "\n")
            f.flush()
            print(f"[label {kind}] seed {seed}: levels {H.levels} omega "
                  f"{np.round(om, 2).tolist()} sweeps {sw.tolist()} label time gain "
                  f"{pc / bc:.3f} ({time.time() - t0:.0f}s)", flush=True)


def read_labels(kind: str, paths):
    rows = []
    for p in paths:
        for line in Path(p).read_text().splitlines():
            if line.strip():
                r = json.loads(line)
                if r["kind"] == kind:
                    rows.append(r)
    seen = {}
    for r in rows:  # one record per seed (the last one written wins)
        seen[r["seed"]] = r
    return [seen[k] for k in sorted(seen)]


def constant_policy(rows, spec) -> dict:
    """The best fixed setting the labels support, for comparison with the
    network: per level index, the most frequent sweep count and the mean
    damping of the labels that smooth that level."""
    depth = max(len(r["omega"]) for r in rows)
    omega, sweeps = [], []
    for k in range(depth):
        sw = [r["sweeps"][k] for r in rows if k < len(r["sweeps"])]
        vals, counts = np.unique(sw, return_counts=True)
        s_k = int(vals[np.argmax(counts)])
        om = [r["omega"][k] for r in rows if k < len(r["omega"]) and r["sweeps"][k] > 0]
        omega.append(float(np.mean(om)) if (om and s_k > 0) else spec["plain_omega"])
        sweeps.append(s_k)
    return {"omega": omega, "sweeps": sweeps}


def fit(kind: str, label_paths, epochs: int, out: Path) -> dict:
    """Trains the MLP on the GPU (HIP forward / backward / Adam) on the
    labelled levels and stores it under `kind` in `out`."""
    rows = read_labels(kind, label_paths)
    if not rows:
        raise SystemExit(f"no {kind} labels in {label_paths}")
    spec = KINDS[kind]
    X = np.concatenate([normalize(np.array(r["features"])) for r in rows])
    # A level with 0 sweeps has no damping to learn: its target is the plain w.
    T = np.concatenate([encode_targets([o if sw > 0 else spec["plain_omega"]
                                        for o, sw in zip(r["omega"], r["sweeps"])], r["sweeps"])
                        for r in rows])
    net = gs.MLP(5, HIDDEN, 2, seed=1234)
    l0, _ = net.loss_grad(X, T)
    loss = net.train(X, T, epochs=epochs, lr=3e-3)
    print(f"[fit {kind}] seeds {len(rows)} samples {X.shape[0]} loss {l0:.4f} -> {loss:.4f}",
          flush=True)
    res = {"hidden": HIDDEN, "params": net.get_params().tolist(), "samples": int(X.shape[0]),
           "seeds": len(rows), "size": rows[0]["size"],
           "sizes": sorted({int(r["size"]) for r in rows}), "epochs": epochs,
           "loss_initial": l0, "loss_final": loss,
           "label_time_gain_mean": float(np.mean([r["plain_ms"] / r["best_ms"] for r in rows])),
           "constant_policy": constant_policy(rows, spec)}
    res["feature_min"], res["feature_max"] = feature_range(rows)
    doc = json.loads(out.read_text()) if out.exists() else {}
    doc[kind] = res
    out.write_text(json.dumps(doc))
    return res


def train(kind: str, seeds: int, size: int, epochs: int, out: Path) -> dict:
    lab = out.with_suffix(f".{kind}.labels.jsonl")
    if lab.exists():
        lab.unlink()
    label(kind, range(seeds), size, lab)
    return fit(kind, [lab], epochs, out)


def constant_weights(kind: str, L: int, path: Path = MODEL_PATH):
    """The stored constant policy extended to L levels (plain weights past
    the labelled depth)."""
    spec = KINDS[kind]
    c = json.loads(Path(path).read_text())[kind].get("constant_policy")
    if c is None:
        return None
    om = (list(c["omega"]) + [spec["plain_omega"]] * L)[:L]
    sw = (list(c["sweeps"]) + [1] * L)[:L]
    sw[0] = max(sw[0], 1)
    return om, sw


def evaluate(kind: str, seeds, size: int, path: Path) -> dict:
    """Held-out comparison, all in the mixed headline configuration: plain
    weights, the network's weights, and the best constant policy of the
    labels. Each entry is the better of two timed solves."""
    spec = KINDS[kind]
    rows = []
    for seed in seeds:
        A, b, H = training_problem(kind, seed, size)
        L = H.levels
        om, sw = predict(H, kind, path)
        policies = [("plain", spec, [spec["plain_omega"]] * L, [1] * L),
                    ("learned", spec, om.tolist(), sw.tolist())]
        cw = constant_weights(kind, L, path)
        if cw is not None:
            policies.append(("constant", spec, cw[0], cw[1]))
        if "reference" in spec:
            ref = KINDS[spec["reference"]]
            policies.append(("plain_reference", ref, [ref["plain_omega"]] * L, [1] * L))
        nnz = level_nnz(H)
        res = {}
        for name, sp, omk, swk in policies:
            s = make_solver(H, sp, omk, swk)
            s.solve(b, tol=1e-8, maxit=1000)
            best = None
            for _ in range(2):
                x, info = s.solve(b, tol=1e-8, maxit=1000)
                if best is None or info.solve_ms < best.solve_ms:
                    best = info
            used = len(s.describe()["levels"])
            res[name] = dict(iterations=best.iterations, ms=best.solve_ms,
                             work=(best.iterations + 1) * work_per_iteration(nnz, swk),
                             true_relres=best.true_relres, converged=best.converged,
                             omega=[round(float(v), 3) for v in omk[:used - 1]],
                             sweeps=[int(v) for v in swk[:used - 1]])
            del s
        rows.append(res)
        print(f"[eval {kind} {size}] seed {seed}: " + " | ".join(
            f"{k} it {v['iterations']} {v['ms']:.2f} ms" for k, v in res.items()), flush=True)
    agg = {k: float(np.mean([r["plain"][k] / r["learned"][k] for r in rows]))
           for k in ("iterations", "ms", "work")}
    out = {"per_seed": rows, "mean_ratio_plain_over_learned": agg}
    if all("plain_reference" in r for r in rows):
        out["mean_ratio_plain_reference_over_learned"] = {
            k: float(np.mean([r["plain_reference"][k] / r["learned"][k] for r in rows]))
            for k in ("iterations", "ms")}
    if all("constant" in r for r in rows):
        out["mean_ratio_constant_over_learned"] = {
            k: float(np.mean([r["constant"][k] / r["learned"][k] for r in rows]))
            for k in ("iterations", "ms")}
        out["mean_ratio_plain_over_constant"] = {
            k: float(np.mean([r["plain"][k] / r["constant"][k] for r in rows]))
            for k in ("iterations", "ms")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["label", "fit", "train", "evaluate"])
    ap.add_argument("--kind", choices=sorted(KINDS), default="sgs")
    ap.add_argument("--seeds", type=int, default=200)
    ap.add_argument("--first", type=int, default=-1,
                    help="first seed (label: 0, evaluate: 200, the held-out seeds)")
    ap.add_argument("--count", type=int, default=-1, help="seeds (label: 10, evaluate: 20)")
    ap.add_argument("--labels", nargs="*", default=[])
    ap.add_argument("--size", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--out", default=str(MODEL_PATH))
    ap.add_argument("--weights", default=str(MODEL_PATH))
    ap.add_argument("--json-out", default="", help="evaluate: write the full result here")
    a = ap.parse_args()
    size = a.size or DEFAULT_SIZE[a.kind]
    if a.first < 0:
        a.first = 200 if a.cmd == "evaluate" else 0
    if a.count < 0:
        a.count = 20 if a.cmd == "evaluate" else 10
    if a.cmd == "label":
        label(a.kind, range(a.first, a.first + a.count), size, Path(a.out))
    elif a.cmd == "fit":
        fit(a.kind, a.labels, a.epochs, Path(a.out))
    elif a.cmd == "train":
        train(a.kind, a.seeds, size, a.epochs, Path(a.out))
    else:
        r = evaluate(a.kind, range(a.first, a.first + a.count), size, Path(a.weights))
        if a.json_out:
            Path(a.json_out).write_text(json.dumps(r))
        print(json.dumps({k: v for k, v in r.items() if k != "per_seed"}))


if __name__ == "__main__":
    main()
