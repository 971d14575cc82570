"""Parity at the benchmark's full sizes (3D Poisson 128^3, elasticity 64^3,
variable-coefficient Poisson 160^3, heterogeneous elasticity 96^3) against
full FP64 oracle solves on the host: iteration count within +-2, solution
within 1e-6 relative (2-norm), FP64 true relative residual <= 1.5e-8. The
mixed-precision modes with the direct coarse solve (and the learned
per-level weights) are checked against the oracle running the same method
(FP32 V-cycle, FP64 CG, same truncated hierarchy, same weights). GPU.

The 96^3 elasticity oracle solves its 13 047-unknown direct level with an
inverse from LAPACK (numpy.linalg.inv), which the oracle's own scalar
Cholesky inverse is checked against on smaller matrices; its plain solve
takes about 40 s on 16 host threads."""
import numpy as np
import pytest

import gpusolver as gs
import learned
import oracle_api as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def say(msg):
    """Names the case before each GPU step (flushed, so a failing run still
    shows it)."""
    print(f"[fullsize] {msg}", flush=True)


def hier_view(H, levels=None):
    L = H.levels if levels is None else levels
    return orc.HierView([(a.ptr, a.col, a.val) for a in (H.get("A", k) for k in range(L))],
                        [(p.ptr, p.col, p.val) for p in (H.get("P", k) for k in range(L - 1))],
                        [(r.ptr, r.col, r.val) for r in (H.get("R", k) for k in range(L - 1))])


def check(x, info, xo, ito, A, b):
    assert info.converged
    assert abs(info.iterations - ito) <= 2, (info.iterations, ito)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert err <= 1e-6, err
    true = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
    assert true <= 1.5e-8, true
    assert abs(info.true_relres - true) <= 1e-3 * true + 1e-15


@pytest.fixture(scope="module")
def poisson128():
    A = gs.poisson3d(128)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    return A, H, gs.uniform_pm1(1, A.rows)


@pytest.fixture(scope="module")
def elasticity64():
    A, b, _ = gs.elasticity(64, young=1.0, poisson=0.3, seed=1)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=5000)
    return A, H, b


def test_poisson128_plain_fp64(poisson128):
    A, H, b = poisson128
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=1000, kinds=gs.SGS, coarse_steps=2)
    check(x, info, xo, ito, A, b)


def test_poisson128_mixed_direct(poisson128):
    A, H, b = poisson128
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2, coarse_direct=1,
                  direct_truncate=1)
    s.setup(H)
    used = len(s.describe()["levels"])
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H, used), b, tol=1e-8, maxit=1000, mg_double=False,
                            kinds=gs.SGS, coarse_steps=2, direct_coarse=True)
    check(x, info, xo, ito, A, b)


def test_elasticity64_plain_fp64(elasticity64):
    A, H, b = elasticity64
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, coarse_steps=0, omega=0.7, cg_block3=1)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=2000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=2000, kinds=gs.BLOCK_GS,
                            omegas=0.7, coarse_steps=0)
    check(x, info, xo, ito, A, b)


@pytest.mark.parametrize("coarse,level1", [(gs.BLOCK_GS, 0), (gs.BLOCK_JACOBI, 1)])
def test_elasticity64_mixed_learned_shape(elasticity64, coarse, level1):
    """The learned configurations' shape on a full-size elasticity problem:
    mixed precision, direct solve at the first level within max_direct, and
    level 1 either block Gauss-Seidel left unsmoothed or block Jacobi with
    one sweep (the hybrid block smoother), against the oracle running the
    same method."""
    A, H, b = elasticity64
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.BLOCK_GS, coarse_smoother=coarse, coarse_steps=0,
                  omega=0.7, cg_block3=1, coarse_direct=1, direct_truncate=1)
    s.setup(H)
    used = len(s.describe()["levels"])
    sweeps = [1, level1] + [1] * (H.levels - 2)
    omega = [0.96, 0.9] + [0.7] * (H.levels - 2)
    s.set_level_weights(omega=omega, sweeps=sweeps)
    say(f"elasticity_64 mixed, level 1 {'block Jacobi' if level1 else 'unsmoothed'}")
    x, info = s.solve(b, tol=1e-8, maxit=2000)
    kinds = [gs.BLOCK_GS] + [coarse] * (used - 1)
    xo, ito, _, _ = orc.pcg(hier_view(H, used), b, tol=1e-8, maxit=2000, mg_double=False,
                            kinds=kinds, omegas=omega[:used], pres=sweeps[:used],
                            posts=sweeps[:used], coarse_steps=0, direct_coarse=True)
    check(x, info, xo, ito, A, b)


@pytest.fixture(scope="module")
def varcoef160():
    """-div(k grad u) = f on 160^3 cells, k = exp(2 g), g a smoothed Gaussian
    field (sigma 8 cells, unit variance), seed 200 (the first held-out
    seed); f uniform in [-1, 1]."""
    A = gs.varcoef3d(160, seed=200, sigma=8.0, scale=2.0)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    return A, H, gs.uniform_pm1(1, A.rows)


def test_varcoef160_plain_fp64(varcoef160):
    A, H, b = varcoef160
    say(f"varcoef3d_160 plain FP64: n={A.rows} levels={H.levels}")
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=1000, kinds=gs.SGS, coarse_steps=2)
    check(x, info, xo, ito, A, b)


def test_varcoef160_mixed_learned(varcoef160):
    """The configuration's own method: FP32 V-cycle in FP64 CG, FP64 direct
    solve at the first level within max_direct, the smoother family of the
    stored model (SGS on level 0, damped Jacobi below for the hybrid) with
    its learned per-level weights; the oracle runs the same method and
    weights."""
    A, H, b = varcoef160
    kind = learned.config_kind("varcoef3d_160")
    assert kind is not None, "no trained scalar model"
    opts = learned.solver_options(kind)
    s = gs.Solver(gs.PREC_MIXED, coarse_steps=2, coarse_direct=1, direct_truncate=1, **opts)
    s.setup(H)
    used = len(s.describe()["levels"])
    lw = learned.weights_for("varcoef3d_160", H)
    omega, sweeps = np.asarray(lw[0])[:used], np.asarray(lw[1])[:used]
    s.set_level_weights(omega=omega, sweeps=sweeps)
    say(f"varcoef3d_160 mixed+learned ({kind}): levels used {used} omega "
        f"{np.round(omega, 3).tolist()} sweeps {sweeps.tolist()}")
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    kinds = [opts["smoother"]] + [opts.get("coarse_smoother", opts["smoother"])] * (used - 1)
    xo, ito, _, _ = orc.pcg(hier_view(H, used), b, tol=1e-8, maxit=1000, mg_double=False,
                            kinds=kinds, omegas=omega, pres=sweeps, posts=sweeps, coarse_steps=2,
                            direct_coarse=True)
    check(x, info, xo, ito, A, b)


@pytest.fixture(scope="module")
def hetero96():
    """Heterogeneous elasticity on 96^3 vertices (2 626 560 DOF), E =
    exp(1.5 g), g smoothed with sigma 6 cells; the benchmark's workload."""
    A, b, _ = gs.elasticity(96, young=1.0, poisson=0.3, hetero_sigma=6.0, hetero_scale=1.5,
                            seed=1)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=5000)
    return A, H, b


def test_hetero96_plain_fp64(hetero96):
    """The reference's method in FP64: block GS w = 0.7, 1/1 sweeps, the
    coarsest level not smoothed (examples/elasticity/cantileverBeam.cc:
    572-573)."""
    A, H, b = hetero96
    say(f"hetero_elasticity_96 plain FP64: n={A.rows} levels={H.levels}")
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, coarse_steps=0, omega=0.7, cg_block3=1)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=2000)
    del s
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=2000, kinds=gs.BLOCK_GS,
                            omegas=0.7, coarse_steps=0)
    check(x, info, xo, ito, A, b)


def test_hetero96_mixed_learned(hetero96):
    """The headline mode: FP32 V-cycle in FP64 CG, FP64 direct solve at the
    first level within max_direct (13 047 unknowns), the smoother family of
    the stored model (the hybrid block smoother: block GS on level 0, block
    Jacobi below) with its learned per-level weights; the oracle runs the
    same method with the same weights."""
    A, H, b = hetero96
    kind = learned.config_kind("hetero_elasticity_96")
    assert kind is not None, "no trained elasticity model"
    opts = learned.solver_options(kind)
    s = gs.Solver(gs.PREC_MIXED, coarse_steps=0, omega=0.7, cg_block3=1, coarse_direct=1,
                  direct_truncate=1, **opts)
    s.setup(H)
    used = len(s.describe()["levels"])
    lw = learned.weights_for("hetero_elasticity_96", H)
    omega, sweeps = np.asarray(lw[0])[:used], np.asarray(lw[1])[:used]
    s.set_level_weights(omega=omega, sweeps=sweeps)
    say(f"hetero_elasticity_96 mixed+learned ({kind}): levels used {used} omega "
        f"{np.round(omega, 3).tolist()} sweeps {sweeps.tolist()}")
    x, info = s.solve(b, tol=1e-8, maxit=2000)
    del s
    Ac = H.get("A", used - 1).to_scipy().toarray()
    inv = np.linalg.inv(Ac)
    del Ac
    kinds = [opts["smoother"]] + [opts.get("coarse_smoother", opts["smoother"])] * (used - 1)
    xo, ito, _, _ = orc.pcg(hier_view(H, used), b, tol=1e-8, maxit=2000, mg_double=False,
                            kinds=kinds, omegas=omega, pres=sweeps, posts=sweeps,
                            coarse_steps=0, direct_coarse=True, coarse_inv=inv)
    check(x, info, xo, ito, A, b)


def test_hetero96_device_hierarchy(hetero96):
    """The benchmark's hierarchy built with the GPU-side products equals the
    host-built one bit for bit (every A, P, R and aggregate map)."""
    import time
    A, H, b = hetero96
    say("hetero_elasticity_96 device-built hierarchy")
    t0 = time.perf_counter()
    Hd = gs.Hierarchy.smoothed_aggregation(A, block_size=3, eps_strong=0.1, max_levels=10,
                                           min_dofs=5000, device=True)
    say(f"device-built hierarchy in {time.perf_counter() - t0:.2f} s")
    assert Hd.levels == H.levels
    for w, L in (("A", H.levels), ("P", H.levels - 1), ("R", H.levels - 1)):
        for k in range(L):
            x, y = Hd.get(w, k), H.get(w, k)
            np.testing.assert_array_equal(x.ptr, y.ptr)
            np.testing.assert_array_equal(x.col, y.col)
            np.testing.assert_array_equal(x.val.view(np.uint64), y.val.view(np.uint64))
