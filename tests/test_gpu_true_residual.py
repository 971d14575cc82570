"""Mixed precision with the FP64 CG operator: the true residual b - A x is
recomputed in FP64 every `cg_true_residual_every` iterations (32 by default)
and before every stop on convergence, and replaces the recursive residual.
The solve must match the oracle's mixed-precision run (FP64 CG without
replacement) to +-2 iterations and 1e-6 in the solution, every confirmed
stop must hold on b - A x in FP64, and a captured graph, which carries the
replacement kernels in its last iteration only, must give the same bits as
the eager loop, which carries them in every iteration.
"""
import numpy as np
import pytest

import gpusolver as gs
import oracle_api as orc

pytestmark = pytest.mark.gpu


def hier_view(H):
    L = H.levels
    A = [H.get("A", k) for k in range(L)]
    P = [H.get("P", k) for k in range(L - 1)]
    R = [H.get("R", k) for k in range(L - 1)]
    return orc.HierView([(a.ptr, a.col, a.val) for a in A], [(p.ptr, p.col, p.val) for p in P],
                        [(r.ptr, r.col, r.val) for r in R])


@pytest.fixture(scope="module")
def poisson():
    A = gs.poisson2d(200, 180)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    return A, H


@pytest.fixture(scope="module")
def elast():
    E, rhs, ns = gs.elasticity(12)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=500, nullspace=ns)
    return E, rhs, H


def case(kind, poisson, elast):
    if kind == "sgs":
        A, H = poisson
        return A, H, gs.uniform_pm1(21, A.rows), dict(smoother=gs.SGS, coarse_steps=2), \
            dict(kinds=gs.SGS)
    E, rhs, H = elast
    return E, H, rhs, dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1), \
        dict(kinds=gs.BLOCK_GS, omegas=0.7)


def test_period_per_mode(poisson):
    A, H = poisson
    expect = ((gs.PREC_MIXED, {}, 32), (gs.PREC_MIXED, dict(cg_true_residual_every=0), 0),
              (gs.PREC_DOUBLE, {}, 0), (gs.PREC_SINGLE, {}, 0),
              (gs.PREC_MIXED, dict(cg_vcycle_operator=1), 8))
    for prec, kw, every in expect:
        s = gs.Solver(prec, smoother=gs.SGS, coarse_steps=2, **kw)
        s.setup(H)
        assert s.describe()["replace_every"] == every, (prec, kw)


@pytest.mark.parametrize("kind", ["sgs", "block_gs"])
@pytest.mark.parametrize("every", [0, 1, 5, 32])
@pytest.mark.parametrize("gi", [0, 8])
def test_parity_with_true_residual(poisson, elast, kind, every, gi):
    A, H, b, kw, okw = case(kind, poisson, elast)
    s = gs.Solver(gs.PREC_MIXED, cg_true_residual_every=every, graph_iters=gi, **kw)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=1000, cg_double=True,
                            mg_double=False, **okw)
    assert info.converged
    assert abs(info.iterations - ito) <= 2, (info.iterations, ito)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    S = A.to_scipy()
    true = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
    if every > 0:
        # The stop was confirmed on b - A x in FP64.
        assert info.true_relres < 1e-8 * (1 + 1e-9)
        assert true < 1.0001e-8
        h = s.history()
        assert abs(np.sqrt(h[-1]) - info.true_relres) <= 1e-6 * info.true_relres
    else:
        assert info.true_relres <= 1.5e-8 and true <= 1.5e-8


def test_failed_check_resumes_the_iteration(elast):
    """A tolerance below what FP64 attains: the recursive residual passes it,
    the replaced one does not, and the iteration goes on from the replaced
    residual. Without the check the same solve stops early; with it the
    graph (check resolved in its last iteration, the skipped iterations'
    kernels idle) and the eager loop give the same bits."""
    E, rhs, H = elast
    kw = dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1)
    tol, maxit = 1e-15, 80
    s0 = gs.Solver(gs.PREC_MIXED, cg_true_residual_every=0, **kw)
    s0.setup(H)
    _, info0 = s0.solve(rhs, tol=tol, maxit=maxit)
    assert info0.converged and info0.iterations < maxit  # the recursive residual passes
    out = []
    for gi in (0, 8):
        s = gs.Solver(gs.PREC_MIXED, cg_true_residual_every=1000, graph_iters=gi, **kw)
        s.setup(H)
        x, info = s.solve(rhs, tol=tol, maxit=maxit)
        assert not info.converged and info.iterations == maxit
        assert info.loop_iterations == maxit  # every iteration ran, graph or not
        h = s.history()
        # Entries below the tolerance were replaced by the true residual.
        assert h.min() >= tol * tol
        assert np.all(np.isfinite(x))
        out.append((x, info, h))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    assert out[0][1].true_relres == out[1][1].true_relres


def test_period_must_be_non_negative():
    with pytest.raises(gs.GpuSolverError, match="cg_true_residual_every must be non-negative"):
        gs.Solver(gs.PREC_MIXED, cg_true_residual_every=-1)
