"""The graph tail: near convergence the solve loop replays a one-iteration
graph as many times as the last chunk's rate says are left, instead of a
whole chunk of `graph_iters` iterations whose iterations past convergence
are no-op launches. It must give the chunk graph's results bit for bit
(solution, history, iteration counts, true residual) in every precision,
with and without the periodic FP64 residual replacement, and with a
convergence check that fails and resumes inside the tail; and it must not
launch more iterations than whole chunks do.
"""
import numpy as np
import pytest

import gpusolver as gs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def poisson():
    A = gs.poisson2d(200, 180)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    return A, H, gs.uniform_pm1(31, A.rows)


@pytest.fixture(scope="module")
def elast():
    E, rhs, ns = gs.elasticity(12)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=500, nullspace=ns)
    return E, H, rhs


def run(H, b, prec, tail, tol=1e-8, maxit=1000, **kw):
    s = gs.Solver(prec, graph_iters=8, graph_tail=tail, **kw)
    s.setup(H)
    x, info = s.solve(b, tol=tol, maxit=maxit)
    return x, info, s.history(), int(s.result("launched_iterations"))


def same(a, b):
    xa, ia, ha, _ = a
    xb, ib, hb, _ = b
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(ha, hb)
    assert ia.iterations == ib.iterations and ia.loop_iterations == ib.loop_iterations
    assert ia.converged == ib.converged
    assert ia.true_relres == ib.true_relres


CASES = [
    ("sgs", gs.PREC_DOUBLE, {}),
    ("sgs", gs.PREC_MIXED, {}),  # FP64 residual replacement every 32 iterations
    ("sgs", gs.PREC_MIXED, dict(cg_true_residual_every=0)),
    ("sgs", gs.PREC_MIXED, dict(cg_true_residual_every=5)),  # pair in every iteration
    ("sgs", gs.PREC_MIXED, dict(cg_true_residual_every=12)),  # tail off: chunks only
    ("sgs", gs.PREC_SINGLE, {}),
    ("block_gs", gs.PREC_DOUBLE, {}),
    ("block_gs", gs.PREC_MIXED, {}),
    ("block_gs", gs.PREC_MIXED, dict(coarse_direct=1)),
]


@pytest.mark.parametrize("kind,prec,extra", CASES)
def test_tail_gives_the_chunk_bits(poisson, elast, kind, prec, extra):
    if kind == "sgs":
        A, H, b = poisson
        kw = dict(smoother=gs.SGS, coarse_steps=2)
    else:
        A, H, b = elast
        kw = dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1)
    kw.update(extra)
    tol = 1e-5 if prec == gs.PREC_SINGLE else 1e-8
    chunk = run(H, b, prec, 0, tol=tol, **kw)
    tail = run(H, b, prec, 1, tol=tol, **kw)
    assert chunk[1].converged
    same(chunk, tail)
    it = chunk[1].loop_iterations
    # Whole chunks launch a multiple of 8 iterations; the tail launches no
    # more, and at least the iterations that ran.
    assert chunk[3] % 8 == 0 and chunk[3] >= it
    assert it <= tail[3] <= chunk[3]
    if extra.get("cg_true_residual_every") == 12:
        assert tail[3] == chunk[3]


def test_tail_launches_fewer_iterations(poisson):
    """On the 2D Poisson problem the plain FP64 solve stops a few iterations
    into its last chunk; with the tail it launches at most one iteration
    beyond the ones that ran, where whole chunks launch up to seven."""
    A, H, b = poisson
    kw = dict(smoother=gs.SGS, coarse_steps=2)
    chunk = run(H, b, gs.PREC_DOUBLE, 0, **kw)
    tail = run(H, b, gs.PREC_DOUBLE, 1, **kw)
    same(chunk, tail)
    it = tail[1].loop_iterations
    assert it % 8 != 0  # the last chunk is partly idle without the tail
    assert tail[3] <= it + 1 < chunk[3]


def test_failed_check_resumes_inside_the_tail(elast):
    """A tolerance below what FP64 attains, with a replacement period that
    is a multiple of the chunk length (so the tail is on): the recursive
    residual passes the tolerance, the replaced one does not, and the
    iteration resumes; chunks, the tail and the eager loop agree bit for
    bit."""
    E, H, rhs = elast
    kw = dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1,
              cg_true_residual_every=1024)
    tol, maxit = 1e-15, 80
    out = [run(H, rhs, gs.PREC_MIXED, tail, tol=tol, maxit=maxit, **kw) for tail in (0, 1)]
    s = gs.Solver(gs.PREC_MIXED, graph_iters=0, **kw)
    s.setup(H)
    x, info = s.solve(rhs, tol=tol, maxit=maxit)
    out.append((x, info, s.history(), int(s.result("launched_iterations"))))
    for o in out:
        assert not o[1].converged and o[1].loop_iterations == maxit
    same(out[0], out[1])
    same(out[0], out[2])


def test_repeated_short_solve_launches_only_its_iterations():
    """A solve shorter than a chunk (the 2D Poisson problem of 65 536
    unknowns with the direct coarse level: 6 iterations) launches a whole
    chunk the first time; the next solve starts from the last solve's loop
    count and launches only the iterations that run, with the same bits."""
    A = gs.poisson2d(256, 256)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    b = gs.uniform_pm1(1, A.rows)
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2, coarse_direct=1,
                  direct_truncate=1, graph_iters=8)
    s.setup(H)
    x1, i1 = s.solve(b, tol=1e-8)
    l1 = int(s.result("launched_iterations"))
    x2, i2 = s.solve(b, tol=1e-8)
    l2 = int(s.result("launched_iterations"))
    assert i1.converged and i2.converged
    assert i1.loop_iterations < 8 and l1 == 8
    assert l2 == i2.loop_iterations == i1.loop_iterations
    np.testing.assert_array_equal(x1, x2)
    assert i1.true_relres == i2.true_relres
    # New weights drop the graphs and the remembered count.
    s.set_level_weights(omega=[1.0, 1.0], sweeps=[1, 1])
    s.solve(b, tol=1e-8)
    assert int(s.result("launched_iterations")) == 8
