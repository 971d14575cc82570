"""Mixed precision with the CG SpMV on level 0's FP32 V-cycle operator
(`cg_vcycle_operator`): the products and sums stay in FP64, the FP64
operator replaces the recursive residual by b - A x every `cg_replace_every`
iterations and confirms every stop on convergence. The solve must still reach
the 1e-8 relative residual measured with the FP64 operator, with the
iteration count within +-2 of the oracle's mixed-precision run (which uses
the FP64 operator throughout) and the solution within 1e-6 of it.
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


@pytest.mark.parametrize("kind", ["sgs", "block_gs"])
@pytest.mark.parametrize("every", [1, 3, 8])
def test_converges_to_fp64_residual(poisson, elast, kind, every):
    A, H, b, kw, okw = case(kind, poisson, elast)
    s = gs.Solver(gs.PREC_MIXED, cg_vcycle_operator=1, cg_replace_every=every, **kw)
    s.setup(H)
    assert s.describe()["cg_vcycle_operator"] is True
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=1000, cg_double=True,
                            mg_double=False, **okw)
    assert info.converged
    assert abs(info.iterations - ito) <= 2, (info.iterations, ito)
    S = A.to_scipy()
    true = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
    # Every stop is confirmed on b - A x in FP64, so the target holds.
    assert info.true_relres < 1e-8 * (1 + 1e-9)
    assert true < 1.0001e-8
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    # The history's last entry is the replaced (true) residual.
    h = s.history()
    assert abs(np.sqrt(h[-1]) - info.true_relres) <= 1e-6 * info.true_relres


def test_graph_and_eager_loops_agree(elast):
    E, rhs, H = elast
    out = []
    for gi in (0, 8):
        s = gs.Solver(gs.PREC_MIXED, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2,
                      cg_block3=1, cg_vcycle_operator=1, graph_iters=gi)
        s.setup(H)
        out.append(s.solve(rhs, tol=1e-8, maxit=1000))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1].iterations == out[1][1].iterations


def test_maxit_stops_with_replacement(poisson):
    A, H = poisson
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2, cg_vcycle_operator=1,
                  cg_replace_every=2)
    s.setup(H)
    x, info = s.solve(gs.uniform_pm1(5, A.rows), tol=1e-14, maxit=4)
    assert not info.converged and info.iterations == 4


def test_ignored_where_it_does_not_apply(poisson):
    """FP64 mode, and a level 0 on lane-group kernels, keep the CG operator."""
    A, H = poisson
    for prec, kw in ((gs.PREC_DOUBLE, {}), (gs.PREC_MIXED, dict(group_min_scalar=1.0))):
        s = gs.Solver(prec, smoother=gs.SGS, coarse_steps=2, cg_vcycle_operator=1, **kw)
        s.setup(H)
        assert s.describe()["cg_vcycle_operator"] is False
        x, info = s.solve(gs.uniform_pm1(6, A.rows), tol=1e-8, maxit=1000)
        assert info.converged and info.true_relres <= 1.5e-8


def test_replace_every_must_be_positive(poisson):
    with pytest.raises(gs.GpuSolverError, match="cg_replace_every must be at least 1"):
        gs.Solver(gs.PREC_MIXED, cg_replace_every=0)
