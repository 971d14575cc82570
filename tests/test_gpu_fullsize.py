"""Parity at the benchmark's full sizes (3D Poisson 128^3, elasticity 64^3)
against full FP64 oracle solves on the host: iteration count within +-2,
solution within 1e-6 relative (2-norm), FP64 true relative residual <=
1.5e-8. The mixed-precision mode with the direct coarse solve is checked
against the oracle running the same method (FP32 V-cycle, FP64 CG, same
truncated hierarchy). GPU."""
import numpy as np
import pytest

import gpusolver as gs
import oracle_api as orc

pytestmark = pytest.mark.gpu


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


def test_elasticity64_mixed_learned_shape(elasticity64):
    """The headline configuration on a full-size elasticity problem: mixed
    precision, direct solve at the first level within max_direct, level 1
    unsmoothed (the learned model's choice), against the oracle running the
    same method."""
    A, H, b = elasticity64
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.BLOCK_GS, coarse_steps=0, omega=0.7, cg_block3=1,
                  coarse_direct=1, direct_truncate=1)
    s.setup(H)
    used = len(s.describe()["levels"])
    sweeps = [1, 0] + [1] * (H.levels - 2)
    s.set_level_weights(omega=[0.7] * H.levels, sweeps=sweeps)
    x, info = s.solve(b, tol=1e-8, maxit=2000)
    xo, ito, _, _ = orc.pcg(hier_view(H, used), b, tol=1e-8, maxit=2000, mg_double=False,
                            kinds=gs.BLOCK_GS, omegas=0.7, pres=sweeps[:used],
                            posts=sweeps[:used], coarse_steps=0, direct_coarse=True)
    check(x, info, xo, ito, A, b)
