"""The single-workgroup V-cycle tail (k_vcycle_tail, option tail_rows): the
small scalar Gauss-Seidel levels below the finest, down to the coarsest,
run their part of the V-cycle (sweeps, residual, transfers, coarse sweeps)
in one launch. It must give the bits of the launch-per-colour path
(tail_rows = 0) and of the oracle (src/Smoothers/MultigridSmoother.cc:
337-412), for every precision, sweep count, omega, coarse-sweep count, the
reference's aliased coarse correction, per-level weights that leave a tail
level unsmoothed, several V-cycles per application, and whole PCG solves.
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
def p2():
    A = gs.poisson2d(200, 180)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=100)
    return A, H


@pytest.fixture(scope="module")
def v3():
    A = gs.varcoef3d(24, seed=3, sigma=3.0, scale=2.0)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=100)
    return A, H


def pair(prec, H, **kw):
    s = gs.Solver(prec, **kw)
    s.setup(H)
    t = gs.Solver(prec, tail_rows=0, **kw)
    t.setup(H)
    d = s.describe()
    assert d["tail_level"] >= 1, d["tail_level"]
    assert t.describe()["tail_level"] == -1
    return s, t


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED, gs.PREC_SINGLE])
@pytest.mark.parametrize("ref", [0, 1])
@pytest.mark.parametrize("sweeps,coarse", [(1, 2), (2, 1), (1, 1), (3, 3)])
def test_tail_matches_launches_and_oracle(p2, prec, ref, sweeps, coarse):
    A, H = p2
    kw = dict(smoother=gs.SGS, omega=1.2, pre=sweeps, post=sweeps, coarse_steps=coarse,
              reference_vcycle=ref)
    s, t = pair(prec, H, **kw)
    r = gs.uniform_pm1(11, A.rows)
    z = s.vcycle(r)
    np.testing.assert_array_equal(z, t.vcycle(r))
    zo = orc.vcycle(hier_view(H), r, mg_double=prec == gs.PREC_DOUBLE, kinds=gs.SGS, omegas=1.2,
                    pres=sweeps, posts=sweeps, coarse_steps=coarse, reference_vcycle=bool(ref))
    np.testing.assert_array_equal(z, zo)


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_tail_with_unsmoothed_level_and_two_vcycles(v3, prec):
    """Per-level weights (one tail level without sweeps, its own omegas) and
    two V-cycles per application (the second from a nonzero fine x; the
    tail still starts from zero)."""
    A, H = v3
    L = H.levels
    assert L >= 4
    omega = [1.1, 0.9, 1.3, 1.05, 1.2, 1.0, 1.0][:L]
    sweeps = [1, 1, 0, 2, 1, 1, 1][:L]
    out = []
    for tail_rows in (4096, 0):
        s = gs.Solver(prec, smoother=gs.SGS, coarse_steps=2, vcycles=2, tail_rows=tail_rows)
        s.setup(H)
        s.set_level_weights(omega=omega, sweeps=sweeps)
        assert (s.describe()["tail_level"] >= 1) == (tail_rows > 0)
        out.append(s.vcycle(gs.uniform_pm1(4, A.rows)))
    np.testing.assert_array_equal(out[0], out[1])
    zo = orc.vcycle(hier_view(H), gs.uniform_pm1(4, A.rows), mg_double=prec == gs.PREC_DOUBLE,
                    kinds=gs.SGS, omegas=omega, pres=sweeps, posts=sweeps, coarse_steps=2,
                    vcycles=2)
    np.testing.assert_array_equal(out[0], zo)


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
@pytest.mark.parametrize("gi", [0, 8])
def test_tail_pcg_same_bits(v3, prec, gi):
    """Whole solves, captured graph or eager loop: identical iterates, so
    identical iteration counts and solutions."""
    A, H = v3
    b = gs.uniform_pm1(9, A.rows)
    s, t = pair(prec, H, smoother=gs.SGS, coarse_steps=2, graph_iters=gi)
    x1, i1 = s.solve(b, tol=1e-8, maxit=300)
    x2, i2 = t.solve(b, tol=1e-8, maxit=300)
    assert i1.converged and i1.iterations == i2.iterations
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(s.history(), t.history())


def test_tail_limit_and_exclusions(p2):
    """The tail is the longest run of levels ending at the coarsest (below
    the finest) with at most tail_rows unknowns and at most 64 rows in any
    colour (one round of the workgroup's row groups), whose vectors fit
    one workgroup's 160 KiB of LDS beside the 64 KiB of staging slots. Block
    smoothers and a direct coarsest level keep no tail."""
    A, H = p2
    ref = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2, tail_rows=0)
    ref.setup(H)
    lv = ref.describe()["levels"]

    def lds(n):  # one level's LDS block in FP64: x, rhs, res, int32 row offsets
        return (3 * n * 8 + 4 * (n + 1) + 15) // 16 * 16

    def expected(rows):
        k0, used = len(lv), 1024 * 8 * 8
        while (k0 - 1 >= 1 and lv[k0 - 1]["n"] <= rows and
               lv[k0 - 1]["max_color_rows"] <= 64 and
               used + lds(lv[k0 - 1]["n"]) <= 160 * 1024):
            k0 -= 1
            used += lds(lv[k0]["n"])
        return k0 if k0 < len(lv) else -1

    sizes = [v["n"] for v in lv]
    for rows in (sizes[-1], sizes[-2], sizes[0], 4096):
        s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2, tail_rows=rows)
        s.setup(H)
        d = s.describe()
        assert d["tail_level"] == expected(rows), (rows, d["tail_level"])
        assert d["tail_lds_bytes"] <= 160 * 1024
    assert expected(sizes[-1]) == len(lv) - 1
    d = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2, coarse_direct=1)
    d.setup(H)
    assert d.describe()["tail_level"] == -1
    with pytest.raises(gs.GpuSolverError, match="tail_rows must be non-negative"):
        gs.Solver(gs.PREC_DOUBLE, tail_rows=-1)
