"""GPU parity of the solve engine against the CPU oracle, through the C ABI.

Bars (written into each test):
  - SpMV and one full V-cycle (every smoother, reference coarse mode): bit
    for bit equal to the oracle in FP64 and FP32 -- same row orders, no FMA
    contraction, correctly rounded division on both sides.
  - PCG: FP64 solution within 1e-6 relative (2-norm) of the FP64 oracle at
    a 1e-8 relative residual, iteration count within +-2 (reference
    convention), FP64 true relative residual <= 1.5e-8.
"""
import numpy as np
import pytest

import gpusolver as gs
import oracle_api as orc

pytestmark = pytest.mark.gpu


def hier_view(H, perms=None):
    L = H.levels
    A = [H.get("A", k) for k in range(L)]
    P = [H.get("P", k) for k in range(L - 1)]
    R = [H.get("R", k) for k in range(L - 1)]
    hv = orc.HierView([(a.ptr, a.col, a.val) for a in A], [(p.ptr, p.col, p.val) for p in P],
                      [(r.ptr, r.col, r.val) for r in R], perms)
    return hv, A


@pytest.fixture(scope="module")
def cfg1():
    """2D 5-point Poisson, 256 x 256 interior (65 536 unknowns)."""
    A = gs.poisson2d(256, 256)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=1000)
    return A, H


@pytest.fixture(scope="module")
def elast():
    E, rhs, ns = gs.elasticity(12)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=500, nullspace=ns)
    return E, rhs, H


def oracle_flags(prec):
    return dict(cg_double=prec != gs.PREC_SINGLE, mg_double=prec == gs.PREC_DOUBLE)


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_SINGLE])
def test_spmv_bitexact(cfg1, prec):
    A, H = cfg1
    s = gs.Solver(prec)
    s.setup(H)
    x = gs.uniform_pm1(7, A.rows)
    y = s.spmv(x)
    if prec == gs.PREC_SINGLE:
        yo = orc.spmv(A.ptr, A.col, A.val.astype(np.float32), x.astype(np.float32))
    else:
        yo = orc.spmv(A.ptr, A.col, A.val, x)
    np.testing.assert_array_equal(y, yo.astype(np.float64))


@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("lo", [0, 1])
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_SINGLE])
def test_spmv_block_bitexact(elast, prec, lo, pairs):
    """The 3x3-block CG operator (rows in colour order, or in level-0 order
    through a row map; FP64 values as nine planes or as planes of value pairs
    read by the persistent kernel): every row sum bit for bit the oracle's
    CSR row sum."""
    E, rhs, H = elast
    s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, cg_block3=1, cg_level_order=lo,
                  cg_pair_layout=pairs)
    s.setup(H)
    assert s.describe()["cg_pair_layout"] == (pairs == 1 and prec == gs.PREC_DOUBLE)
    x = gs.uniform_pm1(8, E.rows)
    y = s.spmv(x)
    if prec == gs.PREC_SINGLE:
        yo = orc.spmv(E.ptr, E.col, E.val.astype(np.float32), x.astype(np.float32))
    else:
        yo = orc.spmv(E.ptr, E.col, E.val, x)
    np.testing.assert_array_equal(y, yo.astype(np.float64))


@pytest.mark.parametrize("grid", [8, 16, 64, 4096])
def test_persistent_spmv_walks_several_slices(grid):
    """The persistent FP64 CG SpMV with small grids: with 8 workgroups each
    wavefront walks about seven SELL slices of this 13 248-block-row
    operator (its stride loop's later passes); every row sum is bit for bit
    the oracle's. The PCG with it matches the default grid's iterations and
    solution to rounding (p.Ap is grouped by workgroup)."""
    E, rhs, _ = gs.elasticity(24, young=1.0, poisson=0.3, hetero_sigma=2.0, hetero_scale=1.5,
                              seed=2)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=500)
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=0, cg_block3=1,
                  persist_grid=grid)
    s.setup(H)
    assert s.describe()["cg_pair_layout"]
    slices = (E.rows // 3 + 63) // 64
    assert grid > 8 or slices // 8 >= 6 * 4  # >= 6 slices per wavefront at grid 8
    x = gs.uniform_pm1(8, E.rows)
    np.testing.assert_array_equal(s.spmv(x), orc.spmv(E.ptr, E.col, E.val, x))
    xs, info = s.solve(rhs, tol=1e-8, maxit=2000)
    d = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=0, cg_block3=1)
    d.setup(H)
    xd, infod = d.solve(rhs, tol=1e-8, maxit=2000)
    assert info.converged and abs(info.iterations - infod.iterations) <= 1
    assert np.linalg.norm(xs - xd) <= 1e-10 * np.linalg.norm(xd)


# Operator storage: -1 = one thread per row everywhere (SELL-64), 0 = lane
# groups for every operator, None = the default selection by row length.
GROUPS = [-1.0, 0.0, None]


def group_opts(grp):
    return {} if grp is None else dict(group_min_scalar=grp, group_min_block=grp,
                                       full_group_min_scalar=grp, full_group_min_block=grp)


@pytest.mark.parametrize("grp", GROUPS)
@pytest.mark.parametrize("ref", [0, 1])
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED, gs.PREC_SINGLE])
@pytest.mark.parametrize("kind,omega", [(gs.SGS, 1.0), (gs.SGS, 1.3), (gs.DAMPED_JACOBI, 0.7)])
def test_vcycle_bitexact_scalar(cfg1, prec, kind, omega, ref, grp):
    """ref = 1: the reference's coarse-level correction aliasing."""
    A, H = cfg1
    s = gs.Solver(prec, smoother=kind, omega=omega, coarse_steps=2, reference_vcycle=ref,
                  **group_opts(grp))
    s.setup(H)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(11, A.rows)
    z = s.vcycle(r)
    zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=kind, omegas=omega,
                    coarse_steps=2, reference_vcycle=bool(ref))
    np.testing.assert_array_equal(z, zo)


@pytest.mark.parametrize("grp", GROUPS)
@pytest.mark.parametrize("ref", [0, 1])
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_SINGLE])
@pytest.mark.parametrize("kind", [gs.BLOCK_GS, gs.BLOCK_JACOBI])
def test_vcycle_bitexact_block(elast, prec, kind, ref, grp):
    E, rhs, H = elast
    s = gs.Solver(prec, smoother=kind, omega=0.7, coarse_steps=2, reference_vcycle=ref,
                  **group_opts(grp))
    s.setup(H)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(5, E.rows)
    z = s.vcycle(r)
    zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=kind, omegas=0.7,
                    coarse_steps=2, reference_vcycle=bool(ref))
    np.testing.assert_array_equal(z, zo)


@pytest.mark.parametrize("grp", [-1.0, 0.0])
def test_vcycle_signed_zeros_match_oracle(cfg1, elast, grp):
    """Left-out diagonal entries are added as -0.0 inside the kernels; the
    results must match the oracle (which skips them) bit for bit, signs of
    zeros included: residuals with exact zeros, compared as raw bits."""
    A, H = cfg1
    r = gs.uniform_pm1(12, A.rows)
    r[::3] = 0.0
    r[1::7] = -0.0
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2, **group_opts(grp))
    s.setup(H)
    hv, _ = hier_view(H)
    z = s.vcycle(r)
    zo = orc.vcycle(hv, r, kinds=gs.SGS, coarse_steps=2)
    assert np.array_equal(z.view(np.uint64), zo.view(np.uint64))
    np.testing.assert_array_equal(s.vcycle(np.zeros(A.rows)).view(np.uint64),
                                  orc.vcycle(hv, np.zeros(A.rows), kinds=gs.SGS,
                                             coarse_steps=2).view(np.uint64))
    E, rhs, HE = elast
    rb = gs.uniform_pm1(13, E.rows)
    rb[::4] = 0.0
    sb = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2,
                   **group_opts(grp))
    sb.setup(HE)
    hve, _ = hier_view(HE)
    zb = sb.vcycle(rb)
    zbo = orc.vcycle(hve, rb, kinds=gs.BLOCK_GS, omegas=0.7, coarse_steps=2)
    assert np.array_equal(zb.view(np.uint64), zbo.view(np.uint64))


@pytest.mark.parametrize("grp", [-1.0, 0.0])
@pytest.mark.parametrize("steps", [1, 2, 3])
def test_vcycle_bitexact_hybrid(elast, steps, grp):
    """Fine block GS, coarse block Jacobi (MultigridHybridBlockPreconditioner,
    src/Preconditioners/MultigridHybridBlockPreconditioner.cc:9-61), with 1-3
    sweeps on the coarse levels: the first sweep from x = 0 in closed form,
    the others alternating between x and the second buffer, a result left
    in the second buffer read by the prolongation in place. Bit for bit the
    oracle's V-cycle (which runs every sweep as b - A x)."""
    E, rhs, H = elast
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, coarse_smoother=gs.BLOCK_JACOBI,
                  omega=0.7, coarse_steps=steps, **group_opts(grp))
    s.setup(H)
    sweeps = [1] + [steps] * (H.levels - 1)
    s.set_level_weights(sweeps=sweeps)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(6, E.rows)
    kinds = [gs.BLOCK_GS] + [gs.BLOCK_JACOBI] * (H.levels - 1)
    zo = orc.vcycle(hv, r, kinds=kinds, omegas=0.7, pres=sweeps, posts=sweeps,
                    coarse_steps=steps)
    np.testing.assert_array_equal(s.vcycle(r), zo)


@pytest.mark.parametrize("sweeps1", [0, 2])
def test_vcycle_bitexact_hybrid_level_weights(elast, sweeps1):
    """The learned configurations' shape on the block hybrid: level 1 with
    block Jacobi unsmoothed (0 sweeps: the restriction zeroes its x and no
    closed-form sweep is written) or with two sweeps (the second read from
    the other buffer), level 2 with one; per-level damping; bit for bit the
    oracle's V-cycle."""
    E, rhs, H = elast
    s = gs.Solver(gs.PREC_MIXED, smoother=gs.BLOCK_GS, coarse_smoother=gs.BLOCK_JACOBI,
                  omega=0.7, coarse_steps=1)
    s.setup(H)
    L = H.levels
    sweeps = ([1, sweeps1] + [1] * L)[:L]
    omega = ([0.82, 1.01, 0.9] + [0.7] * L)[:L]
    s.set_level_weights(omega=omega, sweeps=sweeps)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(26, E.rows)
    kinds = [gs.BLOCK_GS] + [gs.BLOCK_JACOBI] * (L - 1)
    zo = orc.vcycle(hv, r, mg_double=False, kinds=kinds, omegas=omega, pres=sweeps,
                    posts=sweeps, coarse_steps=1)
    np.testing.assert_array_equal(s.vcycle(r), zo)


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
@pytest.mark.parametrize("steps", [1, 2])
def test_vcycle_bitexact_scalar_jacobi_coarse(cfg1, prec, steps):
    """Scalar levels: SGS on the finest, damped Jacobi below (the coarse
    levels' Jacobi sweeps as in the block hybrid), 1 or 2 sweeps."""
    A, H = cfg1
    s = gs.Solver(prec, smoother=gs.SGS, coarse_smoother=gs.DAMPED_JACOBI, omega=0.8,
                  coarse_steps=steps, tail_rows=0)
    s.setup(H)
    sweeps = [1] + [steps] * (H.levels - 1)
    s.set_level_weights(sweeps=sweeps)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(16, A.rows)
    kinds = [gs.SGS] + [gs.DAMPED_JACOBI] * (H.levels - 1)
    zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=kinds, omegas=0.8,
                    pres=sweeps, posts=sweeps, coarse_steps=steps)
    np.testing.assert_array_equal(s.vcycle(r), zo)


def test_pcg_parity_reference_semantics(cfg1):
    """The reference's own V-cycle (correction aliasing included) in FP64
    against the oracle's restatement of it."""
    A, H = cfg1
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2, reference_vcycle=1)
    s.setup(H)
    hv, _ = hier_view(H)
    b = gs.uniform_pm1(12, A.rows)
    x, xo, info, true = _pcg_parity(s, hv, A, b, gs.PREC_DOUBLE, reference_vcycle=True)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert true <= 1.5e-8


def _pcg_parity(s, hv, A, b, prec, tol=1e-8, **okw):
    x, info = s.solve(b, tol=tol, maxit=1000)
    xo, ito, rro, hist = orc.pcg(hv, b, tol=tol, maxit=1000, **oracle_flags(prec), **okw)
    assert info.converged
    assert abs(info.iterations - ito) <= 2, (info.iterations, ito)
    S = A.to_scipy()
    true = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
    return x, xo, info, true


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_pcg_parity_cfg1(cfg1, prec):
    A, H = cfg1
    s = gs.Solver(prec, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    hv, _ = hier_view(H)
    b = gs.uniform_pm1(1, A.rows)
    x, xo, info, true = _pcg_parity(s, hv, A, b, prec)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert true <= 1.5e-8 and abs(info.true_relres - true) <= 1e-3 * true + 1e-16


def test_pcg_single_precision_matches_reference_semantics(cfg1):
    """FP32 throughout, the reference's precision, tol 1e-6 (its default)."""
    A, H = cfg1
    s = gs.Solver(gs.PREC_SINGLE, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    hv, _ = hier_view(H)
    b = gs.uniform_pm1(2, A.rows)
    x, info = s.solve(b, tol=1e-6, maxit=1000)
    xo, ito, rro, _ = orc.pcg(hv, b, tol=1e-6, cg_double=False, mg_double=False)
    assert abs(info.iterations - ito) <= 2
    assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)


@pytest.mark.parametrize("ordering", ["rcm", "hierarchical"])
def test_pcg_with_level_permutations(cfg1, ordering):
    """Level permutations as the reference drivers pass them (RCM on every
    level, or HierarchicalReordering's per-level list; the CG operator uses
    level 0's, order kept)."""
    A, H = cfg1
    if ordering == "rcm":
        perms = [gs.rcm(H.get("A", k))[0] for k in range(H.levels)]
    else:
        perms = [o2n for o2n, _ in gs.hierarchical_reordering(H)[0]]
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS)
    s.setup(H, perms)
    hv, _ = hier_view(H, perms)
    b = gs.uniform_pm1(3, A.rows)
    x, xo, info, true = _pcg_parity(s, hv, A, b, gs.PREC_DOUBLE)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    # The V-cycle in the permuted level-0 ordering is still bit-exact.
    r = gs.uniform_pm1(4, A.rows)
    np.testing.assert_array_equal(s.vcycle(r), orc.vcycle(hv, r))


def test_pcg_coarse_direct(cfg1):
    A, H = cfg1
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_direct=1)
    s.setup(H)
    hv, _ = hier_view(H)
    b = gs.uniform_pm1(5, A.rows)
    x, xo, info, true = _pcg_parity(s, hv, A, b, gs.PREC_DOUBLE, direct_coarse=True)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert true <= 1.5e-8


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_transfer_row_orders_bitexact(elast, prec):
    """Transfer rows executed in another order (the level's own order, and
    prolongation rows sorted by length within windows, SELL-C-sigma) and
    stored as 3x3 blocks give the same V-cycle bit for bit as scalar rows in
    the colour order with no sorting."""
    E, rhs, H = elast
    r = gs.uniform_pm1(11, E.rows)
    out = []
    for blk, lo, sigma in ((0, 0, 0), (0, 1, 0), (0, 0, 1024), (0, 1, 1024), (0, 1, 37),
                           (1, 0, 0), (1, 1, 0), (1, 1, 1024), (1, 0, 37)):
        s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, transfer_sigma=sigma,
                      transfer_level_order=lo, block_transfers=blk)
        s.setup(H)
        assert all(l["transfer_b3"] == bool(blk) for l in s.describe()["levels"][:-1])
        out.append(s.vcycle(r))
    for z in out[1:]:
        np.testing.assert_array_equal(out[0], z)


@pytest.mark.parametrize("grp", GROUPS)
def test_zero_start_lower_bitexact(cfg1, elast, grp):
    """The forward colours of a sweep from x = 0 read only earlier-colour
    entries: the same V-cycle bit for bit as reading whole rows, for scalar
    SGS and 3x3 block GS, one and two sweeps."""
    A, H = cfg1
    E, rhs, HE = elast
    for mat, hier, kind, om in ((A, H, gs.SGS, 1.1), (E, HE, gs.BLOCK_GS, 0.7)):
        r = gs.uniform_pm1(21, mat.rows)
        for sweeps in (1, 2):
            out = []
            for low in (0, 1, 10):  # none, level 0, every level
                s = gs.Solver(gs.PREC_DOUBLE, smoother=kind, omega=om, pre=sweeps, post=sweeps,
                              coarse_steps=2, zero_start_lower=low, **group_opts(grp))
                s.setup(hier)
                out.append(s.vcycle(r))
            for z in out[1:]:
                assert np.array_equal(out[0].view(np.uint64), z.view(np.uint64))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 333])
def test_direct_level_tiled_product(n):
    """A one-level hierarchy with the direct coarse solve: the V-cycle is
    x = A^-1 r through the lower-tile symmetric product, at sizes on both
    sides of the 64-row tile; checked against numpy's dense solve."""
    import scipy.sparse as sp
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    dense = B @ B.T + n * np.eye(n)
    dense[np.abs(dense) < 0.5] = 0.0  # sparsify off-diagonal entries
    dense = 0.5 * (dense + dense.T) + n * np.eye(n)
    A = gs.Csr.from_scipy(sp.csr_matrix(dense))
    H = gs.Hierarchy.from_levels([A], [], [])
    r = rng.standard_normal(n)
    ref = np.linalg.solve(dense, r)
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_direct=1)
    s.setup(H)
    assert s.describe()["levels"][0]["direct"]
    z = s.vcycle(r)
    assert np.linalg.norm(z - ref) <= 1e-12 * np.linalg.norm(ref)
    m = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_direct=1)
    m.setup(H)
    zm = m.vcycle(r)
    assert np.linalg.norm(zm - ref) <= 1e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_pcg_parity_elasticity_block_gs(elast, prec, pairs):
    E, rhs, H = elast
    s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1,
                  cg_pair_layout=pairs)
    s.setup(H)
    hv, _ = hier_view(H)
    x, xo, info, true = _pcg_parity(s, hv, E, rhs, prec, kinds=gs.BLOCK_GS, omegas=0.7)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert true <= 1.5e-8


def test_graph_and_eager_paths_agree(cfg1):
    """The captured-graph loop and the eager loop run the same kernels."""
    A, H = cfg1
    b = gs.uniform_pm1(8, A.rows)
    outs = []
    for gi in (0, 1, 8):
        s = gs.Solver(gs.PREC_DOUBLE, graph_iters=gi)
        s.setup(H)
        for _ in range(2):  # replays with eager work in between
            x, info = s.solve(b, tol=1e-8)
            outs.append((x, info.iterations))
    for x, it in outs[1:]:
        np.testing.assert_array_equal(x, outs[0][0])
        assert it == outs[0][1]


def test_history_and_iteration_convention(cfg1):
    A, H = cfg1
    s = gs.Solver(gs.PREC_DOUBLE)
    s.setup(H)
    b = gs.uniform_pm1(9, A.rows)
    x, info = s.solve(b, tol=1e-8)
    h = s.history()
    assert len(h) == info.iterations + 2  # initial entry + (iterations + 1) updates
    assert h[-1] < 1e-16 <= h[-2]
    x, info3 = s.solve(b, tol=1e-8, maxit=3)
    assert info3.iterations == 3 and not info3.converged


def test_identity_preconditioner_cg_sanity():
    """tests/metal/mainCGSanityCheck.cc:15-84 on the GPU."""
    import scipy.sparse as sp
    T = sp.diags([[-1.0] * 4, [4.0] * 5, [-1.0] * 4], [-1, 0, 1]).tocsr()
    A = gs.Csr.from_scipy(T)
    H = gs.Hierarchy.from_levels([A], [], [])
    xs = np.arange(1, 6, dtype=np.float64)
    b = T @ xs
    for prec in (gs.PREC_SINGLE, gs.PREC_DOUBLE):
        s = gs.Solver(prec, preconditioner=0)
        s.setup(H)
        x, info = s.solve(b, tol=1e-6, maxit=100)
        assert np.linalg.norm(x - xs) < 1e-4
        assert np.linalg.norm(T @ x - b) < 1e-4


def test_zero_rhs_and_nan():
    import scipy.sparse as sp
    T = gs.Csr.from_scipy(sp.diags([[2.0] * 10], [0]).tocsr())
    H = gs.Hierarchy.from_levels([T], [], [])
    s = gs.Solver(gs.PREC_DOUBLE, preconditioner=0)
    s.setup(H)
    x, info = s.solve(np.zeros(10), x0=np.ones(10))
    np.testing.assert_array_equal(x, np.zeros(10))
    b = np.ones(10)
    b[3] = np.nan
    with pytest.raises(gs.GpuSolverError, match="NaN"):
        s.solve(b)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 130, 700])
def test_device_spd_inverse_matches_numpy(n):
    """The coarse-level inverse built on the GPU (blocked Gauss-Jordan,
    64 x 64 pivot blocks) against numpy, across block-size boundaries."""
    rng = np.random.default_rng(n)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    a = (q * rng.uniform(1.0, 20.0, n)) @ q.T
    inv = gs.dense_spd_inverse_device(a)
    ref = np.linalg.inv(a)
    assert np.max(np.abs(inv - ref)) <= 1e-12 * np.max(np.abs(ref))
    with pytest.raises(gs.GpuSolverError):
        gs.dense_spd_inverse_device(np.array([[1.0, 2.0], [2.0, 1.0]]))


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_pcg_direct_truncate(elast, prec):
    """coarse_direct + direct_truncate: the direct solve moves up to the first
    level within max_direct; the oracle runs the same truncated hierarchy."""
    E, rhs, H = elast
    nl1 = H.get("A", 1).rows
    s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=0, coarse_direct=1,
                  direct_truncate=1, max_direct=nl1, cg_block3=1)
    s.setup(H)
    d = s.describe()
    assert len(d["levels"]) == 2 and d["levels"][1]["direct"]
    A = [H.get("A", k) for k in range(2)]
    hv = orc.HierView([(a.ptr, a.col, a.val) for a in A],
                      [(p.ptr, p.col, p.val) for p in [H.get("P", 0)]],
                      [(r.ptr, r.col, r.val) for r in [H.get("R", 0)]])
    x, xo, info, true = _pcg_parity(s, hv, E, rhs, prec, kinds=gs.BLOCK_GS, omegas=0.7,
                                    coarse_steps=0, direct_coarse=True)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert true <= 1.5e-8


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_SINGLE])
def test_vcycle_bitexact_zero_sweep_level(elast, prec):
    """Per-level weights with level 1 unsmoothed (0 sweeps, a no-op block
    smoother): the level's residual is its right-hand side and is restricted
    directly; bit-exact against the oracle's full V-cycle."""
    E, rhs, H = elast
    L = H.levels
    s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2)
    s.setup(H)
    om = [0.9, 1.1] + [0.7] * (L - 2)
    sw = [1, 0] + [1] * (L - 2)
    s.set_level_weights(omega=om, sweeps=sw)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(9, E.rows)
    zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=gs.BLOCK_GS, omegas=om,
                    pres=sw, posts=sw, coarse_steps=2)
    np.testing.assert_array_equal(s.vcycle(r), zo)


@pytest.fixture(scope="module")
def hetero_small():
    """Heterogeneous coefficients at test size: elasticity on 14^3 vertices
    with E = exp(1.5 g) (g smoothed with sigma 1.5 cells, seed 3), and the
    variable-coefficient Poisson problem on 24^3 cells with k = exp(2 g)
    (sigma 2 cells, seed 4)."""
    E, rhs, _ = gs.elasticity(14, young=1.0, poisson=0.3, hetero_sigma=1.5, hetero_scale=1.5,
                              seed=3)
    HE = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                           min_dofs=300)
    V = gs.varcoef3d(24, seed=4, sigma=2.0, scale=2.0)
    HV = gs.Hierarchy.smoothed_aggregation(V, block_size=1, eps_strong=0.1, max_levels=10,
                                           min_dofs=200)
    return E, rhs, HE, V, HV


@pytest.mark.parametrize("grp", GROUPS)
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED, gs.PREC_SINGLE])
def test_vcycle_bitexact_heterogeneous(hetero_small, prec, grp):
    """Coefficient contrast of several decades (E or k = exp(c g)): block GS
    on the elasticity hierarchy and SGS on the variable-coefficient one,
    bit for bit the oracle's V-cycle."""
    E, rhs, HE, V, HV = hetero_small
    for mat, hier, kind, om, seed in ((E, HE, gs.BLOCK_GS, 0.7, 31), (V, HV, gs.SGS, 1.0, 32)):
        s = gs.Solver(prec, smoother=kind, omega=om, coarse_steps=2, **group_opts(grp))
        s.setup(hier)
        hv, _ = hier_view(hier)
        r = gs.uniform_pm1(seed, mat.rows)
        zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=kind, omegas=om,
                        coarse_steps=2)
        np.testing.assert_array_equal(s.vcycle(r), zo)


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_pcg_parity_heterogeneous(hetero_small, prec):
    E, rhs, HE, V, HV = hetero_small
    for mat, hier, b, kw in ((E, HE, rhs, dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=0,
                                                cg_block3=1)),
                             (V, HV, gs.uniform_pm1(33, V.rows), dict(smoother=gs.SGS,
                                                                       coarse_steps=2))):
        s = gs.Solver(prec, **kw)
        s.setup(hier)
        hv, _ = hier_view(hier)
        x, xo, info, true = _pcg_parity(s, hv, mat, b, prec, kinds=kw["smoother"],
                                        omegas=kw.get("omega", 1.0),
                                        coarse_steps=kw["coarse_steps"])
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
        assert true <= 1.5e-8


@pytest.mark.parametrize("sweeps", [[1, 0, 1], [1, 0, 0], [2, 1, 0]])
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_vcycle_bitexact_unsmoothed_sgs_levels(cfg1, prec, sweeps):
    """Per-level weights that leave coarse symmetric Gauss-Seidel levels
    unsmoothed (0 sweeps: the restricted residual goes straight down and the
    prolongated correction straight back): bit for bit the oracle's V-cycle
    with the same per-level sweeps; the finest level still refuses 0."""
    A, H = cfg1
    L = H.levels
    sw = sweeps + [1] * (L - len(sweeps))
    om = [1.1, 0.9, 1.2] + [1.0] * (L - 3)
    s = gs.Solver(prec, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    s.set_level_weights(omega=om, sweeps=sw)
    hv, _ = hier_view(H)
    r = gs.uniform_pm1(19, A.rows)
    zo = orc.vcycle(hv, r, mg_double=prec == gs.PREC_DOUBLE, kinds=gs.SGS, omegas=om, pres=sw,
                    posts=sw, coarse_steps=2)
    np.testing.assert_array_equal(s.vcycle(r), zo)
    with pytest.raises(gs.GpuSolverError, match="at least one smoothing step"):
        s.set_level_weights(sweeps=[0] + sw[1:])
