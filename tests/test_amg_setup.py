"""Smoothed-aggregation hierarchy and problem generators. CPU only.

The reference builds its hierarchy with amgcl (unpinned master, absent
here), so aggregate / P / R / A_c values are "parity unpinned"; these tests
pin the invariants the reference's own AMG tests check
(examples/amgSolver/tests/MultigridUnitTests.cc:239-392) and what the
mathematics fixes: aggregates partition the aggregated rows, R = P^T,
A_c = R A P, A_c symmetric positive definite, the tentative prolongation
reproduces the near-nullspace, levels shrink.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs


def hier_lists(H):
    A = [H.get("A", k).to_scipy() for k in range(H.levels)]
    P = [H.get("P", k).to_scipy() for k in range(H.levels - 1)]
    R = [H.get("R", k).to_scipy() for k in range(H.levels - 1)]
    return A, P, R


@pytest.fixture(scope="module")
def poisson_hier():
    A = gs.poisson2d(64, 64)
    return A, gs.Hierarchy.smoothed_aggregation(A, min_dofs=100)


def test_levels_shrink_and_stop(poisson_hier):
    A, H = poisson_hier
    As, Ps, Rs = hier_lists(H)
    sizes = [a.shape[0] for a in As]
    assert sizes[0] == 4096
    assert all(b < a for a, b in zip(sizes, sizes[1:]))
    assert sizes[-1] <= 100 and sizes[-2] > 100


def test_galerkin_and_transpose(poisson_hier):
    _, H = poisson_hier
    As, Ps, Rs = hier_lists(H)
    for k in range(len(Ps)):
        assert abs(Rs[k] - Ps[k].T).max() == 0
        G = (Rs[k] @ As[k] @ Ps[k]).toarray()
        assert np.abs(As[k + 1].toarray() - G).max() <= 1e-12 * np.abs(G).max()


def test_coarse_operators_spd(poisson_hier):
    _, H = poisson_hier
    As, _, _ = hier_lists(H)
    for a in As[1:]:
        d = a.toarray()
        assert np.abs(d - d.T).max() <= 1e-12 * np.abs(d).max()
        np.linalg.cholesky(d)


def test_aggregates_partition(poisson_hier):
    _, H = poisson_hier
    As, Ps, _ = hier_lists(H)
    for k in range(len(Ps)):
        agg = H.aggregates(k)
        nc = As[k + 1].shape[0]
        assert agg.min() >= -1 and agg.max() == nc - 1
        assert set(np.unique(agg[agg >= 0])) == set(range(nc))


def test_prolongation_smoothing_preserves_constants_away_from_boundary():
    """P = (I - w D_f^-1 A_f) P_tent; A_f annihilates constants in rows
    whose weak connections were lumped and that touch no Dirichlet
    boundary, so P 1 = 1 there."""
    A = gs.poisson2d(32, 32)
    H = gs.Hierarchy.smoothed_aggregation(A, min_dofs=50)
    P = H.get("P", 0).to_scipy()
    p1 = P @ np.ones(P.shape[1])
    xy = np.indices((32, 32)).reshape(2, -1)
    interior = (xy[0] > 1) & (xy[0] < 30) & (xy[1] > 1) & (xy[1] < 30)
    np.testing.assert_allclose(p1[interior], 1.0, rtol=1e-12)


def test_vector_problem_with_nullspace():
    E, rhs, ns = gs.elasticity(8)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, min_dofs=100, nullspace=ns)
    As, Ps, Rs = hier_lists(H)
    assert len(As) >= 2
    for k in range(len(Ps)):
        assert As[k + 1].shape[0] % 6 == 0  # six modes per aggregate
        G = (Rs[k] @ As[k] @ Ps[k]).toarray()
        assert np.abs(As[k + 1].toarray() - G).max() <= 1e-12 * np.abs(G).max()
    # The coarse space spans the near-nullspace: project the rigid modes.
    P = Ps[0].toarray()
    coef, *_ = np.linalg.lstsq(P, ns, rcond=None)
    # Prolongation smoothing perturbs the tentative basis, so the fit is
    # close but not exact; the tentative basis itself is exact (QR).
    assert np.linalg.norm(P @ coef - ns) <= 0.2 * np.linalg.norm(ns)


def test_generator_sizes_match_configs():
    A = gs.poisson2d(256, 256)
    assert (A.rows, A.nnz) == (65536, 326656)
    A = gs.poisson3d(128)
    assert (A.rows, A.nnz) == (2097152, 14581760)


def test_elasticity_counts_64():
    """774 144 free DOF and 3 774 202 3x3 blocks for the 64^3-vertex cube."""
    E, rhs, ns = gs.elasticity(64)
    assert E.rows == 774144
    gp, gc = gs.block_graph(E, 3)
    assert gc.size + E.rows // 3 == 3774202
    assert E.nnz == 33967818


def test_elasticity_operator_properties():
    E, rhs, ns = gs.elasticity(6)
    S = E.to_scipy()
    assert abs(S - S.T).max() <= 1e-14 * abs(S).max()
    w = np.linalg.eigvalsh(S.toarray())
    assert w.min() > 0
    # Rigid translations are in the nullspace of every element; rows of
    # vertices not adjacent to the clamped face see zero force.
    f = S @ ns[:, 0]
    nv = 6
    xs = np.repeat(np.arange(1, nv), 3)  # x index per DOF in free ordering is periodic
    # free vertex ids: x in 1..nv-1 fastest
    x_of_vertex = np.tile(np.arange(1, nv), nv * nv)
    x_of_dof = np.repeat(x_of_vertex, 3)
    far = x_of_dof >= 2
    assert np.abs(f[far]).max() <= 1e-12
    del xs


def test_elasticity_element_matrix_reference_formula():
    """One tet, compared with the reference formula written in numpy
    (examples/elasticity/LinearElasticity.ipp:239-266)."""
    E, _, _ = gs.elasticity(2, young=1.0, poisson=0.3)
    # 2^3 vertices, x = 0 face clamped -> 4 free vertices (x = 1).
    assert E.rows == 12
    lam = 0.3 / (1.3 * 0.4)
    mu = 1 / 2.6
    K = np.zeros((24, 24))
    corners = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    vid = {c: i for i, c in enumerate(corners)}
    import itertools
    for perm in itertools.permutations(range(3)):
        c = [0, 0, 0]
        pts = [tuple(c)]
        for ax in perm:
            c[ax] = 1
            pts.append(tuple(c))
        X = np.array(pts, float)
        Dm = (X[1:] - X[0]).T
        V = abs(np.linalg.det(Dm)) / 6
        G = np.linalg.inv(Dm)  # rows are grad N1..N3
        g = np.vstack([-G.sum(0), G])
        for i in range(4):
            for j in range(4):
                H = mu * (g[j] @ g[i]) * np.eye(3) + mu * np.outer(g[j], g[i]) + lam * np.outer(g[i], g[j])
                a, b = vid[pts[i]], vid[pts[j]]
                K[3 * a:3 * a + 3, 3 * b:3 * b + 3] += V * H
    free = [vid[c] for c in corners if c[0] == 1]
    dofs = np.concatenate([[3 * v, 3 * v + 1, 3 * v + 2] for v in free])
    np.testing.assert_allclose(E.to_scipy().toarray(), K[np.ix_(dofs, dofs)], rtol=1e-12, atol=1e-15)


def test_varcoef_operator():
    A = gs.varcoef3d(12, seed=3, sigma=2.0, scale=2.0)
    S = A.to_scipy()
    assert abs(S - S.T).max() <= 1e-13 * abs(S).max()
    # Row sums are the Dirichlet boundary terms: zero in the interior.
    rs = np.asarray(S.sum(axis=1)).ravel()
    idx = np.indices((12, 12, 12)).reshape(3, -1)
    interior = np.all((idx > 0) & (idx < 11), axis=0)
    assert np.abs(rs[interior]).max() <= 1e-12 * abs(S).max()
    assert rs[~interior].min() > 0


def test_uniform_stream_reproducible_in_numpy():
    """The SplitMix64 counter stream, restated with numpy uint64 arithmetic."""
    seed, n = 12345, 1000
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    np.testing.assert_array_equal(gs.uniform_pm1(seed, n), 2 * u - 1)


def test_gaussian_field_moments():
    g = gs.gaussian_field(24, 24, 24, seed=5, sigma=3.0)
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1) < 1e-12
    g2 = g.reshape(24, 24, 24)
    # smoothed: neighbouring cells strongly correlated
    c = np.corrcoef(g2[:, :, :-1].ravel(), g2[:, :, 1:].ravel())[0, 1]
    assert c > 0.8
