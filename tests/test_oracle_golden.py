"""Pins the CPU oracle against the reference's own known-answer tests and
against closed forms. CPU only.

Golden vectors (from the reference's test sources):
  - tests/cuda/testCudaGaussSeidel.cc:15-111   one colour   -> {1, 0, 3, 0}
  - tests/cuda/testCudaGaussSeidel.cc:114-222  two colours  -> {1, 3, 3, 4.75}
  - tests/cuda/testCudaBlockOps.cc:15-78       block inverses -> {2, 6, 12, 9, 10, 18}
  - tests/cuda/testCudaSpmv.cc:17-56           dot = 32
  - tests/cuda/testCudaSpmv.cc:123-195         SpMV -> {6, 13, 19}
  - tests/cuda/testCudaSpmv.cc:197-273         updated values -> {8, 14}
  - tests/metal/DotProductTests.cc:27-140      dot sizes vs double, tol 1e-4 max(1, |e|)
  - tests/metal/mainCGSanityCheck.cc:15-84     5x5 CG, ||x - x*|| < 1e-4
  - tests/metal/gaussSeidelSymmUnittest.cc     8x8 Poisson SGS, r^T M^-1 r > 0
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_api as orc


def csr(dense):
    m = sp.csr_matrix(np.asarray(dense, np.float64))
    m.sort_indices()
    return (m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64))


TRI4 = [[4, -1, 0, 0], [-1, 4, -1, 0], [0, -1, 4, -1], [0, 0, -1, 4]]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gs_one_color(dt):
    ptr, col, val = csr(TRI4)
    order = np.array([0, 2, 1, 3], np.int32)
    rhs = np.array([4, 8, 12, 16], dt)
    x = np.zeros(4, dt)
    orc.gs_color(ptr, col, val.astype(dt), order, 0, 2, rhs, x, 1.0)
    np.testing.assert_array_equal(x, np.array([1, 0, 3, 0], dt))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gs_two_colors(dt):
    ptr, col, val = csr(TRI4)
    order = np.array([0, 2, 1, 3], np.int32)
    rhs = np.array([4, 8, 12, 16], dt)
    x = np.zeros(4, dt)
    orc.gs_color(ptr, col, val.astype(dt), order, 0, 2, rhs, x, 1.0)
    orc.gs_color(ptr, col, val.astype(dt), order, 2, 2, rhs, x, 1.0)
    np.testing.assert_array_equal(x, np.array([1, 3, 3, 4.75], dt))


def test_gs_zero_diagonal_raises():
    ptr, col, val = csr([[0, 1], [1, 2]])
    x = np.zeros(2)
    with pytest.raises(RuntimeError, match="zero diagonal"):
        orc.gs_color(ptr, col, val, np.array([0, 1], np.int32), 0, 2, np.ones(2), x, 1.0)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_apply_block_inverses(dt):
    inv = np.array([2, 0, 0, 0, 3, 0, 0, 0, 4, 1, 1, 0, 0, 2, 0, 0, 0, 3], dt)
    res = np.array([1, 2, 3, 4, 5, 6], dt)
    np.testing.assert_array_equal(orc.apply_block_inverses(inv, res),
                                  np.array([2, 6, 12, 9, 10, 18], dt))


def test_dot_32():
    a = np.array([1, 2, 3], np.float32)
    b = np.array([4, 5, 6], np.float32)
    assert orc.dot(a, b) == 32.0


@pytest.mark.parametrize("n", [1, 17, 255, 256, 257, 1000, 100000])
def test_dot_sizes(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = rng.uniform(-1, 1, n).astype(np.float32)
    e = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    assert abs(orc.dot(x, y) - e) <= 1e-4 * max(1.0, abs(e))
    assert abs(orc.dot(x.astype(np.float64), y.astype(np.float64)) - e) <= 1e-12 * max(1.0, abs(e))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_spmv_known(dt):
    ptr, col, val = csr([[4, 1, 0], [1, 3, 2], [0, 2, 5]])
    y = orc.spmv(ptr, col, val.astype(dt), np.array([1, 2, 3], dt))
    np.testing.assert_array_equal(y, np.array([6, 13, 19], dt))


def test_spmv_updated_values():
    ptr, col, _ = csr([[2, 1], [1, 3]])
    y = orc.spmv(ptr, col, np.array([4, 2, 2, 6], np.float64), np.array([1.0, 2.0]))
    np.testing.assert_array_equal(y, [8.0, 14.0])


def _dirichlet_laplacian(dims):
    import gpusolver as gs
    if len(dims) == 2:
        return gs.poisson2d(*dims)
    return gs.poisson3d(*dims)


@pytest.mark.parametrize("dims", [(13, 9), (7, 6, 5)])
def test_spmv_sine_eigenvectors(dims):
    """The discrete Dirichlet Laplacian has exact sine eigenvectors; A v = lambda v."""
    A = _dirichlet_laplacian(dims)
    idx = np.indices(dims[::-1]).reshape(len(dims), -1)[::-1]  # x fastest
    modes = [2, 3, 1][:len(dims)]
    v = np.ones(A.rows)
    lam = 0.0
    for d, (m, nd) in enumerate(zip(modes, dims)):
        v *= np.sin(np.pi * m * (idx[d] + 1) / (nd + 1))
        lam += 2 - 2 * np.cos(np.pi * m / (nd + 1))
    y = orc.spmv(A.ptr, A.col, A.val, v)
    assert np.linalg.norm(y - lam * v) <= 1e-12 * np.linalg.norm(lam * v)


def test_gs_diagonal_is_exact_solve():
    d = np.array([2.0, 4.0, 8.0, 5.0])
    ptr, col, val = csr(np.diag(d))
    x = np.zeros(4)
    b = np.array([1.0, 2.0, 3.0, 4.0])
    orc.gs_color(ptr, col, val, np.arange(4, dtype=np.int32), 0, 4, b, x, 1.0)
    np.testing.assert_array_equal(x, b / d)


def test_block_gs_block_diagonal_is_exact_solve():
    rng = np.random.default_rng(0)
    blocks = []
    for _ in range(3):
        m = rng.uniform(-1, 1, (3, 3))
        blocks.append(m @ m.T + 3 * np.eye(3))
    D = sp.block_diag(blocks).toarray()
    ptr, col, val = csr(D)
    inv = np.concatenate([np.linalg.inv(b).ravel() for b in blocks])
    b = rng.uniform(-1, 1, 9)
    x = np.zeros(9)
    orc.block_gs_color(ptr, col, val, np.arange(3, dtype=np.int32), inv, 0, 3, b, x, 1.0)
    np.testing.assert_allclose(x, np.linalg.solve(D, b), rtol=1e-12)


def _hier_single(A):
    """One-level hierarchy view of a CSR matrix."""
    return orc.HierView([(A[0], A[1], A[2])], [], [])


def test_cg_sanity_5x5():
    """tests/metal/mainCGSanityCheck.cc:15-84 with the identity preconditioner."""
    T = np.diag([4.0] * 5) + np.diag([-1.0] * 4, 1) + np.diag([-1.0] * 4, -1)
    xs = np.arange(1, 6, dtype=np.float64)
    b = T @ xs
    hv = _hier_single(csr(T))
    x, it, rr, hist = orc.pcg(hv, b, tol=1e-6, maxit=100, prec=0, cg_double=False, mg_double=False)
    assert np.linalg.norm(x - xs) < 1e-4
    assert np.linalg.norm(T @ x - b) < 1e-4


def test_cg_iteration_convention():
    """k PCG updates report k-1 (src/CGSolver.cc:257-260): with the identity
    preconditioner and 3 distinct eigenvalues CG is exact after 3 updates."""
    A = np.diag([1.0, 1.0, 2.0, 2.0, 3.0, 3.0])
    b = np.ones(6)
    x, it, rr, hist = orc.pcg(_hier_single(csr(A)), b, tol=1e-10, maxit=50, prec=0)
    assert it == 2
    assert len(hist) == 4  # initial + 3 updates
    np.testing.assert_allclose(x, b / np.diag(A), rtol=1e-12)


def test_cg_maxit_reports_maxit():
    import gpusolver as gs
    A = gs.poisson2d(20, 20)
    b = np.ones(A.rows)
    x, it, rr, hist = orc.pcg(_hier_single((A.ptr, A.col, A.val)), b, tol=1e-14, maxit=3, prec=0)
    assert it == 3 and len(hist) == 4


def test_cg_zero_rhs_returns_zero():
    A = np.diag([2.0, 3.0])
    x, it, rr, hist = orc.pcg(_hier_single(csr(A)), np.zeros(2), x0=np.ones(2), prec=0)
    np.testing.assert_array_equal(x, np.zeros(2))
    assert it == 0


def test_cg_nan_rhs_raises():
    A = np.diag([2.0, 3.0])
    with pytest.raises(RuntimeError, match="NaN"):
        orc.pcg(_hier_single(csr(A)), np.array([np.nan, 1.0]), prec=0)


def _poisson8():
    """tests/metal/gaussSeidelSymmUnittest.cc:32-70."""
    import gpusolver as gs
    return gs.poisson2d(8, 8)


def test_sgs_preconditioner_is_spd():
    """SGS as a one-level 'multigrid' (coarsest level = smoother only):
    r^T M^-1 r > 0 and M^-1 symmetric in FP64."""
    A = _poisson8()
    hv = orc.HierView([(A.ptr, A.col, A.val)], [], [])
    n = A.rows
    rhs = np.sin(0.17 * np.arange(n)) + 0.5
    z = orc.vcycle(hv, rhs, coarse_steps=1)
    assert rhs @ z > 0
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    Mu, Mv = orc.vcycle(hv, u, coarse_steps=2), orc.vcycle(hv, v, coarse_steps=2)
    assert abs(u @ Mv - v @ Mu) <= 1e-12 * abs(u @ Mv)


def test_sgs_matches_sequential_gs_on_bipartite_grid():
    """On a 2-coloured grid, one forward + backward multicolour sweep from
    zero equals red-black Gauss-Seidel written with scipy triangular solves."""
    A = _poisson8()
    S = A.to_scipy()
    n = A.rows
    b = np.sin(0.17 * np.arange(n)) + 0.5
    hv = orc.HierView([(A.ptr, A.col, A.val)], [], [])
    z = orc.vcycle(hv, b, coarse_steps=1)
    xy = np.indices((8, 8)).reshape(2, -1)
    red = ((xy[0] + xy[1]) % 2 == 0)
    perm = np.concatenate([np.where(red)[0], np.where(~red)[0]])
    P = S[perm][:, perm].toarray()
    bp = b[perm]
    L = np.tril(P)
    U = np.triu(P)
    x = np.linalg.solve(L, bp)                       # forward sweep from 0
    x = x + np.linalg.solve(U, bp - P @ x)            # backward sweep
    xr = np.empty(n)
    xr[perm] = x
    np.testing.assert_allclose(z, xr, rtol=1e-12, atol=1e-14)


def _hier(A, levels_min_dofs=100, block=1):
    import gpusolver as gs
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=block, min_dofs=levels_min_dofs)
    L = H.levels
    Al = [H.get("A", k) for k in range(L)]
    return orc.HierView([(a.ptr, a.col, a.val) for a in Al],
                        [(p.ptr, p.col, p.val) for p in (H.get("P", k) for k in range(L - 1))],
                        [(r.ptr, r.col, r.val) for r in (H.get("R", k) for k in range(L - 1))]), L


@pytest.mark.parametrize("kind,omega", [(1, 1.0), (1, 1.3), (0, 0.7)])
def test_standard_vcycle_is_symmetric(kind, omega):
    """Textbook V-cycle with pre = post symmetric smoothing is a symmetric
    operator: u^T M v = v^T M u in FP64 (pins sweep order, R = P^T use,
    x += P e)."""
    import gpusolver as gs
    A = gs.poisson2d(64, 64)
    hv, L = _hier(A)
    assert L >= 3
    rng = np.random.default_rng(9)
    u, v = rng.uniform(-1, 1, A.rows), rng.uniform(-1, 1, A.rows)
    kw = dict(kinds=kind, omegas=omega, coarse_steps=2)
    Mu, Mv = orc.vcycle(hv, u, **kw), orc.vcycle(hv, v, **kw)
    assert abs(u @ Mv - v @ Mu) <= 1e-11 * abs(u @ Mv)
    assert u @ Mu > 0


def test_reference_vcycle_aliasing_only_below_the_finest_level():
    """The reference prolongates into the coarse level's own `solution`
    vector (src/Smoothers/MultigridSmoother.cc:401-404): with two levels
    nothing is aliased and both variants agree bit for bit; with three or
    more they differ. With symmetric sweeps both stay symmetric operators:
    the reference variant is 2 (I - S A) C (I - A S) + S with the sweep S
    symmetric, the standard one S + S - S A S + (I - S A) C (I - A S)."""
    import gpusolver as gs
    A = gs.poisson2d(48, 48)
    H2 = gs.Hierarchy.smoothed_aggregation(A, min_dofs=100, max_levels=1)
    assert H2.levels == 2
    hv2 = orc.HierView([(a.ptr, a.col, a.val) for a in (H2.get("A", k) for k in range(2))],
                       [(H2.get("P", 0).ptr, H2.get("P", 0).col, H2.get("P", 0).val)],
                       [(H2.get("R", 0).ptr, H2.get("R", 0).col, H2.get("R", 0).val)])
    r = np.sin(0.3 * np.arange(A.rows))
    np.testing.assert_array_equal(orc.vcycle(hv2, r, reference_vcycle=True),
                                  orc.vcycle(hv2, r, reference_vcycle=False))
    hv, L = _hier(A, 50)
    assert L >= 3
    za = orc.vcycle(hv, r, reference_vcycle=True)
    zb = orc.vcycle(hv, r, reference_vcycle=False)
    assert np.abs(za - zb).max() > 1e-6
    rng = np.random.default_rng(1)
    u, v = rng.uniform(-1, 1, A.rows), rng.uniform(-1, 1, A.rows)
    Mu = orc.vcycle(hv, u, reference_vcycle=True)
    Mv = orc.vcycle(hv, v, reference_vcycle=True)
    assert abs(u @ Mv - v @ Mu) <= 1e-11 * abs(u @ Mv)


# ---------------------------------------------------------------------------
# Coarsest-level direct solve: dense SPD inverse through Cholesky.
# ---------------------------------------------------------------------------
def _spd(n, seed, lo=1.0, hi=10.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    return (q * rng.uniform(lo, hi, n)) @ q.T


@pytest.mark.parametrize("n", [1, 2, 5, 40])
def test_oracle_dense_spd_inverse_matches_numpy(n):
    a = _spd(n, n)
    inv = orc.dense_spd_inverse(a)
    ref = np.linalg.inv(a)
    assert np.max(np.abs(inv - ref)) <= 1e-12 * np.max(np.abs(ref))
    np.testing.assert_allclose(a @ inv, np.eye(n), atol=1e-12)


def test_oracle_dense_spd_inverse_closed_forms():
    assert orc.dense_spd_inverse(np.array([[4.0]]))[0, 0] == 0.25
    a = np.array([[2.0, 1.0], [1.0, 3.0]])  # inverse = [[3, -1], [-1, 2]] / 5
    np.testing.assert_allclose(orc.dense_spd_inverse(a), np.array([[3, -1], [-1, 2]]) / 5.0,
                               rtol=1e-15, atol=1e-16)


def test_oracle_dense_spd_inverse_rejects_indefinite():
    with pytest.raises(RuntimeError):
        orc.dense_spd_inverse(np.array([[1.0, 2.0], [2.0, 1.0]]))


# ---------------------------------------------------------------------------
# V-cycle operator in closed form (src/Smoothers/MultigridSmoother.cc:337-412).
# ---------------------------------------------------------------------------
def _sweep_matrix(A, classes, block, omega, sweeps):
    """Dense approximate inverse B of `sweeps` symmetric multicolour sweeps
    from x = 0 (x = B b). One colour update is x <- x + K_c (b - A x) with
    K_c = w E_c D_c^-1 E_c^T, D_c the colour's diagonal (3x3 diagonal
    blocks for block GS): rows of one colour do not couple, so the
    reference's (1 - w) x + w D^-1 (b - off-diagonal sum) is exactly this.
    The error propagates as I - K_c A; a symmetric sweep runs the colours
    forward, then backward, and B = (I - G) A^-1 for the product G."""
    n = A.shape[0]
    I = np.eye(n)
    Ks = []
    for rows in classes:
        idx = np.concatenate([np.arange(block * r, block * r + block) for r in rows]) \
            if block > 1 else np.asarray(rows)
        K = np.zeros((n, n))
        K[np.ix_(idx, idx)] = omega * np.linalg.inv(A[np.ix_(idx, idx)])
        Ks.append(K)
    G = I.copy()
    for _ in range(sweeps):
        for K in Ks:
            G = (I - K @ A) @ G
        for K in reversed(Ks):
            G = (I - K @ A) @ G
    return (I - G) @ np.linalg.inv(A)


def _classes(Ad, block):
    m = sp.csr_matrix(Ad)
    if block == 1:
        ptr, col = m.indptr.astype(np.int32), m.indices.astype(np.int32)
    else:
        ptr, col = orc.block_graph(m.indptr.astype(np.int32), m.indices.astype(np.int32), block)
    _, verts, offs, nc = orc.coloring(ptr, col)
    return [verts[offs[c]:offs[c + 1]] for c in range(nc)]


def _vcycle_closed_form(As, Ps, Rs, block, omega, coarse_steps, aliasing):
    """M with z = M r for the V-cycle over dense levels As (finest first):
    pre-smoothing from zero, residual, restriction, coarse correction,
    prolongation, post-smoothing. Standard: x += P e. The reference's
    storage on levels below the finest (aliasing): the prolongated
    correction overwrites the level's solution and is added to itself,
    x = 2 P e, dropping the pre-smoothed x."""
    def level(k):
        A = As[k]
        if k == len(As) - 1:
            return _sweep_matrix(A, _classes(A, block), block, omega, coarse_steps)
        B = _sweep_matrix(A, _classes(A, block), block, omega, 1)
        C = Ps[k] @ level(k + 1) @ Rs[k]
        I = np.eye(A.shape[0])
        X2 = 2.0 * C @ (I - A @ B) if (aliasing and k > 0) else B + C @ (I - A @ B)
        return X2 + B @ (I - A @ X2)
    return level(0)


@pytest.mark.parametrize("aliasing", [False, True])
@pytest.mark.parametrize("levels", [2, 3])
@pytest.mark.parametrize("problem", ["sgs", "block_gs"])
def test_vcycle_matches_closed_form_operator(problem, levels, aliasing):
    """orc.vcycle applied to the unit vectors equals the dense V-cycle
    operator built from matrix products (two- and three-level hierarchies,
    both correction variants) to FP64 rounding; a dropped term, a swapped
    pre / post order or an aliasing slip changes entries by O(1)."""
    import gpusolver as gs
    if problem == "sgs":
        A, block, omega, kind = gs.poisson2d(12, 12), 1, 1.2, 1
        H = gs.Hierarchy.smoothed_aggregation(A, min_dofs=10, max_levels=levels - 1)
    else:
        A, _, _ = gs.elasticity(4, young=1.0, poisson=0.3)
        block, omega, kind = 3, 0.7, 3
        H = gs.Hierarchy.smoothed_aggregation(A, block_size=3, min_dofs=6, max_levels=levels - 1)
    assert H.levels == levels
    Al = [H.get("A", k) for k in range(levels)]
    Pl = [H.get("P", k) for k in range(levels - 1)]
    Rl = [H.get("R", k) for k in range(levels - 1)]
    hv = orc.HierView([(a.ptr, a.col, a.val) for a in Al], [(p.ptr, p.col, p.val) for p in Pl],
                      [(r.ptr, r.col, r.val) for r in Rl])
    M = _vcycle_closed_form([a.to_scipy().toarray() for a in Al],
                            [p.to_scipy().toarray() for p in Pl],
                            [r.to_scipy().toarray() for r in Rl], block, omega, 2, aliasing)
    n = A.rows
    Mo = np.stack([orc.vcycle(hv, e, kinds=kind, omegas=omega, coarse_steps=2,
                              reference_vcycle=aliasing) for e in np.eye(n)], axis=1)
    scale = np.abs(M).max()
    assert np.abs(Mo - M).max() <= 1e-11 * scale
    if levels == 3:
        # The two variants differ on three levels, so the check tells them apart.
        Mother = _vcycle_closed_form([a.to_scipy().toarray() for a in Al],
                                     [p.to_scipy().toarray() for p in Pl],
                                     [r.to_scipy().toarray() for r in Rl], block, omega, 2,
                                     not aliasing)
        assert np.abs(Mo - Mother).max() > 1e-6 * scale


def test_oracle_caller_coarse_inverse_matches_own():
    """A caller-supplied coarsest inverse (LAPACK) gives the oracle the same
    PCG iterations and solution, to rounding, as its own Cholesky inverse."""
    import gpusolver as gs
    A = gs.poisson2d(40, 40)
    H = gs.Hierarchy.smoothed_aggregation(A, min_dofs=100)
    L = H.levels
    hv = orc.HierView([(a.ptr, a.col, a.val) for a in (H.get("A", k) for k in range(L))],
                      [(p.ptr, p.col, p.val) for p in (H.get("P", k) for k in range(L - 1))],
                      [(r.ptr, r.col, r.val) for r in (H.get("R", k) for k in range(L - 1))])
    b = np.cos(0.1 * np.arange(A.rows))
    inv = np.linalg.inv(H.get("A", L - 1).to_scipy().toarray())
    for mg_double in (True, False):
        x1, it1, _, _ = orc.pcg(hv, b, mg_double=mg_double, direct_coarse=True)
        x2, it2, _, _ = orc.pcg(hv, b, mg_double=mg_double, direct_coarse=True, coarse_inv=inv)
        assert it1 == it2
        assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x1)
    with pytest.raises(ValueError):
        orc.pcg(hv, b, direct_coarse=True, coarse_inv=np.eye(3))
