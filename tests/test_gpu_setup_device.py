"""GPU side of the smoothed-aggregation setup against the host, bit for bit:
the sparse products (gs_csr_spgemm, gs_csr_transpose with on_device = 1)
and whole hierarchies (gs_hierarchy_build_sa_device). The host products are
checked against scipy on the CPU (test_setup_artifacts.py). GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs

pytestmark = pytest.mark.gpu


def same(a: gs.Csr, b: gs.Csr):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a.ptr, b.ptr)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.val.view(np.uint64), b.val.view(np.uint64))


def rand_csr(m, n, density, seed, empty_rows=()):
    rng = np.random.default_rng(seed)
    M = sp.random(m, n, density=density, format="lil", random_state=seed,
                  data_rvs=lambda k: rng.uniform(-1, 1, k))
    for r in empty_rows:
        M[r, :] = 0
    M = M.tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    return gs.Csr.from_scipy(M)


CASES = [
    ("square", (300, 300, 0.02), (300, 300, 0.02)),
    ("rectangular", (500, 120, 0.03), (120, 80, 0.05)),
    ("empty rows", (200, 150, 0.02), (150, 200, 0.03)),
    ("long rows", (40, 2000, 0.5), (2000, 3000, 0.02)),  # ~20 000 products per row
    # Distinct columns per output row in (128, 256] and (256, 1024]: the
    # wavefront-per-row kernels' middle and large LDS tables.
    ("mid rows", (120, 400, 0.075), (400, 300, 0.03)),
    ("wide rows", (120, 400, 0.075), (400, 800, 0.03)),
]


@pytest.mark.parametrize("name,sa,sb", CASES, ids=[c[0] for c in CASES])
def test_device_spgemm_and_transpose_match_host(name, sa, sb):
    A = rand_csr(*sa, seed=1, empty_rows=(0, 5, 17) if name == "empty rows" else ())
    B = rand_csr(*sb, seed=2, empty_rows=(3, 4) if name == "empty rows" else ())
    C_host = gs.spgemm(A, B)
    same(gs.spgemm(A, B, device=True), C_host)
    same(gs.transpose(A, device=True), gs.transpose(A))
    same(gs.transpose(C_host, device=True), gs.transpose(C_host))


def test_device_spgemm_signed_zero_first_product():
    """A column whose first product is -0.0 keeps it (the sum starts from
    the first product, not from +0), as on the host."""
    A = gs.Csr.from_scipy(sp.csr_matrix(np.array([[-1.0, 2.0]])))
    B = gs.Csr(np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32),
               np.array([0.0, -0.0]), (2, 1))
    C = gs.spgemm(A, B, device=True)
    same(C, gs.spgemm(A, B))
    assert np.signbit(C.val[0])  # (-1 * 0) + (2 * -0) = -0 + -0 = -0


def test_device_spgemm_row_classes_are_covered():
    """The CASES above reach every path of the device product: rows of at
    most 128, 256 and 1024 distinct columns (the LDS kernels) and longer
    ones (the global-table kernels, after the LDS count gives up)."""
    seen = set()
    for name, sa, sb in CASES:
        A = rand_csr(*sa, seed=1, empty_rows=(0, 5, 17) if name == "empty rows" else ())
        B = rand_csr(*sb, seed=2, empty_rows=(3, 4) if name == "empty rows" else ())
        C = gs.spgemm(A, B)
        for m in np.diff(C.ptr):
            seen.add(0 if m == 0 else 1 if m <= 128 else 2 if m <= 256 else 3 if m <= 1024 else 4)
    assert seen == {0, 1, 2, 3, 4}


def test_device_spgemm_repeated_columns_in_b():
    """A B row that names a column twice (the products are added in their
    stored order, as on the host): the device takes the sequential path."""
    A = gs.Csr.from_scipy(sp.csr_matrix(np.array([[1.5, -2.0, 0.25], [0.0, 3.0, 1.0]])))
    B = gs.Csr(np.array([0, 3, 5, 6], np.int32), np.array([1, 0, 1, 2, 2, 0], np.int32),
               np.array([1e16, 1.0, -1e16, 3.0, 5.0, 7.0]), (3, 3))
    same(gs.spgemm(A, B, device=True), gs.spgemm(A, B))


def _hier_same(Hd, Hh):
    assert Hd.levels == Hh.levels
    for k in range(Hh.levels):
        same(Hd.get("A", k), Hh.get("A", k))
    for k in range(Hh.levels - 1):
        same(Hd.get("P", k), Hh.get("P", k))
        same(Hd.get("R", k), Hh.get("R", k))
        np.testing.assert_array_equal(Hd.aggregates(k), Hh.aggregates(k))


@pytest.mark.parametrize("problem", ["elasticity", "elasticity_nullspace", "varcoef", "poisson2d"])
def test_device_hierarchy_matches_host(problem):
    kw = {}
    if problem.startswith("elasticity"):
        A, _, ns = gs.elasticity(16, young=1.0, poisson=0.3, hetero_sigma=2.0, hetero_scale=1.5,
                                 seed=3)
        kw = dict(block_size=3, min_dofs=300)
        if problem.endswith("nullspace"):
            kw["nullspace"] = ns
    elif problem == "varcoef":
        A = gs.varcoef3d(40, seed=4, sigma=3.0, scale=2.0)
        kw = dict(block_size=1, min_dofs=200)
    else:
        A = gs.poisson2d(100, 100)
        kw = dict(block_size=1, min_dofs=100)
    Hh = gs.Hierarchy.smoothed_aggregation(A, **kw)
    Hd = gs.Hierarchy.smoothed_aggregation(A, device=True, **kw)
    assert Hh.levels >= 3
    _hier_same(Hd, Hh)


@pytest.fixture(scope="module")
def elast_h():
    E, rhs, _ = gs.elasticity(14, young=1.0, poisson=0.3, hetero_sigma=1.5, hetero_scale=1.5,
                              seed=5)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=300)
    return E, rhs, H


OPTS = [dict(), dict(group_min_block=0.0), dict(group_min_block=-1.0), dict(cg_pair_layout=0),
        dict(cg_level_order=0), dict(zero_start_lower=10), dict(smoother=gs.BLOCK_JACOBI),
        dict(coarse_direct=1, direct_truncate=1, max_direct=600)]


@pytest.mark.parametrize("opts", OPTS, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()) or "default")
@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED, gs.PREC_SINGLE])
def test_device_pack_matches_host_pack(elast_h, prec, opts):
    """3x3-block operators packed on the GPU (device_pack = 1) give the same
    SpMV, V-cycle and PCG solve, bit for bit, as the host packers (0):
    SELL and lane-group storage, value pairs or planes, level-order row
    maps, earlier-colour operators on every level, block Jacobi, a direct
    coarse level."""
    E, rhs, H = elast_h
    kw = dict(smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1)
    kw.update(opts)
    x = gs.uniform_pm1(21, E.rows)
    out = []
    for dp in (0, 1):
        s = gs.Solver(prec, device_pack=dp, **kw)
        s.setup(H)
        xs, info = s.solve(rhs, tol=1e-8, maxit=500)
        out.append((s.spmv(x), s.vcycle(x), xs, info.iterations, s.describe()["levels"]))
    (y0, z0, x0, i0, l0), (y1, z1, x1, i1, l1) = out
    assert np.array_equal(y0.view(np.uint64), y1.view(np.uint64))
    assert np.array_equal(z0.view(np.uint64), z1.view(np.uint64))
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64)) and i0 == i1
    assert l0 == l1


def test_device_pack_with_level_permutations(elast_h):
    """Block-preserving level permutations (block RCM per level): the CG
    operator is then uploaded and packed on its own; same results as the
    host packers bit for bit."""
    E, rhs, H = elast_h
    perms = []
    for k in range(H.levels):
        Ak = H.get("A", k)
        gp, gc = gs.block_graph(Ak, 3)
        G = gs.Csr(gp, gc, np.ones(len(gc)), (len(gp) - 1, len(gp) - 1))
        node_o2n, _ = gs.rcm(G)
        perms.append(gs.expand_block_permutation(node_o2n, 3)[0])
    x = gs.uniform_pm1(22, E.rows)
    out = []
    for dp in (0, 1):
        s = gs.Solver(gs.PREC_MIXED, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=2, cg_block3=1,
                      device_pack=dp)
        s.setup(H, perms)
        xs, info = s.solve(rhs, tol=1e-8, maxit=500)
        out.append((s.spmv(x), s.vcycle(x), xs, info.iterations))
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert out[0][3] == out[1][3]
