"""HierarchicalReordering (src/ReorderingStrategies/HierarchicalReordering.cc:
13-278) must match the reference bit for bit. CPU only.

The reference needs Eigen (AMGHierarchy's matrices), so it does not build
here. The oracle restates it (std::sort on the (parent rank, RCM rank) key,
as the reference); the product computes the same permutation by a parallel
arg-max and a counting sort. Both are pinned by hand-derived answers on small
hierarchies that exercise every rule: the first strongest entry wins a tie,
a later larger entry wins, weights compare in FP32 (a double-precision
comparison picks another parent in the first case), NaN entries and empty
columns leave a DOF without a parent and put it last, and each finer level
follows the coarse level's final (not its RCM) order. A single-level
hierarchy reduces to RCM, which the compiled reference pins.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs
import oracle_api as orc


def path_matrix(order, n):
    """Symmetric matrix with unit diagonal and the path order[0]-order[1]-...;
    columns sorted within each row."""
    m = sp.lil_matrix((n, n))
    for i in range(n):
        m[i, i] = 2.0
    for a, b in zip(order[:-1], order[1:]):
        m[a, b] = m[b, a] = -1.0
    return sp.csr_matrix(m)


def csr_rows(shape, rows):
    """CSR with the given per-row (col, value) lists kept in that order."""
    ptr, col, val = [0], [], []
    for r in rows:
        for c, v in r:
            col.append(c)
            val.append(v)
        ptr.append(len(col))
    return (np.array(ptr, np.int32), np.array(col, np.int32), np.array(val, np.float64), shape)


def gs_csr(t):
    ptr, col, val, shape = t
    return gs.Csr(ptr, col, val, shape)


def transpose(t):
    ptr, col, val, (r, c) = t
    m = sp.csr_matrix((val, col, ptr), shape=(r, c)).T.tocsr()
    return (m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64), (c, r))


def from_scipy(m):
    m = sp.csr_matrix(m)
    return (m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64),
            m.shape)


# Fine level: the path 2-0-4-1-5-3. Degrees 1 at 2 and 3, so RCM starts at 2,
# Cuthill-McKee order 2,0,4,1,5,3, reversed: new_to_old = 3,5,1,4,0,2.
A0 = from_scipy(path_matrix([2, 0, 4, 1, 5, 3], 6))
# Coarse level: the path 1-0-2; RCM starts at 1, order 1,0,2, reversed:
# new_to_old = 2,0,1, old_to_new = 1,2,0.
A1 = from_scipy(path_matrix([1, 0, 2], 3))
# R0 (3 x 6). Fine 0: 0.1 at row 0 ties 0.10000000001 at row 2 in FP32 (row 0
# wins; in FP64 row 2 would). Fine 1: 0.25 at row 0, then 0.5 at row 1 (row 1
# wins). Fine 2: only NaN (no parent). Fine 3: 0.3 at rows 1 and 2 (row 1).
# Fine 4: 0.5 at rows 0 and 2 (row 0). Fine 5: |-0.75| at row 1.
R0 = csr_rows((3, 6), [[(0, 0.1), (4, 0.5), (1, 0.25)],
                       [(1, 0.5), (5, -0.75), (3, 0.3)],
                       [(3, 0.3), (4, 0.5), (0, 0.10000000001), (2, float("nan"))]])
A2 = from_scipy(np.array([[1.0]]))
# R1 (1 x 3): coarse DOF 0 has no parent, 1 and 2 share the single parent.
R1 = csr_rows((1, 3), [[(1, 2.0), (2, 1.0)]])


def hier(As, Rs):
    return gs.Hierarchy.from_levels([gs_csr(a) for a in As], [gs_csr(transpose(r)) for r in Rs],
                                    [gs_csr(r) for r in Rs])


def oracle(As, Rs):
    return orc.hierarchical_reordering([(a[0], a[1]) for a in As], [(r[0], r[1], r[2]) for r in Rs])


def both(As, Rs):
    return gs.hierarchical_reordering(hier(As, Rs)), oracle(As, Rs)


def as_lists(perms):
    return [(o.tolist(), n.tolist()) for o, n in perms]


def test_two_levels_hand_derived():
    # Parents 0->0, 1->1, 2->none, 3->1, 4->0, 5->1; coarse ranks 0->1,
    # 1->2, 2->0. Groups: rank 1 {0, 4} in RCM order 4, 0; rank 2 {1, 3, 5}
    # in RCM order 3, 5, 1; then the orphan 2.
    expect = [([1, 4, 5, 2, 0, 3], [4, 0, 3, 5, 1, 2]), ([1, 2, 0], [2, 0, 1])]
    for perms, orphans in both([A0, A1], [R0]):
        assert as_lists(perms) == expect
        assert orphans.tolist() == [1]


def test_three_levels_follow_the_coarse_final_order():
    # Level 1: coarse 0 has no parent -> last; 1 and 2 in level-1 RCM order
    # (2 before 1): new_to_old 2,1,0. Level 0 then ranks coarse 1 before
    # coarse 0 (plain level-1 RCM would rank 0 first, as in the two-level
    # case): groups {1,3,5} -> 3,5,1, {0,4} -> 4,0, then 2.
    expect = [([4, 2, 5, 0, 3, 1], [3, 5, 1, 4, 0, 2]), ([2, 1, 0], [2, 1, 0]), ([0], [0])]
    for perms, orphans in both([A0, A1, A2], [R0, R1]):
        assert as_lists(perms) == expect
        assert orphans.tolist() == [1, 1]


def test_fp64_comparison_would_differ():
    """The FP32 tie on fine DOF 0 matters: giving row 2 a weight that is
    larger in FP32 too moves DOF 0 into coarse 2's group, which ranks first."""
    ptr, col, val, shape = R0
    val = val.copy()
    val[list(col[ptr[2]:ptr[3]]).index(0) + ptr[2]] = 0.2
    R0b = (ptr, col, val, shape)
    for perms, _ in both([A0, A1], [R0b]):
        assert perms[0][1].tolist() == [0, 4, 3, 5, 1, 2]


def test_single_level_is_rcm():
    A = gs.poisson2d(13, 11)
    ptr, col = A.ptr, A.col
    (perms, orph) = gs.hierarchical_reordering(gs.Hierarchy.from_levels([A], [], []))
    o2n, n2o = orc.rcm(ptr, col)
    assert perms[0][0].tolist() == o2n.tolist() and perms[0][1].tolist() == n2o.tolist()
    assert len(orph) == 0
    L = orc.ref_lib()
    if L is not None:
        ro, rn = orc.ref_rcm(L, ptr, col)
        assert rn.tolist() == n2o.tolist()


def hierarchy_arrays(H):
    As = [from_scipy(H.get("A", k).to_scipy()) for k in range(H.levels)]
    Rs = [from_scipy(H.get("R", k).to_scipy()) for k in range(H.levels - 1)]
    return As, Rs


def check_permutation(o2n, n2o):
    n = len(o2n)
    assert sorted(n2o.tolist()) == list(range(n))
    assert (o2n[n2o] == np.arange(n)).all()


@pytest.mark.parametrize("case", ["poisson2d", "elasticity"])
def test_sa_hierarchy_matches_oracle(case):
    if case == "poisson2d":
        A = gs.poisson2d(60, 50)
        H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                              min_dofs=50)
    else:
        E, _, ns = gs.elasticity(8)
        H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                              min_dofs=100, nullspace=ns)
    assert H.levels >= 3
    perms, orph = gs.hierarchical_reordering(H)
    As, Rs = hierarchy_arrays(H)
    operms, oorph = oracle(As, Rs)
    assert as_lists(perms) == as_lists(operms)
    assert orph.tolist() == oorph.tolist()
    for o2n, n2o in perms:
        check_permutation(o2n, n2o)


def test_ties_and_nans_at_scale_match_oracle():
    """Many FP32 ties, NaNs and empty columns across a parallel scan: the
    product's packed-key arg-max must keep the first strongest entry."""
    rng = np.random.default_rng(7)
    nf, nc = 6000, 700
    fine = sp.random(nf, nf, density=6 / nf, random_state=3, format="csr")
    fine = ((fine + fine.T) + sp.eye(nf)).tocsr()
    coarse = sp.random(nc, nc, density=5 / nc, random_state=4, format="csr")
    coarse = ((coarse + coarse.T) + sp.eye(nc)).tocsr()
    rows = []
    choices = np.array([0.5, -0.5, 0.25, 1.0, -1.0, 0.1, 0.10000000001, np.nan, 0.0, -0.0])
    for c in range(nc):
        cols = rng.choice(nf - 40, size=rng.integers(0, 30), replace=False)  # last 40 never hit
        rows.append([(int(j), float(rng.choice(choices))) for j in cols])
    R = csr_rows((nc, nf), rows)
    As, Rs = [from_scipy(fine), from_scipy(coarse)], [R]
    (perms, orph), (operms, oorph) = both(As, Rs)
    assert as_lists(perms) == as_lists(operms)
    assert orph.tolist() == oorph.tolist() and orph[0] >= 40
    check_permutation(*perms[0])


def test_errors():
    h = gs.Hierarchy.from_levels([], [], [])
    with pytest.raises(gs.GpuSolverError, match="HierarchicalReordering: empty hierarchy."):
        gs.hierarchical_reordering(h)
    with pytest.raises(gs.GpuSolverError, match="expected one restriction matrix between"):
        gs.hierarchical_reordering(gs.Hierarchy.from_levels([gs_csr(A0), gs_csr(A1)], [], []))
    rect = from_scipy(sp.random(6, 5, density=0.5, random_state=1))
    R_rect = csr_rows((3, 6), [[(0, 1.0)], [(1, 1.0)], [(2, 1.0)]])
    with pytest.raises(gs.GpuSolverError, match="fine matrix must be square"):
        gs.hierarchical_reordering(hier([rect, A1], [R_rect]))
    R_cols = csr_rows((3, 5), [[(0, 1.0)], [(1, 1.0)], [(2, 1.0)]])
    with pytest.raises(gs.GpuSolverError, match="restriction dimensions do not match fine matrix"):
        gs.hierarchical_reordering(hier([A0, A1], [R_cols]))
    R_rows = csr_rows((2, 6), [[(0, 1.0)], [(1, 1.0)]])
    with pytest.raises(gs.GpuSolverError, match="restriction rows do not match the coarse level"):
        gs.hierarchical_reordering(hier([A0, A1], [R_rows]))
