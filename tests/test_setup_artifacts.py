"""Setup artifacts must match the reference bit for bit: colouring, RCM,
permutations, ELL packing, both permutation flavours, block graphs. CPU only.

Pinned by tests/golden/setup_fixtures.json (written by the reference's own
GraphColoring.cc / RCMReordering.cc via tests/golden/make_golden.py), by the
live reference build in oracle/_ref when present, and by the invariants of
tests/metal/graphColoringTest.cc:28-201 with its exact colour counts
(:222-505).
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import gpusolver as gs
import oracle_api as orc
from graphs import small_graphs, large_patterns

FIX = json.loads((Path(__file__).parent / "golden" / "setup_fixtures.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


def as_csr(ptr, col):
    return gs.Csr(ptr, col, np.ones(len(col)), (len(ptr) - 1, len(ptr) - 1))


def assert_valid_coloring(ptr, col, colors):
    n = len(ptr) - 1
    for i in range(n):
        for k in range(ptr[i], ptr[i + 1]):
            j = col[k]
            if j != i:
                assert colors[i] != colors[j]


@pytest.mark.parametrize("name", sorted(FIX["small"]))
def test_coloring_small_matches_reference(name):
    f = FIX["small"][name]
    ptr = np.array(f["ptr"], np.int32)
    col = np.array(f["col"], np.int32)
    for impl in (gs.graph_coloring(as_csr(ptr, col)), orc.coloring(ptr, col)):
        colors, verts, offs, nc = impl
        assert nc == f["ncolors"]
        assert colors.tolist() == f["colors"]
        assert verts.tolist() == f["color_vertices"]
        assert offs.tolist() == f["color_offsets"]
    assert_valid_coloring(ptr, col, np.array(f["colors"]))


def test_reference_color_counts():
    """Exact counts of tests/metal/graphColoringTest.cc:222-505."""
    expect = {"path5": 2, "triangle": 3, "cycle4": 2, "k5": 5, "grid10x10": 2}
    for k, v in expect.items():
        assert FIX["small"][k]["ncolors"] == v


@pytest.mark.parametrize("name", sorted(FIX["small"]))
def test_rcm_small_matches_reference(name):
    f = FIX["small"][name]
    ptr = np.array(f["ptr"], np.int32)
    col = np.array(f["col"], np.int32)
    for o2n, n2o in (gs.rcm(as_csr(ptr, col)), orc.rcm(ptr, col)):
        assert o2n.tolist() == f["rcm_old_to_new"]
        assert n2o.tolist() == f["rcm_new_to_old"]


def test_large_patterns_match_reference_digests():
    pats = large_patterns()
    for name, f in FIX["large"].items():
        ptr, col = pats[name]
        assert len(ptr) - 1 == f["n"] and len(col) == f["nnz"]
        colors, verts, offs, nc = gs.graph_coloring(as_csr(ptr, col))
        assert nc == f["ncolors"]
        assert sha(colors) == f["colors_sha256"]
        assert sha(verts) == f["color_vertices_sha256"]
        assert offs.tolist() == f["color_offsets"]
        if "rcm_old_to_new_sha256" in f:
            assert sha(gs.rcm(as_csr(ptr, col))[0]) == f["rcm_old_to_new_sha256"]


def test_live_reference_build_agrees():
    L = orc.ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for name, (ptr, col) in small_graphs().items():
        a = orc.ref_coloring(L, ptr, col)
        b = gs.graph_coloring(as_csr(ptr, col))
        for x, y in zip(a[:3], b[:3]):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(orc.ref_rcm(L, ptr, col)[0], gs.rcm(as_csr(ptr, col))[0])


def test_permutation_validation():
    gs.permutation_check([2, 0, 1], [1, 2, 0])
    with pytest.raises(gs.GpuSolverError, match="not inverses"):
        gs.permutation_check([2, 0, 1], [0, 1, 2])
    with pytest.raises(gs.GpuSolverError, match="invalid index"):
        gs.permutation_check([3, 0, 1], [1, 2, 0])
    L = orc.ref_lib()
    if L is not None:
        ok = np.array([2, 0, 1], np.int32), np.array([1, 2, 0], np.int32)
        bad = np.array([2, 0, 1], np.int32), np.array([0, 1, 2], np.int32)
        assert L.ref_permutation_valid(3, *ok) == 0
        assert L.ref_permutation_valid(3, *bad) != 0


def _random_sym(n, seed, density=0.08):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    m = sp.random(n, n, density=density, random_state=rng, format="csr")
    m = m + m.T + sp.eye(n) * 5
    m = m.tocsr()
    m.sort_indices()
    return gs.Csr.from_scipy(m)


@pytest.mark.parametrize("seed", [0, 1])
def test_permute_keep_order_matches_oracle(seed):
    A = _random_sym(80, seed)
    o2n = np.random.default_rng(seed).permutation(A.rows).astype(np.int32)
    B = gs.permute_keep_order(A, o2n)
    optr, ocol, oval = orc.permute_keep_order(A.ptr, A.col, A.val, o2n)
    np.testing.assert_array_equal(B.ptr, optr)
    np.testing.assert_array_equal(B.col, ocol)
    np.testing.assert_array_equal(B.val, oval)
    # P A P^T as a matrix
    S = A.to_scipy().toarray()
    n2o = np.argsort(o2n)
    np.testing.assert_array_equal(B.to_scipy().toarray(), S[n2o][:, n2o])


@pytest.mark.parametrize("seed", [2, 3])
def test_permute_sorted_matches_oracle(seed):
    A = _random_sym(70, seed)
    rng = np.random.default_rng(seed)
    ro, co = rng.permutation(A.rows).astype(np.int32), rng.permutation(A.rows).astype(np.int32)
    B = gs.permute_sorted(A, ro, co)
    optr, ocol, oval = orc.permute_sorted(A.ptr, A.col, A.val, ro, co, A.rows)
    np.testing.assert_array_equal(B.ptr, optr)
    np.testing.assert_array_equal(B.col, ocol)
    np.testing.assert_array_equal(B.val, oval)
    for r in range(B.rows):
        assert np.all(np.diff(B.col[B.ptr[r]:B.ptr[r + 1]]) > 0)


def test_permute_sorted_sums_duplicates():
    # Two stored entries that land on the same (row, col) after relabelling
    # cannot arise from a permutation, so duplicates come from the input.
    A = gs.Csr(np.array([0, 2, 3], np.int32), np.array([1, 1, 0], np.int32),
               np.array([1.5, 2.5, 4.0]), (2, 2))
    B = gs.permute_sorted(A, np.array([0, 1], np.int32), np.array([0, 1], np.int32))
    assert B.ptr.tolist() == [0, 1, 2] and B.val.tolist() == [4.0, 4.0]


@pytest.mark.parametrize("width", [1, 3, 5, 8])
def test_ell_pack_matches_oracle(width):
    A = _random_sym(50, 7, density=0.1)
    ours = gs.ell_pack(A, width)
    theirs = orc.ell_pack(A.ptr, A.col, A.val, width)
    for a, b in zip(ours, theirs):
        np.testing.assert_array_equal(a, b)
    ec, ev, op, oc, ov = ours
    n = A.rows
    lens = np.diff(A.ptr)
    assert len(oc) == np.maximum(lens - width, 0).sum()
    assert np.all(np.diff(op) >= 0)
    # slot-major layout, padding (0, 0.0)
    for r in range(n):
        m = min(width, lens[r])
        for s in range(width):
            if s < m:
                assert ec[s * n + r] == A.col[A.ptr[r] + s] and ev[s * n + r] == A.val[A.ptr[r] + s]
            else:
                assert ec[s * n + r] == 0 and ev[s * n + r] == 0.0
    # ELL + overflow reproduces the CSR product bit for bit (same order)
    x = np.random.default_rng(1).uniform(-1, 1, n)
    y = np.zeros(n)
    for r in range(n):
        s = 0.0
        for k in range(min(width, lens[r])):
            s += ev[k * n + r] * x[ec[k * n + r]]
        for k in range(op[r], op[r + 1]):
            s += ov[k] * x[oc[k]]
        y[r] = s
    np.testing.assert_array_equal(y, orc.spmv(A.ptr, A.col, A.val, x))


def test_ell_width_zero_raises():
    A = _random_sym(10, 1)
    with pytest.raises(gs.GpuSolverError, match="ELL width must be nonzero"):
        gs.ell_pack(A, 0)


def test_block_graph_matches_oracle():
    E, _, _ = gs.elasticity(5)
    gp, gc = gs.block_graph(E, 3)
    op, oc = orc.block_graph(E.ptr, E.col, 3)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gc, oc)
    # Symmetric, no self loops, sorted.
    nb = E.rows // 3
    edges = {(i, j) for i in range(nb) for j in gc[gp[i]:gp[i + 1]]}
    assert all((j, i) in edges for i, j in edges)
    assert all(i != j for i, j in edges)


@pytest.mark.parametrize("seed", [0, 1])
def test_block_graph_nonsymmetric_pattern(seed):
    """A one-sided coupling still yields the edge in both directions
    (src/BlockUtils.cc:126-216); checked against the oracle and against a
    set-based construction."""
    rng = np.random.default_rng(seed)
    n = 3 * 97
    rows = [sorted(set(rng.integers(0, n, rng.integers(0, 9)).tolist()) | {i}) for i in range(n)]
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    gp, gc = gs.block_graph(as_csr(ptr, col), 3)
    op, oc = orc.block_graph(ptr, col, 3)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gc, oc)
    adj = [set() for _ in range(n // 3)]
    for i, r in enumerate(rows):
        for c in r:
            if i // 3 != c // 3:
                adj[i // 3].add(c // 3)
                adj[c // 3].add(i // 3)
    assert gp.tolist() == np.cumsum([0] + [len(a) for a in adj]).tolist()
    assert gc.tolist() == [c for a in adj for c in sorted(a)]


def test_expand_block_permutation():
    node = np.array([2, 0, 1], np.int32)
    o2n, n2o = gs.expand_block_permutation(node, 3)
    assert o2n.tolist() == [6, 7, 8, 0, 1, 2, 3, 4, 5]
    assert n2o.tolist() == [3, 4, 5, 6, 7, 8, 0, 1, 2]


def test_inverse_diag_blocks3_accuracy():
    E, _, _ = gs.elasticity(4)
    inv = gs.inverse_diag_blocks3(E).reshape(-1, 3, 3)
    D = E.to_scipy().toarray()
    for b in range(E.rows // 3):
        blk = D[3 * b:3 * b + 3, 3 * b:3 * b + 3]
        np.testing.assert_allclose(blk @ inv[b], np.eye(3), atol=1e-12 * np.linalg.cond(blk))


def test_inverse_diag_blocks3_not_spd_raises():
    A = gs.Csr.from_scipy(__import__("scipy.sparse", fromlist=["x"]).csr_matrix(
        np.diag([1.0, -1.0, 1.0])))
    with pytest.raises(gs.GpuSolverError, match="not SPD"):
        gs.inverse_diag_blocks3(A)


def test_sell_layout_in_bounds_for_every_operator():
    """Every stored column of every device operator (padding included) lies
    inside [0, cols): the prolongators of a Poisson hierarchy have empty
    rows (strength-isolated coarse nodes) and rectangular shapes."""
    A = gs.poisson2d(256, 256)
    H = gs.Hierarchy.smoothed_aggregation(A, min_dofs=1000)
    ops = [H.get("A", k) for k in range(H.levels)]
    ops += [H.get(w, k) for w in ("P", "R") for k in range(H.levels - 1)]
    # A tall prolongation-like operator with empty rows (strength-isolated
    # fine nodes get none) whose row indices exceed the column count.
    rows = np.arange(300)
    lens = np.where(rows % 7 == 0, 0, 2)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.array([c for r in rows for c in ((r % 10, (r + 3) % 10) if lens[r] else ())], np.int32)
    ops.append(gs.Csr(ptr, col, np.ones(col.size), (300, 10)))
    for M in ops:
        slots, lo, hi = gs.sell_layout(M)
        assert slots >= M.nnz and 0 <= lo and hi < M.shape[1]


def test_sell_layout_diagonal_only_rows():
    import scipy.sparse as sp
    D = gs.Csr.from_scipy(sp.csr_matrix(np.diag(np.arange(1.0, 70.0))))
    slots, lo, hi = gs.sell_layout(D)
    assert slots >= D.nnz and 0 <= lo and hi < 69
    # Mixed lengths force padding; a row whose only entry is its diagonal
    # must be padded with another column.
    M = sp.lil_matrix((70, 70))
    for i in range(70):
        M[i, i] = 2.0
        if i % 3 == 0 and i + 1 < 70:
            M[i, i + 1] = -1.0
    slots, lo, hi = gs.sell_layout(gs.Csr.from_scipy(M.tocsr()))
    assert slots > M.nnz and hi < 70


def test_host_spgemm_and_transpose_match_scipy():
    """gs_csr_spgemm / gs_csr_transpose on the host (the products of the
    smoothed-aggregation setup): structure equal to scipy's, values equal to
    rounding, rows sorted; rectangular shapes and empty rows included."""
    import scipy.sparse as sp
    A = sp.random(60, 40, density=0.1, format="lil", random_state=1)
    A[7, :] = 0
    A = A.tocsr()
    B = sp.random(40, 30, density=0.15, format="csr", random_state=2)
    for M in (A, B):
        M.sort_indices()
    a, b = gs.Csr.from_scipy(A), gs.Csr.from_scipy(B)
    C = gs.spgemm(a, b).to_scipy()
    ref = (A @ B).tocsr()
    ref.sort_indices()
    assert (abs(C - ref) > 1e-14).nnz == 0
    assert all(np.all(np.diff(C.indices[C.indptr[i]:C.indptr[i + 1]]) > 0) for i in range(60))
    T = gs.transpose(a).to_scipy()
    np.testing.assert_array_equal(T.toarray(), A.T.toarray())
    assert T.shape == (40, 60)
