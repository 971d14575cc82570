"""Binary CSR and hierarchy files (gs_csr_save / gs_csr_load,
gs_hierarchy_save / gs_hierarchy_load). CPU only.

The layout is pinned byte for byte against a file assembled here with
struct.pack from the documented field list, so a writer on the reference
side (from Eigen's outerIndexPtr / innerIndexPtr / valuePtr) can produce it
without this library.
"""
import struct

import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs


def same_csr(a, b):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a.ptr, b.ptr)
    np.testing.assert_array_equal(a.col, b.col)
    assert a.val.tobytes() == b.val.tobytes()  # bit for bit, NaN and -0.0 included


def test_csr_layout_byte_for_byte(tmp_path):
    A = gs.Csr(np.array([0, 2, 3], np.int32), np.array([0, 1, 1], np.int32),
               np.array([4.0, -1.5, 2.0]), (2, 3))
    A.save(tmp_path / "a.bin")
    expect = (b"GSCSR001" + struct.pack("<qqq", 2, 3, 3) + struct.pack("<3i", 0, 2, 3)
              + struct.pack("<3i", 0, 1, 1) + struct.pack("<3d", 4.0, -1.5, 2.0))
    assert (tmp_path / "a.bin").read_bytes() == expect
    same_csr(gs.load_csr(tmp_path / "a.bin"), A)


def test_hierarchy_layout_byte_for_byte(tmp_path):
    A0 = gs.Csr(np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([2.0, 3.0]),
                (2, 2))
    A1 = gs.Csr(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([5.0]), (1, 1))
    P = gs.Csr(np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32), np.array([1.0, 1.0]),
               (2, 1))
    R = gs.Csr(np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, 1.0]),
               (1, 2))
    gs.Hierarchy.from_levels([A0, A1], [P], [R]).save(tmp_path / "h.bin")

    def rec(m):
        return (struct.pack("<qqq", m.shape[0], m.shape[1], m.nnz)
                + m.ptr.astype("<i4").tobytes() + m.col.astype("<i4").tobytes()
                + m.val.astype("<f8").tobytes())
    expect = (b"GSHIER01" + struct.pack("<qq", 2, 1) + rec(A0) + rec(A1) + rec(P) + rec(R)
              + struct.pack("<q", 0))
    assert (tmp_path / "h.bin").read_bytes() == expect


def test_csr_round_trip_special_values(tmp_path):
    m = sp.random(50, 40, density=0.1, random_state=2, format="csr")
    v = m.data.copy()
    v[:4] = [np.nan, np.inf, -0.0, -np.inf]
    A = gs.Csr(m.indptr.astype(np.int32), m.indices.astype(np.int32), v, m.shape)
    A.save(tmp_path / "m.bin")
    same_csr(gs.load_csr(tmp_path / "m.bin"), A)


def test_hierarchy_round_trip(tmp_path):
    E, _, ns = gs.elasticity(7)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=100, nullspace=ns)
    H.save(tmp_path / "h.bin")
    G = gs.Hierarchy.load(tmp_path / "h.bin")
    assert G.levels == H.levels >= 3
    for k in range(H.levels):
        same_csr(G.get("A", k), H.get("A", k))
    for k in range(H.levels - 1):
        same_csr(G.get("P", k), H.get("P", k))
        same_csr(G.get("R", k), H.get("R", k))
        np.testing.assert_array_equal(G.aggregates(k), H.aggregates(k))
    np.testing.assert_array_equal(G.features(3), H.features(3))


def test_load_errors(tmp_path):
    A = gs.poisson2d(6, 5)
    A.save(tmp_path / "a.bin")
    data = (tmp_path / "a.bin").read_bytes()
    with pytest.raises(gs.GpuSolverError, match="is not a hierarchy file"):
        gs.Hierarchy.load(tmp_path / "a.bin")
    (tmp_path / "t.bin").write_bytes(data[:-5])
    with pytest.raises(gs.GpuSolverError, match="truncated file"):
        gs.load_csr(tmp_path / "t.bin")
    (tmp_path / "x.bin").write_bytes(data + b"\0")
    with pytest.raises(gs.GpuSolverError, match="trailing bytes"):
        gs.load_csr(tmp_path / "x.bin")
    (tmp_path / "h.bin").write_bytes(b"GSCSR001" + struct.pack("<qqq", 1 << 40, 3, 3))
    with pytest.raises(gs.GpuSolverError, match="corrupt matrix header"):
        gs.load_csr(tmp_path / "h.bin")
    (tmp_path / "c.bin").write_bytes(b"GSCSR001" + struct.pack("<qqq", 1, 1, 1)
                                     + struct.pack("<2i", 0, 1) + struct.pack("<i", 7)
                                     + struct.pack("<d", 1.0))
    with pytest.raises(gs.GpuSolverError, match="column index out of bounds"):
        gs.load_csr(tmp_path / "c.bin")
    (tmp_path / "g.bin").write_bytes(b"GSHIER01" + struct.pack("<qq", 1 << 50, 0))
    with pytest.raises(gs.GpuSolverError, match="corrupt hierarchy header"):
        gs.Hierarchy.load(tmp_path / "g.bin")
    with pytest.raises(gs.GpuSolverError, match="cannot open"):
        gs.load_csr(tmp_path / "missing.bin")
