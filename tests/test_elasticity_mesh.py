"""The cantilever driver's elasticity system on a caller's tet mesh
(gs_problem_elasticity_mesh, examples/elasticity/cantileverBeam.cc:275-348)
and the MEDIT .mesh reader. CPU only.

Pins: on the Kuhn-split unit cube the general assembly reproduces the
separately written structured generator (same pattern, same values to
rounding); rigid-body motions produce no force away from the clamp; the
element matrices scale exactly with E; the lumped mass of an interior vertex
is h^3 (24 tets of volume h^3 / 6, a quarter each); inverted tets add no
mass, as LinearElasticity.ipp:31-47 skips them.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs

PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def kuhn_mesh(nx, ny=None, nz=None, h=None):
    """Box of nx x ny x nz vertices, 6 Kuhn tets per hex (corner 000, then
    the axes in permutation order; odd permutations with vertices 1 and 2
    swapped so every tet is positively oriented); vertex id
    (z * ny + y) * nx + x."""
    ny = ny or nx
    nz = nz or nx
    h = h or 1.0 / (nx - 1)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    V = np.stack([x.ravel() * h, y.ravel() * h, z.ravel() * h], axis=1)
    vid = lambda x, y, z: (z * ny + y) * nx + x  # noqa: E731
    T = []
    for cz in range(nz - 1):
        for cy in range(ny - 1):
            for cx in range(nx - 1):
                for k, p in enumerate(PERMS):
                    c = [0, 0, 0]
                    tet = [vid(cx, cy, cz)]
                    for ax in p:
                        c[ax] = 1
                        tet.append(vid(cx + c[0], cy + c[1], cz + c[2]))
                    if k in (1, 2, 5):  # odd permutations
                        tet[1], tet[2] = tet[2], tet[1]
                    T.append(tet)
    return V, np.array(T, np.int32)


def test_matches_structured_generator():
    nv = 6
    V, T = kuhn_mesh(nv)
    A, rhs, ns = gs.elasticity_mesh(V, T, young=1.0, poisson=0.3)
    B, _, _ = gs.elasticity(nv, young=1.0, poisson=0.3)
    np.testing.assert_array_equal(A.ptr, B.ptr)
    np.testing.assert_array_equal(A.col, B.col)
    np.testing.assert_allclose(A.val, B.val, rtol=0, atol=1e-14 * np.abs(B.val).max())


def test_rigid_modes_and_symmetry():
    V, T = kuhn_mesh(7, 5, 4, h=0.25)
    A, rhs, ns = gs.elasticity_mesh(V, T)
    S = A.to_scipy()
    assert abs(S - S.T).max() <= 1e-12 * abs(S).max()
    # Rows whose vertex has no clamped neighbour: rigid motions give no force.
    clamped = V[:, 0] <= V[:, 0].min() + 1e-4 * np.ptp(V[:, 0])
    free = np.flatnonzero(~clamped)
    near = set()
    for t in T:
        if clamped[t].any():
            near.update(int(v) for v in t)
    far = [3 * f + k for f, v in enumerate(free) if v not in near for k in range(3)]
    assert len(far) > 0
    R = S @ ns
    assert abs(R[far]).max() <= 1e-9 * abs(S).max()


def test_young_scales_exactly_and_per_tet():
    V, T = kuhn_mesh(4)
    A1, _, _ = gs.elasticity_mesh(V, T, young=1.0)
    A2, _, _ = gs.elasticity_mesh(V, T, young=2.0)
    A3, _, _ = gs.elasticity_mesh(V, T, young=7.0, young_per_tet=np.full(len(T), 2.0))
    np.testing.assert_array_equal(A2.val, 2.0 * A1.val)
    np.testing.assert_array_equal(A3.val, A2.val)


def test_orientation():
    V, T = kuhn_mesh(3)
    X = V[T]
    det = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]],
                                 axis=2))
    assert (det > 0).all()


def test_lumped_mass_load():
    nv = 6
    h = 1.0 / (nv - 1)
    V, T = kuhn_mesh(nv)
    A, rhs, _ = gs.elasticity_mesh(V, T, density=2.0, gravity=-9.81)
    free = np.flatnonzero(V[:, 0] > 1e-4)
    interior = [f for f, v in enumerate(free)
                if all(0 < V[v, d] < 1 for d in range(3))]
    assert interior
    for f in interior:
        assert rhs[3 * f] == 0 and rhs[3 * f + 1] == 0
        assert rhs[3 * f + 2] == pytest.approx(2.0 * -9.81 * h ** 3, rel=1e-12)
    # Every tet inverted (vertices 1 and 2 swapped): same stiffness, no mass.
    Ti = T[:, [0, 2, 1, 3]]
    Ai, rhsi, _ = gs.elasticity_mesh(V, Ti, density=2.0)
    np.testing.assert_allclose(Ai.val, A.val, rtol=0, atol=1e-12 * np.abs(A.val).max())
    assert not rhsi.any()


def test_errors():
    V, T = kuhn_mesh(3)
    with pytest.raises(gs.GpuSolverError, match="No vertices found on clamped side."):
        gs.elasticity_mesh(V, T, clamp_rel=-1.0)
    bad = T.copy()
    bad[0, 0] = len(V)
    with pytest.raises(gs.GpuSolverError, match="vertex index out of range"):
        gs.elasticity_mesh(V, bad)
    flat = T.copy()
    flat[0] = [0, 1, 2, 2]
    with pytest.raises(gs.GpuSolverError, match="degenerate tetrahedron"):
        gs.elasticity_mesh(V, flat)


def test_medit_round_trip_and_parse(tmp_path):
    V, T = kuhn_mesh(3, h=0.1)
    F = np.array([[0, 1, 3]], np.int32)
    gs.write_mesh(tmp_path / "m.mesh", V, T, F)
    V2, T2, F2 = gs.read_mesh(tmp_path / "m.mesh")
    np.testing.assert_array_equal(V2, V)
    np.testing.assert_array_equal(T2, T)
    np.testing.assert_array_equal(F2, F)
    (tmp_path / "h.mesh").write_text(
        "MeshVersionFormatted 1\n# a comment\nDimension 3\nVertices\n4\n"
        "0 0 0 1\n1 0 0 1\n0 1 0 1\n0 0 1 1\nTetrahedra\n1\n1 2 3 4 0\nEnd\n")
    V3, T3, F3 = gs.read_mesh(tmp_path / "h.mesh")
    assert V3.shape == (4, 3) and T3.tolist() == [[0, 1, 2, 3]] and F3.shape == (0, 3)
    (tmp_path / "t.mesh").write_text("Vertices\n3\n0 0 0 1\n")
    with pytest.raises(ValueError, match="truncated"):
        gs.read_mesh(tmp_path / "t.mesh")
