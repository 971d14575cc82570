"""A cantilever beam given as a tet mesh (the reference drivers' input
path): gs_problem_elasticity_mesh's system solved on the GPU with the
driver's preconditioner (SA(3, 0.1), multigrid block Gauss-Seidel 1/1,
omega 0.7, no coarsest-level sweeps, examples/elasticity/cantileverBeam.cc:
413-415,572-573) matches the oracle's PCG on the same hierarchy.

FP64: iterations within +-2 of the oracle. Mixed precision (FP32 V-cycle):
this weakly preconditioned beam needs 130 iterations in FP64 and ~154 with
the FP32 V-cycle, and there the count follows rounding: relative
perturbations of 1e-15 in b move the oracle's own mixed count over 151-155
on the CPU. The mixed bar is therefore the FP64-verified residual, the
solution against the FP64 oracle, and the count within 5% of the oracle's.
"""
import numpy as np
import pytest

import gpusolver as gs
import oracle_api as orc
from test_elasticity_mesh import kuhn_mesh

pytestmark = pytest.mark.gpu


def hier_view(H):
    L = H.levels
    A = [H.get("A", k) for k in range(L)]
    P = [H.get("P", k) for k in range(L - 1)]
    R = [H.get("R", k) for k in range(L - 1)]
    return orc.HierView([(a.ptr, a.col, a.val) for a in A], [(p.ptr, p.col, p.val) for p in P],
                        [(r.ptr, r.col, r.val) for r in R])


@pytest.mark.parametrize("prec", [gs.PREC_DOUBLE, gs.PREC_MIXED])
def test_cantilever_mesh_pcg(tmp_path, prec):
    V, T = kuhn_mesh(25, 7, 7, h=0.25)
    gs.write_mesh(tmp_path / "beam.mesh", V, T)
    V, T, _ = gs.read_mesh(tmp_path / "beam.mesh")
    A, b, _ = gs.elasticity_mesh(V, T)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=500)
    assert H.levels >= 2
    s = gs.Solver(prec, smoother=gs.BLOCK_GS, omega=0.7, coarse_steps=0)
    s.setup(H)
    x, info = s.solve(b, tol=1e-8, maxit=1000)
    xo, ito, _, _ = orc.pcg(hier_view(H), b, tol=1e-8, maxit=1000, kinds=gs.BLOCK_GS, omegas=0.7,
                            coarse_steps=0, mg_double=prec == gs.PREC_DOUBLE)
    assert info.converged
    if prec == gs.PREC_DOUBLE:
        assert abs(info.iterations - ito) <= 2, (info.iterations, ito)
    else:
        assert abs(info.iterations - ito) <= 0.05 * ito, (info.iterations, ito)
    assert info.true_relres <= 1.5e-8
    x64 = xo if prec == gs.PREC_DOUBLE else orc.pcg(
        hier_view(H), b, tol=1e-8, maxit=1000, kinds=gs.BLOCK_GS, omegas=0.7, coarse_steps=0)[0]
    assert np.linalg.norm(x - x64) <= 1e-6 * np.linalg.norm(x64)
    # The beam sags under gravity: the free end moves down.
    tip = np.flatnonzero(V[V[:, 0] > 1e-4 * np.ptp(V[:, 0]), 0] == V[:, 0].max())
    assert (x[3 * tip + 2] < 0).all()
