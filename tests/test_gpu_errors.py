"""Setup-time error paths of the solver, through the C ABI: each case raises
the reference's message (std::runtime_error text) before any kernel runs.

Messages: src/Smoothers/SymmetricGaussSeidelSmoother.cc:14-16 (omega),
src/Smoothers/BlockGaussSeidelSmoother.cc (omega, DOFs divisible by 3),
src/Preconditioners/MultigridGaussSeidelPreconditioner.cc (smoothing steps),
src/CPU/CPUBackend.cc:238-268 (zero diagonal),
src/Smoothers/DampedJacobiSmoother.cc (missing diagonal),
src/CGSolver.cc (square matrix), src/Smoothers/MultigridSmoother.cc:22-335
(transfer shapes); the dense direct-solve limit has no reference
counterpart (the coarse direct solve is new).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    """2D Poisson 24 x 24 (576 unknowns, not a multiple of 3) and its SA
    hierarchy."""
    A = gs.poisson2d(24, 24)
    H = gs.Hierarchy.smoothed_aggregation(A, block_size=1, eps_strong=0.1, max_levels=10,
                                          min_dofs=50)
    assert H.levels >= 2
    return A, H


def one_level(m):
    return gs.Hierarchy.from_levels([gs.Csr.from_scipy(sp.csr_matrix(m))], [], [])


def test_sgs_invalid_omega(small):
    A, H = small
    with pytest.raises(gs.GpuSolverError, match="SymmetricGaussSeidelSmoother: invalid omega"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, omega=2.5).setup(H)


def test_block_gs_invalid_omega():
    E, rhs, ns = gs.elasticity(4)
    H = gs.Hierarchy.smoothed_aggregation(E, block_size=3, eps_strong=0.1, max_levels=10,
                                          min_dofs=50, nullspace=ns)
    with pytest.raises(gs.GpuSolverError, match="BlockGaussSeidelSmoother: omega must be in"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, omega=0.0).setup(H)


def test_block_gs_dofs_not_multiple_of_3(small):
    A, H = small
    with pytest.raises(gs.GpuSolverError, match="number of DOFs must be divisible by 3"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.BLOCK_GS, omega=0.7).setup(H)


def test_sgs_needs_a_smoothing_step(small):
    A, H = small
    with pytest.raises(gs.GpuSolverError, match="at least one smoothing step"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, pre=0, post=0).setup(H)


def test_sgs_zero_diagonal():
    m = np.diag([4.0, 0.0, 4.0, 4.0])
    m[0, 1] = m[1, 0] = -1.0
    m[2, 3] = m[3, 2] = -1.0
    with pytest.raises(gs.GpuSolverError, match="zero diagonal entry"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=1).setup(one_level(m))


def test_damped_jacobi_missing_diagonal():
    m = sp.lil_matrix((3, 3))
    m[0, 0] = 2.0
    m[1, 2] = m[2, 1] = 1.0
    m[2, 2] = 2.0
    H = gs.Hierarchy.from_levels([gs.Csr.from_scipy(m.tocsr())], [], [])
    with pytest.raises(gs.GpuSolverError, match="missing diagonal entry"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.DAMPED_JACOBI, coarse_steps=1).setup(H)


def test_non_square_level():
    m = sp.random(6, 5, density=0.5, format="csr", random_state=1)
    H = gs.Hierarchy.from_levels([gs.Csr.from_scipy(m)], [], [])
    with pytest.raises(gs.GpuSolverError, match="matrix must be square"):
        gs.Solver(gs.PREC_DOUBLE).setup(H)


def test_transfer_shape_mismatch(small):
    A, H = small
    A0, A1 = H.get("A", 0), H.get("A", 1)
    P = H.get("P", 0).to_scipy()
    bad_P = gs.Csr.from_scipy(P[:, :-1])  # one coarse column short
    bad_R = gs.Csr.from_scipy(P[:, :-1].T.tocsr())
    Hb = gs.Hierarchy.from_levels([A0, A1], [bad_P], [bad_R])
    with pytest.raises(gs.GpuSolverError, match="transfer operator shape mismatch"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2).setup(Hb)


def test_direct_level_above_limit(small):
    A, H = small
    n_last = H.get("A", H.levels - 1).rows
    with pytest.raises(gs.GpuSolverError, match="above the dense direct-solve limit"):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_direct=1,
                  max_direct=n_last - 1).setup(H)


def test_errors_leave_the_solver_usable(small):
    """After a failed setup the same process still sets up and solves."""
    A, H = small
    with pytest.raises(gs.GpuSolverError):
        gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, omega=3.0).setup(H)
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    b = gs.uniform_pm1(3, A.rows)
    x, info = s.solve(b, tol=1e-8, maxit=200)
    assert info.converged and info.true_relres <= 1.5e-8


def test_vector_lengths_are_checked(small):
    """src/CGSolver.cc:90-98: a right-hand side or initial guess of another
    length than the matrix is rejected before anything is read; so are
    wrong-length vectors for one V-cycle or one SpMV."""
    A, H = small
    s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2)
    s.setup(H)
    assert s.size() == A.rows
    b = gs.uniform_pm1(3, A.rows)
    with pytest.raises(gs.GpuSolverError, match="RHS size does not match matrix"):
        s.solve(b[:-1])
    with pytest.raises(gs.GpuSolverError, match="RHS size does not match matrix"):
        s.solve(np.concatenate([b, b]))
    with pytest.raises(gs.GpuSolverError, match="initial guess size does not match matrix"):
        s.solve(b, x0=np.zeros(A.rows + 5))
    with pytest.raises(gs.GpuSolverError, match="vector size does not match the finest level"):
        s.vcycle(b[:10])
    with pytest.raises(gs.GpuSolverError, match="encodeSpmv: size mismatch"):
        s.spmv(np.zeros(A.rows - 1))
    x, info = s.solve(b)  # still usable
    assert info.converged


def test_fallback_preconditions(small):
    """A guarded solve refuses a plain solver that is not set up, was set up
    for another size, or has not converged once (no rate to compare with)."""
    A, H = small
    b = gs.uniform_pm1(5, A.rows)
    learned = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2)
    learned.setup(H)
    plain = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2)
    learned.set_fallback(plain, lag=1.5)
    with pytest.raises(gs.GpuSolverError, match="fallback solver is not set up"):
        learned.solve(b)
    plain.setup(H)
    with pytest.raises(gs.GpuSolverError, match="no converged solve"):
        learned.solve(b)
    B = gs.poisson2d(20, 20)
    HB = gs.Hierarchy.smoothed_aggregation(B, min_dofs=50)
    other = gs.Solver(gs.PREC_MIXED, smoother=gs.SGS, coarse_steps=2)
    other.setup(HB)
    other.solve(gs.uniform_pm1(6, B.rows))
    learned.set_fallback(other, lag=1.5)
    with pytest.raises(gs.GpuSolverError, match="another matrix size"):
        learned.solve(b)
    plain.solve(b)
    assert plain.result("log_rate") < 0.0
    learned.set_fallback(plain, lag=1.5)
    x, info = learned.solve(b)
    assert info.converged and not learned.fell_back()


def test_rejected_level_weights_change_nothing(small):
    """An invalid per-level entry (omega >= 2 on the last smoothed level) is
    rejected before any entry is applied: the next solve reproduces the
    previous result bit for bit (eager and captured-graph paths)."""
    A, H = small
    b = gs.uniform_pm1(7, A.rows)
    for gi in (0, 8):
        s = gs.Solver(gs.PREC_DOUBLE, smoother=gs.SGS, coarse_steps=2, graph_iters=gi)
        s.setup(H)
        L = len(s.describe()["levels"])
        s.set_level_weights(omega=[1.2] * L)
        x1, i1 = s.solve(b)
        bad = [0.8] * L
        bad[L - 2] = 2.5
        with pytest.raises(gs.GpuSolverError, match="omega must be in"):
            s.set_level_weights(omega=bad)
        assert [lv["omega"] for lv in s.describe()["levels"]] == pytest.approx([1.2] * L)
        x2, i2 = s.solve(b)
        np.testing.assert_array_equal(x1, x2)
        assert i1.iterations == i2.iterations
