"""The Backend plugin on HIP, with the reference's own kernel tests
(tests/cuda/*.cc, tests/metal/DotProductTests.cc, ScalarKernelTests.cc) as
known answers, plus bit-exact agreement with the CPU oracle on larger
inputs for the row-wise kernels (same summation order, no contraction)."""
import numpy as np
import pytest
import scipy.sparse as sp

import gpusolver as gs
import oracle_api as orc

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]


def csr(dense):
    m = sp.csr_matrix(np.asarray(dense, np.float64))
    m.sort_indices()
    return gs.Csr.from_scipy(m)


TRI4 = [[4, -1, 0, 0], [-1, 4, -1, 0], [0, -1, 4, -1], [0, 0, -1, 4]]


@pytest.mark.parametrize("dt", DTYPES)
def test_gauss_seidel_one_color(dt):
    """tests/cuda/testCudaGaussSeidel.cc:15-111."""
    be = gs.Backend(dt)
    A = be.create_csr_matrix(csr(TRI4))
    colors = be.create_index_vector([0, 2, 1, 3])
    rhs = be.create_vector([4, 8, 12, 16])
    x = be.create_vector(np.zeros(4))
    e = be.create_encoder()
    be.encode_gauss_seidel_color(e, A, colors, 0, 2, rhs, x, 1.0)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(x.download(), np.array([1, 0, 3, 0], dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_gauss_seidel_full_two_color_sweep(dt):
    """tests/cuda/testCudaGaussSeidel.cc:114-222."""
    be = gs.Backend(dt)
    A = be.create_csr_matrix(csr(TRI4))
    colors = be.create_index_vector([0, 2, 1, 3])
    rhs = be.create_vector([4, 8, 12, 16])
    x = be.create_vector(4)
    e = be.create_encoder()
    be.encode_set_zero(e, x)
    be.encode_gauss_seidel_color(e, A, colors, 0, 2, rhs, x, 1.0)
    be.encode_gauss_seidel_color(e, A, colors, 2, 2, rhs, x, 1.0)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(x.download(), np.array([1, 3, 3, 4.75], dt))


def test_gauss_seidel_zero_diagonal_raises():
    be = gs.Backend(np.float32)
    A = be.create_csr_matrix(csr([[0, 1], [1, 2]]))
    colors = be.create_index_vector([0, 1])
    rhs = be.create_vector([1, 1])
    x = be.create_vector(2)
    e = be.create_encoder()
    be.encode_gauss_seidel_color(e, A, colors, 0, 2, rhs, x, 1.0)
    with pytest.raises(gs.GpuSolverError, match="zero diagonal"):
        be.submit_and_wait(e)


@pytest.mark.parametrize("dt", DTYPES)
def test_apply_two_inverse_blocks(dt):
    """tests/cuda/testCudaBlockOps.cc:15-78."""
    be = gs.Backend(dt)
    inv = be.create_vector([2, 0, 0, 0, 3, 0, 0, 0, 4, 1, 1, 0, 0, 2, 0, 0, 0, 3])
    res = be.create_vector([1, 2, 3, 4, 5, 6])
    corr = be.create_vector(6)
    e = be.create_encoder()
    be.encode_apply_block_inverses3x3(e, inv, res, corr)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(corr.download(), np.array([2, 6, 12, 9, 10, 18], dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_spmv_known_and_updated_values(dt):
    """tests/cuda/testCudaSpmv.cc:123-273."""
    be = gs.Backend(dt)
    A = be.create_csr_matrix(csr([[4, 1, 0], [1, 3, 2], [0, 2, 5]]))
    x = be.create_vector([1, 2, 3])
    y = be.create_vector(3)
    be.spmv(A, x, y)
    np.testing.assert_array_equal(y.download(), np.array([6, 13, 19], dt))
    B = be.create_csr_matrix(csr([[2, 1], [1, 3]]))
    B.update_values([4, 2, 2, 6])
    x2 = be.create_vector([1, 2])
    y2 = be.create_vector(2)
    be.spmv(B, x2, y2)
    np.testing.assert_array_equal(y2.download(), np.array([8, 14], dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_dot_32(dt):
    be = gs.Backend(dt)
    a, b, s = be.create_vector([1, 2, 3]), be.create_vector([4, 5, 6]), be.create_scalar()
    be.dot(a, b, s)
    assert s.download() == 32.0


@pytest.mark.parametrize("n", [1, 17, 255, 256, 257, 1000, 100000])
def test_dot_sizes(n):
    """tests/metal/DotProductTests.cc:27-140 (FP32, tol 1e-4 max(1, |e|))."""
    be = gs.Backend(np.float32)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = rng.uniform(-1, 1, n).astype(np.float32)
    s = be.create_scalar()
    be.dot(be.create_vector(x), be.create_vector(y), s)
    e = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    assert abs(s.download() - e) <= 1e-4 * max(1.0, abs(e))


def test_dot_size_mismatch_raises():
    be = gs.Backend(np.float32)
    with pytest.raises(gs.GpuSolverError, match="size mismatch"):
        be.dot(be.create_vector(3), be.create_vector(4), be.create_scalar())


def test_scalar_chain_stays_on_device():
    """tests/cuda/testCudaScalarOps.cc:183-241: -sqrt(3 * 4) without a host
    round trip between the steps."""
    be = gs.Backend(np.float32)
    a, b, p, r, n = (be.create_scalar(3.0), be.create_scalar(4.0), be.create_scalar(),
                     be.create_scalar(), be.create_scalar())
    e = be.create_encoder()
    be.encode_scalar_multiply(e, a, b, p)
    be.encode_scalar_sqrt(e, p, r)
    be.encode_scalar_negate(e, r, n)
    be.submit_and_wait(e)
    assert n.download() == pytest.approx(-np.sqrt(np.float32(12.0)), rel=1e-6)
    q = be.create_scalar()
    be.scalar_divide(a, b, q)
    assert q.download() == np.float32(0.75)
    be.scalar_set(5.5, q)
    be.scalar_copy(q, r)
    assert r.download() == 5.5


def test_several_vector_ops_one_wait():
    """tests/cuda/testCudaVectorOps.cc:270-322."""
    be = gs.Backend(np.float32)
    x = be.create_vector([1, 2, 3, 4])
    y = be.create_vector([10, 20, 30, 40])
    s = be.create_vector([2, 2, 0.5, 1])
    alpha = be.create_scalar(3.0)
    e = be.create_encoder()
    be.encode_axpy(e, 2.0, x, y)         # y = [12, 24, 36, 48]
    be.encode_scale(e, y, 0.5)            # [6, 12, 18, 24]
    be.encode_scale_vector(e, s, y)       # [12, 24, 9, 24]
    be.encode_axpy(e, alpha, x, y)        # [15, 30, 18, 36]
    be.encode_scale(e, x, alpha)          # x = [3, 6, 9, 12]
    z = be.create_vector(4)
    be.encode_copy(e, x, z)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(y.download(), [15, 30, 18, 36])
    np.testing.assert_array_equal(z.download(), [3, 6, 9, 12])


@pytest.mark.parametrize("dt", DTYPES)
def test_spmv_bitexact_vs_oracle(dt):
    A = gs.poisson3d(20)
    rng = np.random.default_rng(0)
    A.val = A.val * rng.uniform(0.5, 1.5, A.nnz)
    be = gs.Backend(dt)
    dA = be.create_csr_matrix(A)
    x = rng.uniform(-1, 1, A.rows)
    y = be.create_vector(A.rows)
    be.spmv(dA, be.create_vector(x), y)
    np.testing.assert_array_equal(y.download(), orc.spmv(A.ptr, A.col, A.val.astype(dt), x.astype(dt)))


@pytest.mark.parametrize("width", [3, 5, 7])
def test_ell_spmv_matches_csr(width):
    A = gs.poisson3d(12)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, A.rows)
    be = gs.Backend(np.float64)
    y = be.create_vector(A.rows)
    be.spmv(be.create_ell_matrix(A, width), be.create_vector(x), y)
    np.testing.assert_array_equal(y.download(), orc.spmv(A.ptr, A.col, A.val, x))


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("width", [3, 7])
def test_ell_packed_matches_library_packing(dt, width):
    """createELLMatrix from arrays packed on the host (what HostELLMatrix
    holds) gives the matrix the library packs itself, bit for bit."""
    A = gs.poisson3d(12)
    x = np.random.default_rng(2).uniform(-1, 1, A.rows)
    be = gs.Backend(dt)
    ec, ev, op, oc, ov = gs.ell_pack(A, width)
    P = be.create_ell_packed(A.rows, A.rows, width, ec, ev.astype(dt), op, oc, ov.astype(dt))
    assert P.info() == (A.rows, A.rows, width, oc.size)
    y1, y2 = be.create_vector(A.rows), be.create_vector(A.rows)
    xd = be.create_vector(x)
    be.spmv(P, xd, y1)
    be.spmv(be.create_ell_matrix(A, width), xd, y2)
    np.testing.assert_array_equal(y1.download(), y2.download())
    np.testing.assert_array_equal(y1.download(),
                                  orc.spmv(A.ptr, A.col, A.val.astype(dt), x.astype(dt)))


def test_ell_packed_checks_indices():
    be = gs.Backend(np.float32)
    one = np.ones(2, np.float32)
    with pytest.raises(gs.GpuSolverError, match="column index out of bounds"):
        be.create_ell_packed(2, 2, 1, [0, 2], one, [0, 0, 0], [], [])
    with pytest.raises(gs.GpuSolverError, match="overflow row pointer does not match"):
        be.create_ell_packed(2, 2, 1, [0, 1], one, [0, 0, 1], [], [])
    with pytest.raises(gs.GpuSolverError, match="ELL width must be nonzero"):
        be.create_ell_packed(2, 2, 0, [], [], [0, 0, 0], [], [])


def test_device_object_sizes_and_updates():
    A = gs.poisson2d(9, 7)
    be = gs.Backend(np.float32)
    m = be.create_csr_matrix(A)
    assert m.info() == (A.rows, A.rows, A.nnz)
    iv = be.create_index_vector(np.arange(5, dtype=np.int32))
    assert iv.size() == 5
    iv.update_values([4, 3, 2, 1, 0])
    np.testing.assert_array_equal(iv.download(), [4, 3, 2, 1, 0])
    with pytest.raises(gs.GpuSolverError, match="DeviceIndexVector::updateValues: size mismatch"):
        iv.update_values([1, 2, 3])


@pytest.mark.parametrize("dt", DTYPES)
def test_gs_color_sweep_bitexact_vs_oracle(dt):
    A = gs.poisson2d(40, 30)
    colors, verts, offs, nc = gs.graph_coloring(A)
    be = gs.Backend(dt)
    dA = be.create_csr_matrix(A)
    dv = be.create_index_vector(verts)
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, A.rows).astype(dt)
    x = np.zeros(A.rows, dt)
    dx, db = be.create_vector(x), be.create_vector(b)
    e = be.create_encoder()
    for c in list(range(nc)) + list(range(nc - 1, -1, -1)):
        be.encode_gauss_seidel_color(e, dA, dv, offs[c], offs[c + 1] - offs[c], db, dx, 1.2)
        orc.gs_color(A.ptr, A.col, A.val.astype(dt), verts, offs[c], offs[c + 1] - offs[c], b, x,
                     1.2)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(dx.download(), x)


@pytest.mark.parametrize("dt", DTYPES)
def test_block_gs_sweep_bitexact_vs_oracle(dt):
    E, rhs, _ = gs.elasticity(6)
    gp, gc = gs.block_graph(E, 3)
    colors, verts, offs, nc = gs.graph_coloring(gs.Csr(gp, gc, np.ones(len(gc)), (len(gp) - 1,) * 2))
    inv = gs.inverse_diag_blocks3(E).astype(dt)
    be = gs.Backend(dt)
    dA, dv, di = be.create_csr_matrix(E), be.create_index_vector(verts), be.create_vector(inv)
    b = rhs.astype(dt)
    x = np.zeros(E.rows, dt)
    dx, db = be.create_vector(x), be.create_vector(b)
    e = be.create_encoder()
    for c in list(range(nc)) + list(range(nc - 1, -1, -1)):
        be.encode_block_gauss_seidel_color3x3(e, dA, dv, di, offs[c], offs[c + 1], db, dx, 0.7)
        orc.block_gs_color(E.ptr, E.col, E.val.astype(dt), verts, inv, offs[c], offs[c + 1], b, x,
                           0.7)
    be.submit_and_wait(e)
    np.testing.assert_array_equal(dx.download(), x)


def test_block_gs_invalid_range_raises():
    be = gs.Backend(np.float32)
    E, rhs, _ = gs.elasticity(3)
    dA = be.create_csr_matrix(E)
    dv = be.create_index_vector(np.arange(E.rows // 3))
    di = be.create_vector(np.zeros(3 * E.rows))
    dx, db = be.create_vector(E.rows), be.create_vector(E.rows)
    e = be.create_encoder()
    with pytest.raises(gs.GpuSolverError, match="invalid color range"):
        be.encode_block_gauss_seidel_color3x3(e, dA, dv, di, 3, 2, db, dx, 0.7)


def test_device_addresses_and_stream():
    """CudaDeviceVector::data / CudaDeviceScalar::data /
    CudaEncoder::stream: stable, distinct device addresses and the
    encoder's own stream."""
    be = gs.Backend(np.float64)
    x, y, s = be.create_vector(np.ones(10)), be.create_vector(10), be.create_scalar(2.0)
    e1, e2 = be.create_encoder(), be.create_encoder()
    px, py, ps = x.data(), y.data(), s.data()
    assert px and py and ps and len({px, py, ps}) == 3
    assert x.data() == px and s.data() == ps
    assert e1.stream() and e1.stream() != e2.stream()
