"""gs_dot_noalloc, the allocation-free variant of gs_encode_dot, through the
C ABI: on device arrays and a device scalar owned by the caller (here the
Backend's vectors, reached through gs_dvec_data / gs_dscalar_data), on the
encoder's stream and on the null stream, it returns gs_encode_dot's sum bit
for bit; two calls in a row, the second on fewer values, so partial sums the
first call left in the shared scratch array must not leak into the second.
GPU."""
import ctypes as C

import numpy as np
import pytest

import gpusolver as gs

pytestmark = pytest.mark.gpu

# Value types and their GS_DTYPE_* codes (include/gpusolver.h).
DTYPES = [(np.float64, 1), (np.float32, 0)]


def noalloc_dot(dtype_code, x, y, n, out, stream):
    return gs.lib.gs_dot_noalloc(dtype_code, C.c_void_p(x.data()), C.c_void_p(y.data()), n,
                                 C.c_void_p(out.data()), C.c_void_p(stream))


@pytest.mark.parametrize("dt,code", DTYPES)
def test_noalloc_dot_matches_encode_dot(dt, code):
    be = gs.Backend(dt)
    e = be.create_encoder()
    rng = np.random.default_rng(17)
    # 700 000 values: more than 2048 workgroups' worth, so the capped grid
    # strides; then 1000 values (4 partials) on the same scratch array.
    for n, stream in ((700_000, e.stream()), (1000, e.stream()), (1000, 0), (1, 0)):
        xv = rng.uniform(-1, 1, n).astype(dt)
        yv = rng.uniform(-1, 1, n).astype(dt)
        x, y = be.create_vector(xv), be.create_vector(yv)
        ref, out = be.create_scalar(), be.create_scalar(7.0)
        be.dot(x, y, ref)
        print(f"[no_alloc] dtype {np.dtype(dt).name} n {n} stream {'encoder' if stream else 'null'}",
              flush=True)
        assert noalloc_dot(code, x, y, n, out, stream) == 0, gs.lib.gs_last_error()
        be.submit_and_wait(e)
        got, want = out.download(), ref.download()
        assert np.array_equal(np.array([got], dt).view(np.uint8), np.array([want], dt).view(np.uint8)), \
            (n, got, want)
        exact = float(np.dot(xv.astype(np.float64), yv.astype(np.float64)))
        assert abs(got - exact) <= (1e-4 if dt == np.float32 else 1e-12) * max(1.0, abs(exact))


def test_noalloc_dot_rejects_bad_arguments():
    be = gs.Backend(np.float64)
    x, out = be.create_vector(4), be.create_scalar()
    assert noalloc_dot(1, x, x, -1, out, 0) != 0
    assert "negative size" in gs.lib.gs_last_error().decode()
    assert gs.lib.gs_dot_noalloc(5, C.c_void_p(x.data()), C.c_void_p(x.data()), 4,
                                 C.c_void_p(out.data()), None) != 0
    assert "unknown value type" in gs.lib.gs_last_error().decode()
    assert gs.lib.gs_dot_noalloc(1, None, None, 4, C.c_void_p(out.data()), None) != 0
    assert "must not be null" in gs.lib.gs_last_error().decode()
    # n = 0 sums nothing.
    assert noalloc_dot(1, x, x, 0, out, 0) == 0
    assert out.download() == 0.0
