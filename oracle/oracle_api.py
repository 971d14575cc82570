"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
setup sources built into oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY:
imported by tests/, by __graft_entry__.smoke() and by the CPU-baseline leg
of bench.py, always as the checker, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libref.so"

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64

KIND = {"jacobi": 0, "sgs": 1, "block_jacobi": 2, "block_gs": 3}


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if ref and Path("/root/reference/src/GraphColoring.cc").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _lib() -> C.CDLL:
    if not ORACLE_SO.exists():
        build(ref=False)
    L = C.CDLL(str(ORACLE_SO))
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_set_parallel_dot": (None, [C.c_int]),
        "orc_num_threads": (C.c_int, []),
        "orc_spmv_d": (C.c_int, [_i64, _i64, _i32p, _i32p, _f64p, _f64p, _f64p]),
        "orc_dense_spd_inverse": (C.c_int, [_i64, _f64p]),
        "orc_spmv_f": (C.c_int, [_i64, _i64, _i32p, _i32p, _f32p, _f32p, _f32p]),
        "orc_gs_color_d": (C.c_int, [_i64, _i32p, _i32p, _f64p, _i32p, _i64, _i64, _f64p, _f64p,
                                     C.c_double]),
        "orc_gs_color_f": (C.c_int, [_i64, _i32p, _i32p, _f32p, _i32p, _i64, _i64, _f32p, _f32p,
                                     C.c_float]),
        "orc_block_gs_color_d": (C.c_int, [_i64, _i32p, _i32p, _f64p, _i32p, _f64p, _i64, _i64,
                                           _f64p, _f64p, C.c_double]),
        "orc_block_gs_color_f": (C.c_int, [_i64, _i32p, _i32p, _f32p, _i32p, _f32p, _i64, _i64,
                                           _f32p, _f32p, C.c_float]),
        "orc_apply_block_inverses_d": (C.c_int, [_f64p, _f64p, _f64p, _i64]),
        "orc_apply_block_inverses_f": (C.c_int, [_f32p, _f32p, _f32p, _i64]),
        "orc_dot_d": (C.c_double, [_f64p, _f64p, _i64]),
        "orc_dot_f": (C.c_float, [_f32p, _f32p, _i64]),
        "orc_coloring": (C.c_int, [_i64, _i32p, _i32p, _i32p, _i32p, _i32p,
                                   C.POINTER(C.c_int32)]),
        "orc_rcm": (C.c_int, [_i64, _i32p, _i32p, _i32p, _i32p]),
        "orc_hierarchical_reordering": (C.c_int, [C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _i32p,
                                                  _i32p, _vp]),
        "orc_ell_pack_d": (C.c_int, [_i64, _i32p, _i32p, _f64p, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                     C.POINTER(_i64)]),
        "orc_permute_keep_order": (C.c_int, [_i64, _i32p, _i32p, _f64p, _i32p, _i32p, _i32p, _i32p,
                                             _f64p]),
        "orc_permute_sorted": (C.c_int, [_i64, _i64, _i32p, _i32p, _f64p, _i32p, _i32p, _i32p,
                                         _i32p, _f64p, C.POINTER(_i64)]),
        "orc_block_graph": (C.c_int, [_i64, _i32p, _i32p, C.c_int32, _vp, _vp, C.POINTER(_i64)]),
        "orc_pcg": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _i32p, _f64p, _i32p, _i32p,
                              C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp, _f64p, _f64p,
                              C.c_double,
                              _i64, C.POINTER(_i64), C.POINTER(C.c_double), _f64p,
                              C.POINTER(_i64)]),
        "orc_vcycle": (C.c_int, [_vp, C.c_int32, _i32p, _f64p, _i32p, _i32p, C.c_int32, C.c_int32,
                                 C.c_int32, C.c_int32, _vp, _f64p, _f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _lib()


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())


def dense_spd_inverse(a):
    """FP64 inverse of a dense SPD matrix through its Cholesky factor."""
    m = np.ascontiguousarray(a, np.float64).copy()
    _check(lib.orc_dense_spd_inverse(m.shape[0], m.reshape(-1)))
    return m


def spmv(ptr, col, val, x, ncols=None):
    n = len(ptr) - 1
    if val.dtype == np.float32:
        y = np.empty(n, np.float32)
        _check(lib.orc_spmv_f(n, ncols or n, ptr, col, val, np.ascontiguousarray(x, np.float32), y))
    else:
        y = np.empty(n, np.float64)
        _check(lib.orc_spmv_d(n, ncols or n, ptr, col, val, np.ascontiguousarray(x, np.float64), y))
    return y


def gs_color(ptr, col, val, order, begin, count, rhs, x, omega):
    """One colour of Gauss-Seidel in place on x (src/CPU/CPUBackend.cc:238-268)."""
    n = len(ptr) - 1
    f = lib.orc_gs_color_f if val.dtype == np.float32 else lib.orc_gs_color_d
    _check(f(n, ptr, col, val, order, begin, count, rhs, x, omega))


def block_gs_color(ptr, col, val, order, dinv, begin, end, rhs, x, omega):
    n = len(ptr) - 1
    f = lib.orc_block_gs_color_f if val.dtype == np.float32 else lib.orc_block_gs_color_d
    _check(f(n, ptr, col, val, order, dinv, begin, end, rhs, x, omega))


def apply_block_inverses(inv, res):
    out = np.empty_like(res)
    f = lib.orc_apply_block_inverses_f if res.dtype == np.float32 else lib.orc_apply_block_inverses_d
    _check(f(inv, res, out, len(res)))
    return out


def dot(x, y):
    if x.dtype == np.float32:
        return lib.orc_dot_f(x, y, len(x))
    return lib.orc_dot_d(x, y, len(x))


def coloring(ptr, col):
    n = len(ptr) - 1
    colors = np.empty(n, np.int32)
    verts = np.empty(n, np.int32)
    offs = np.empty(n + 1, np.int32)
    nc = C.c_int32()
    _check(lib.orc_coloring(n, ptr, col, colors, verts, offs, C.byref(nc)))
    return colors, verts, offs[:nc.value + 1].copy(), nc.value


def rcm(ptr, col):
    n = len(ptr) - 1
    o2n = np.empty(n, np.int32)
    n2o = np.empty(n, np.int32)
    _check(lib.orc_rcm(n, ptr, col, o2n, n2o))
    return o2n, n2o


def hierarchical_reordering(A, R):
    """A: per-level (ptr, col) of square matrices, finest first; R: per-
    transfer (ptr, col, val) restrictions (n[k+1] x n[k]). Returns the
    per-level (old_to_new, new_to_old) pairs and the orphan counts."""
    L = len(A)
    n = np.array([len(a[0]) - 1 for a in A], np.int64)
    keep = []

    def arr(xs, dt):
        xs = [np.ascontiguousarray(x, dt) for x in xs]
        keep.extend(xs)
        p = (_vp * max(len(xs), 1))(*[x.ctypes.data for x in xs])
        keep.append(p)
        return p

    ap, ac = arr([a[0] for a in A], np.int32), arr([a[1] for a in A], np.int32)
    rp = arr([r[0] for r in R], np.int32)
    rc = arr([r[1] for r in R], np.int32)
    rv = arr([r[2] for r in R], np.float64)
    tot = int(n.sum())
    o2n = np.empty(tot, np.int32)
    n2o = np.empty(tot, np.int32)
    orph = np.zeros(max(L - 1, 1), np.int64)
    _check(lib.orc_hierarchical_reordering(L, n.ctypes.data, ap, ac, rp, rc, rv, o2n, n2o,
                                           orph.ctypes.data))
    out, off = [], 0
    for k in range(L):
        out.append((o2n[off:off + n[k]].copy(), n2o[off:off + n[k]].copy()))
        off += int(n[k])
    return out, orph[:L - 1].copy()


def ell_pack(ptr, col, val, width):
    n = len(ptr) - 1
    ov = _i64()
    _check(lib.orc_ell_pack_d(n, ptr, col, val, width, None, None, None, None, None, C.byref(ov)))
    ec = np.empty(n * width, np.int32)
    ev = np.empty(n * width, np.float64)
    op = np.empty(n + 1, np.int32)
    oc = np.empty(max(ov.value, 1), np.int32)
    ovv = np.empty(max(ov.value, 1), np.float64)
    _check(lib.orc_ell_pack_d(n, ptr, col, val, width, ec.ctypes.data_as(_vp),
                              ev.ctypes.data_as(_vp), op.ctypes.data_as(_vp),
                              oc.ctypes.data_as(_vp), ovv.ctypes.data_as(_vp), C.byref(ov)))
    return ec, ev, op, oc[:ov.value].copy(), ovv[:ov.value].copy()


def permute_keep_order(ptr, col, val, o2n):
    n = len(ptr) - 1
    o2n = np.ascontiguousarray(o2n, np.int32)
    n2o = np.empty_like(o2n)
    n2o[o2n] = np.arange(n, dtype=np.int32)
    optr = np.empty(n + 1, np.int32)
    ocol = np.empty(len(col), np.int32)
    oval = np.empty(len(col), np.float64)
    _check(lib.orc_permute_keep_order(n, ptr, col, val, o2n, n2o, optr, ocol, oval))
    return optr, ocol, oval


def permute_sorted(ptr, col, val, ro2n, co2n, ncols):
    n = len(ptr) - 1
    optr = np.empty(n + 1, np.int32)
    ocol = np.empty(max(len(col), 1), np.int32)
    oval = np.empty(max(len(col), 1), np.float64)
    nnz = _i64()
    _check(lib.orc_permute_sorted(n, ncols, ptr, col, val, np.ascontiguousarray(ro2n, np.int32),
                                  np.ascontiguousarray(co2n, np.int32), optr, ocol, oval,
                                  C.byref(nnz)))
    return optr, ocol[:nnz.value].copy(), oval[:nnz.value].copy()


def block_graph(ptr, col, bs=3):
    n = len(ptr) - 1
    nnz = _i64()
    _check(lib.orc_block_graph(n, ptr, col, bs, None, None, C.byref(nnz)))
    gp = np.empty(n // bs + 1, np.int32)
    gc = np.empty(max(nnz.value, 1), np.int32)
    _check(lib.orc_block_graph(n, ptr, col, bs, gp.ctypes.data_as(_vp), gc.ctypes.data_as(_vp),
                               C.byref(nnz)))
    return gp, gc[:nnz.value].copy()


class _Hier(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("n", C.POINTER(C.c_int64)),
                ("aptr", C.POINTER(_vp)), ("acol", C.POINTER(_vp)), ("aval", C.POINTER(_vp)),
                ("perm", C.POINTER(_vp)), ("pcols", C.POINTER(C.c_int64)),
                ("pptr", C.POINTER(_vp)), ("pcol", C.POINTER(_vp)), ("pval", C.POINTER(_vp)),
                ("rptr", C.POINTER(_vp)), ("rcol", C.POINTER(_vp)), ("rval", C.POINTER(_vp))]


class HierView:
    """Packs lists of level / transfer CSR arrays (each a (ptr, col, val)
    tuple of numpy arrays) into the oracle's hierarchy struct."""

    def __init__(self, A, P, R, perms=None):
        self.keep = []
        L = len(A)

        def arr(items):
            a = (_vp * max(L, 1))()
            for i, it in enumerate(items):
                if it is None:
                    a[i] = None
                else:
                    it = np.ascontiguousarray(it)
                    self.keep.append(it)
                    a[i] = it.ctypes.data_as(_vp)
            return a

        n = np.array([len(a[0]) - 1 for a in A], np.int64)
        self.keep.append(n)
        h = _Hier()
        h.nlev = L
        h.n = n.ctypes.data_as(C.POINTER(C.c_int64))
        self._arrs = [arr([a[0] for a in A]), arr([a[1] for a in A]), arr([a[2] for a in A]),
                      arr(perms if perms is not None else [None] * L),
                      arr([p[0] for p in P]), arr([p[1] for p in P]), arr([p[2] for p in P]),
                      arr([r[0] for r in R]), arr([r[1] for r in R]), arr([r[2] for r in R])]
        (h.aptr, h.acol, h.aval, h.perm, h.pptr, h.pcol, h.pval, h.rptr, h.rcol,
         h.rval) = [C.cast(a, C.POINTER(_vp)) for a in self._arrs]
        h.pcols = C.cast(None, C.POINTER(C.c_int64))
        self.h = h
        self.nlev = L

    def ptr(self):
        return C.cast(C.pointer(self.h), _vp)


def _level_params(nlev, kinds, omegas, pres, posts):
    def full(v, dt):
        if np.isscalar(v):
            return np.full(nlev, v, dt)
        return np.ascontiguousarray(v, dt)
    return full(kinds, np.int32), full(omegas, np.float64), full(pres, np.int32), full(posts, np.int32)


def _coarse_inv_arg(coarse_inv, hier: HierView):
    """The caller's dense FP64 inverse of the coarsest matrix (row-major,
    in the coarsest level's permuted order), or None for the oracle's own
    Cholesky inverse."""
    if coarse_inv is None:
        return None, None
    a = np.ascontiguousarray(coarse_inv, np.float64)
    nc = int(hier.h.n[hier.nlev - 1])
    if a.shape != (nc, nc):
        raise ValueError(f"coarse inverse must be {nc} x {nc}, got {a.shape}")
    return a, a.ctypes.data_as(_vp)


def pcg(hier: HierView, b, x0=None, tol=1e-8, maxit=1000, prec=1, cg_double=True,
        mg_double=True, kinds=1, omegas=1.0, pres=1, posts=1, coarse_steps=2,
        direct_coarse=False, vcycles=1, reference_vcycle=False, coarse_inv=None):
    """Restated CGSolver::solve with the multigrid (prec=1) or identity (0)
    preconditioner. reference_vcycle reproduces the reference's correction
    aliasing on coarse levels (solution = 2 P e); off, the textbook
    x += P e. coarse_inv (with direct_coarse): a precomputed inverse of the
    coarsest matrix. Returns x, iterations, relres2, history."""
    n = len(b)
    keep, ci = _coarse_inv_arg(coarse_inv, hier)
    k, o, pr, po = _level_params(hier.nlev, kinds, omegas, pres, posts)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64, copy=True)
    it = _i64()
    rr = C.c_double()
    hist = np.empty(maxit + 1, np.float64)
    hl = _i64()
    _check(lib.orc_pcg(hier.ptr(), prec, int(cg_double), int(mg_double), k, o, pr, po,
                       coarse_steps, int(direct_coarse), vcycles, int(reference_vcycle), ci,
                       np.ascontiguousarray(b, np.float64), x, tol, maxit, C.byref(it),
                       C.byref(rr), hist, C.byref(hl)))
    return x, it.value, rr.value, hist[:hl.value].copy()


def vcycle(hier: HierView, r, mg_double=True, kinds=1, omegas=1.0, pres=1, posts=1,
           coarse_steps=2, direct_coarse=False, vcycles=1, reference_vcycle=False,
           coarse_inv=None):
    k, o, pr, po = _level_params(hier.nlev, kinds, omegas, pres, posts)
    keep, ci = _coarse_inv_arg(coarse_inv, hier)
    z = np.empty(len(r), np.float64)
    _check(lib.orc_vcycle(hier.ptr(), int(mg_double), k, o, pr, po, coarse_steps,
                          int(direct_coarse), vcycles, int(reference_vcycle), ci,
                          np.ascontiguousarray(r, np.float64), z))
    return z


# ---------------------------------------------------------------------------
# The reference's own setup sources (oracle/_ref), when built.
# ---------------------------------------------------------------------------
def ref_lib():
    if not REF_SO.exists():
        return None
    L = C.CDLL(str(REF_SO))
    L.ref_coloring.restype = C.c_int
    L.ref_coloring.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _i32p, C.POINTER(C.c_int32)]
    L.ref_rcm.restype = C.c_int
    L.ref_rcm.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p]
    L.ref_permutation_valid.restype = C.c_int
    L.ref_permutation_valid.argtypes = [_i64, _i32p, _i32p]
    return L


def ref_coloring(L, ptr, col):
    n = len(ptr) - 1
    colors = np.empty(n, np.int32)
    verts = np.empty(n, np.int32)
    offs = np.empty(n + 1, np.int32)
    nc = C.c_int32()
    assert L.ref_coloring(n, ptr, col, colors, verts, offs, C.byref(nc)) == 0
    return colors, verts, offs[:nc.value + 1].copy(), nc.value


def ref_rcm(L, ptr, col):
    n = len(ptr) - 1
    o2n = np.empty(n, np.int32)
    n2o = np.empty(n, np.int32)
    assert L.ref_rcm(n, ptr, col, o2n, n2o) == 0
    return o2n, n2o
