"""ctypes binding of libgpusolver.so (C ABI in include/gpusolver.h).

Thin host-side plumbing for the test suite and the benchmark: CSR arrays
in numpy, the problem generators, the AMG hierarchy and the GPU solver.
The library is loaded from this directory; importing fails loudly when it
has not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libgpusolver.so"

PREC_DOUBLE, PREC_MIXED, PREC_SINGLE = 0, 1, 2
DAMPED_JACOBI, SGS, BLOCK_JACOBI, BLOCK_GS = 0, 1, 2, 3

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32


class GpuSolverError(RuntimeError):
    pass


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python {HERE / 'build.py'}` "
            "(the solver has no CPU fallback)")
    return C.CDLL(str(LIB_PATH))


lib = _load()

_SIGS = {
    "gs_last_error": (C.c_char_p, []),
    "gs_version": (C.c_int, []),
    "gs_csr_create": (C.c_int, [_i64, _i64, _i64, _i32p, _i32p, _f64p, C.POINTER(_vp)]),
    "gs_csr_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "gs_csr_export": (C.c_int, [_vp, _i32p, _i32p, _f64p]),
    "gs_csr_destroy": (C.c_int, [_vp]),
    "gs_csr_permute_keep_order": (C.c_int, [_vp, _i32p, C.POINTER(_vp)]),
    "gs_csr_permute_sorted": (C.c_int, [_vp, _i32p, _i32p, C.POINTER(_vp)]),
    "gs_ell_pack": (C.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64)]),
    "gs_graph_coloring": (C.c_int, [_i64, _i32p, _i32p, _i32p, _i32p, _i32p, C.POINTER(_i32)]),
    "gs_rcm": (C.c_int, [_i64, _i32p, _i32p, _i32p, _i32p]),
    "gs_permutation_check": (C.c_int, [_i64, _i32p, _i32p]),
    "gs_block_graph": (C.c_int, [_vp, _i32, _vp, _vp, C.POINTER(_i64)]),
    "gs_expand_block_permutation": (C.c_int, [_i64, _i32, _i32p, _i32p, _i32p]),
    "gs_inverse_diag_blocks3": (C.c_int, [_vp, _f64p]),
    "gs_dense_spd_inverse": (C.c_int, [_i64, _f64p]),
    "gs_dense_spd_inverse_device": (C.c_int, [_i64, _f64p]),
    "gs_problem_poisson2d": (C.c_int, [_i32, _i32, C.POINTER(_vp)]),
    "gs_problem_poisson3d": (C.c_int, [_i32, _i32, _i32, C.POINTER(_vp)]),
    "gs_problem_varcoef3d": (C.c_int, [_i32, C.c_uint64, C.c_double, C.c_double, C.POINTER(_vp)]),
    "gs_problem_elasticity": (C.c_int, [_i32, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, C.POINTER(_vp), _vp, _vp]),
    "gs_uniform_pm1": (C.c_int, [C.c_uint64, _i64, _f64p]),
    "gs_gaussian_field": (C.c_int, [_i32, _i32, _i32, C.c_uint64, C.c_double, _f64p]),
    "gs_hierarchy_build_sa": (C.c_int, [_vp, _i32, C.c_double, _i32, _i64, _vp, _i32,
                                        C.POINTER(_vp)]),
    "gs_hierarchy_build_sa_device": (C.c_int, [_vp, _i32, C.c_double, _i32, _i64, _vp, _i32,
                                               C.POINTER(_vp)]),
    "gs_csr_spgemm": (C.c_int, [_vp, _vp, _i32, C.POINTER(_vp)]),
    "gs_csr_transpose": (C.c_int, [_vp, _i32, C.POINTER(_vp)]),
    "gs_hierarchy_create": (C.c_int, [C.POINTER(_vp)]),
    "gs_problem_elasticity_mesh": (C.c_int, [_i64, _f64p, _i64, _i32p, C.c_double, C.c_double, _vp,
                                             C.c_double, C.c_double, C.c_double, C.POINTER(_vp),
                                             _f64p, _f64p]),
    "gs_csr_save": (C.c_int, [_vp, C.c_char_p]),
    "gs_csr_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "gs_hierarchy_save": (C.c_int, [_vp, C.c_char_p]),
    "gs_hierarchy_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "gs_hierarchy_push_level": (C.c_int, [_vp, _vp]),
    "gs_hierarchy_push_transfer": (C.c_int, [_vp, _vp, _vp]),
    "gs_hierarchy_levels": (C.c_int, [_vp, C.POINTER(_i32)]),
    "gs_hierarchy_get": (C.c_int, [_vp, _i32, _i32, C.POINTER(_vp)]),
    "gs_hierarchy_aggregates": (C.c_int, [_vp, _i32, _i32p]),
    "gs_hierarchy_level_rows": (C.c_int, [_vp, _i32, C.POINTER(_i64)]),
    "gs_hierarchical_reordering": (C.c_int, [_vp, _i32p, _i32p, _vp]),
    "gs_hierarchy_destroy": (C.c_int, [_vp]),
    "gs_solver_create": (C.c_int, [_i32, C.POINTER(_vp)]),
    "gs_solver_set_int": (C.c_int, [_vp, C.c_char_p, _i64]),
    "gs_solver_set_real": (C.c_int, [_vp, C.c_char_p, C.c_double]),
    "gs_solver_set_level_weights": (C.c_int, [_vp, _i32, _vp, _vp]),
    "gs_solver_setup": (C.c_int, [_vp, _vp, _vp]),
    "gs_solver_solve": (C.c_int, [_vp, _i64, _f64p, _i64, _f64p, C.c_double, _i64, C.POINTER(_i64),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "gs_solver_result": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_double)]),
    "gs_solver_history": (C.c_int, [_vp, _f64p, _i64, C.POINTER(_i64)]),
    "gs_solver_vcycle": (C.c_int, [_vp, _i64, _f64p, _f64p]),
    "gs_solver_spmv": (C.c_int, [_vp, _i64, _f64p, _f64p]),
    "gs_solver_size": (C.c_int, [_vp, C.POINTER(_i64)]),
    "gs_solver_time_kernel": (C.c_int, [_vp, _i32, _i32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
    "gs_solver_describe": (C.c_int, [_vp, C.c_char_p, _i64]),
    "gs_solver_destroy": (C.c_int, [_vp]),
}

_VP = C.POINTER(_vp)
_SIGS.update({
    "gs_backend_create": (C.c_int, [_i32, _VP]),
    "gs_backend_destroy": (C.c_int, [_vp]),
    "gs_backend_last_error": (C.c_char_p, []),
    "gs_backend_create_encoder": (C.c_int, [_vp, _VP]),
    "gs_backend_submit": (C.c_int, [_vp, _vp]),
    "gs_backend_submit_and_wait": (C.c_int, [_vp, _vp]),
    "gs_encoder_destroy": (C.c_int, [_vp]),
    "gs_encoder_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "gs_dvec_data": (C.c_int, [_vp, C.POINTER(_vp)]),
    "gs_dscalar_data": (C.c_int, [_vp, C.POINTER(_vp)]),
    "gs_dot_noalloc": (C.c_int, [_i32, _vp, _vp, _i64, _vp, _vp]),
    "gs_backend_create_vector": (C.c_int, [_vp, _i64, _vp, _VP]),
    "gs_dvec_size": (C.c_int, [_vp, C.POINTER(_i64)]),
    "gs_dvec_update": (C.c_int, [_vp, _vp, _i64]),
    "gs_dvec_download": (C.c_int, [_vp, _vp]),
    "gs_dvec_destroy": (C.c_int, [_vp]),
    "gs_backend_create_index_vector": (C.c_int, [_vp, _i64, _i32p, _VP]),
    "gs_divec_download": (C.c_int, [_vp, _i32p]),
    "gs_divec_destroy": (C.c_int, [_vp]),
    "gs_backend_create_scalar": (C.c_int, [_vp, C.c_double, _VP]),
    "gs_dscalar_download": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "gs_dscalar_destroy": (C.c_int, [_vp]),
    "gs_backend_create_csr": (C.c_int, [_vp, _vp, _VP]),
    "gs_dcsr_update_values": (C.c_int, [_vp, _f64p, _i64]),
    "gs_dcsr_destroy": (C.c_int, [_vp]),
    "gs_backend_create_ell": (C.c_int, [_vp, _vp, _i32, _VP]),
    "gs_dell_destroy": (C.c_int, [_vp]),
    "gs_divec_size": (C.c_int, [_vp, C.POINTER(_i64)]),
    "gs_divec_update": (C.c_int, [_vp, _i32p, _i64]),
    "gs_dcsr_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "gs_backend_create_ell_packed": (C.c_int, [_vp, _i64, _i64, _i64, _i32p, _vp, _i64, _i32p,
                                               _vp, _vp, _VP]),
    "gs_dell_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                               C.POINTER(_i64)]),
    "gs_encode_scale": (C.c_int, [_vp, _vp, C.c_double]),
    "gs_encode_scale_dev": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_copy": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_scale_vector": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_set_zero": (C.c_int, [_vp, _vp]),
    "gs_encode_axpy": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "gs_encode_axpy_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_spmv": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_spmv_ell": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_gauss_seidel_color": (C.c_int, [_vp, _vp, _vp, _i64, _i64, _vp, _vp, C.c_double]),
    "gs_encode_dot": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_scalar_divide": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_scalar_multiply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_scalar_set": (C.c_int, [_vp, C.c_double, _vp]),
    "gs_encode_scalar_copy": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_scalar_negate": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_scalar_sqrt": (C.c_int, [_vp, _vp, _vp]),
    "gs_encode_apply_block_inverses3x3": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_encode_block_gauss_seidel_color3x3": (C.c_int, [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp,
                                                        C.c_double]),
})

_SIGS.update({
    "gs_solver_set_fallback": (C.c_int, [_vp, _vp, C.c_double]),
    "gs_csr_sell_layout": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_i32)]),
    "gs_hierarchy_features": (C.c_int, [_vp, _i32, _f64p]),
    "gs_mlp_create": (C.c_int, [_i32, _i32, _i32, C.c_uint64, _VP]),
    "gs_mlp_num_params": (C.c_int, [_vp, C.POINTER(_i64)]),
    "gs_mlp_get_params": (C.c_int, [_vp, _f64p]),
    "gs_mlp_set_params": (C.c_int, [_vp, _f64p]),
    "gs_mlp_forward": (C.c_int, [_vp, _f64p, _i64, _f64p]),
    "gs_mlp_loss_grad": (C.c_int, [_vp, _f64p, _f64p, _i64, C.POINTER(C.c_double), _f64p]),
    "gs_mlp_adam_step": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "gs_mlp_train": (C.c_int, [_vp, _f64p, _f64p, _i64, _i64, C.c_double,
                               C.POINTER(C.c_double)]),
    "gs_mlp_destroy": (C.c_int, [_vp]),
})

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def check(rc: int) -> None:
    if rc != 0:
        raise GpuSolverError(lib.gs_last_error().decode())


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


# ---------------------------------------------------------------------------
# CSR
# ---------------------------------------------------------------------------
@dataclass
class Csr:
    """CSR matrix with int32 indices and FP64 values (host)."""
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    shape: tuple

    @property
    def nnz(self) -> int:
        return int(self.col.size)

    @property
    def rows(self) -> int:
        return int(self.shape[0])

    @classmethod
    def from_scipy(cls, m) -> "Csr":
        m = m.tocsr()
        return cls(np.ascontiguousarray(m.indptr, np.int32), np.ascontiguousarray(m.indices, np.int32),
                   np.ascontiguousarray(m.data, np.float64), tuple(m.shape))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.ptr), shape=self.shape)

    def handle(self) -> "CsrHandle":
        h = _vp()
        check(lib.gs_csr_create(self.shape[0], self.shape[1], self.nnz,
                                np.ascontiguousarray(self.ptr, np.int32),
                                np.ascontiguousarray(self.col, np.int32),
                                np.ascontiguousarray(self.val, np.float64), C.byref(h)))
        return CsrHandle(h)

    def save(self, path) -> None:
        """Binary CSR file (gs_csr_save)."""
        hd = self.handle()  # keeps the C object alive during the call
        check(lib.gs_csr_save(hd.h, os.fsencode(path)))


def load_csr(path) -> Csr:
    """Reads a file written by Csr.save (gs_csr_load)."""
    return _from_handle_fn(lib.gs_csr_load, os.fsencode(path))


class CsrHandle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.gs_csr_destroy(self.h)
            self.h = None

    def to_csr(self) -> Csr:
        r, c, z = _i64(), _i64(), _i64()
        check(lib.gs_csr_info(self.h, C.byref(r), C.byref(c), C.byref(z)))
        ptr = np.empty(r.value + 1, np.int32)
        col = np.empty(max(z.value, 1), np.int32)
        val = np.empty(max(z.value, 1), np.float64)
        check(lib.gs_csr_export(self.h, ptr, col, val))
        return Csr(ptr, col[:z.value].copy(), val[:z.value].copy(), (r.value, c.value))


def _from_handle_fn(fn, *args) -> Csr:
    h = _vp()
    check(fn(*args, C.byref(h)))
    return CsrHandle(h).to_csr()


def poisson2d(nx: int, ny: int) -> Csr:
    return _from_handle_fn(lib.gs_problem_poisson2d, nx, ny)


def poisson3d(nx: int, ny: int | None = None, nz: int | None = None) -> Csr:
    return _from_handle_fn(lib.gs_problem_poisson3d, nx, ny or nx, nz or nx)


def varcoef3d(n: int, seed: int, sigma: float = 8.0, scale: float = 2.0) -> Csr:
    return _from_handle_fn(lib.gs_problem_varcoef3d, n, seed, sigma, scale)


def elasticity(nv: int, young: float = 1.0, poisson: float = 0.3, hetero_sigma: float = 0.0,
               hetero_scale: float = 1.5, seed: int = 0):
    """Returns (A, rhs, nullspace[n x 6])."""
    h = _vp()
    # Size the outputs: free DOFs = 3 * nv^2 * (nv - 1).
    n = 3 * nv * nv * (nv - 1)
    rhs = np.empty(n, np.float64)
    ns = np.empty(n * 6, np.float64)
    check(lib.gs_problem_elasticity(nv, young, poisson, hetero_sigma, hetero_scale, seed,
                                    C.byref(h), rhs.ctypes.data_as(_vp), ns.ctypes.data_as(_vp)))
    A = CsrHandle(h).to_csr()
    assert A.rows == n
    return A, rhs, ns.reshape(n, 6)


def elasticity_mesh(V, T, young: float = 1.0e5, poisson: float = 0.35, young_per_tet=None,
                    clamp_rel: float = 1.0e-4, density: float = 1.0, gravity: float = -9.81):
    """The cantilever driver's system on a tet mesh (V: n x 3 vertices, T:
    m x 4 zero-based tets; gs_problem_elasticity_mesh). Defaults are the
    driver's (examples/elasticity/cantileverBeam.cc:291-348). Returns
    (A, rhs, nullspace[n x 6])."""
    V = np.ascontiguousarray(V, np.float64)
    T = np.ascontiguousarray(T, np.int32)
    if V.ndim != 2 or V.shape[1] != 3 or T.ndim != 2 or T.shape[1] != 4:
        raise ValueError("V must be n x 3 and T m x 4")
    ypt = None
    if young_per_tet is not None:
        ypt = np.ascontiguousarray(young_per_tet, np.float64)
        if ypt.shape != (T.shape[0],):
            raise ValueError("young_per_tet must hold one value per tet")
    rhs = np.empty(3 * V.shape[0], np.float64)
    ns = np.empty(18 * V.shape[0], np.float64)
    h = _vp()
    check(lib.gs_problem_elasticity_mesh(V.shape[0], V, T.shape[0], T, young, poisson,
                                         None if ypt is None else ypt.ctypes.data_as(_vp),
                                         clamp_rel, density, gravity, C.byref(h), rhs, ns))
    A = CsrHandle(h).to_csr()
    n = A.rows
    return A, rhs[:n].copy(), ns[:6 * n].reshape(n, 6).copy()


def read_mesh(path):
    """Tetrahedral mesh in the MEDIT ASCII format (.mesh: the format the
    reference's drivers load with igl::readMESH): returns V (n x 3), T (m x 4,
    zero-based) and the boundary triangles F (k x 3, zero-based)."""
    with open(path) as fh:
        tok = [t for line in fh for t in line.split("#", 1)[0].split()]
    V = np.zeros((0, 3))
    T = np.zeros((0, 4), np.int32)
    F = np.zeros((0, 3), np.int32)
    i = 0
    while i < len(tok):
        key = tok[i]
        i += 1
        if key in ("MeshVersionFormatted", "Dimension"):
            i += 1
        elif key in ("Vertices", "Triangles", "Tetrahedra", "Edges", "Corners", "Ridges",
                     "RequiredVertices"):
            n = int(tok[i])
            i += 1
            width = {"Vertices": 4, "Triangles": 4, "Tetrahedra": 5, "Edges": 3}.get(key, 1)
            if n < 0 or i + n * width > len(tok):
                raise ValueError(f"{path}: truncated {key} section")
            rows = np.array(tok[i:i + n * width], dtype=np.float64).reshape(n, width)
            i += n * width
            if key == "Vertices":
                V = rows[:, :3].copy()
            elif key == "Triangles":
                F = rows[:, :3].astype(np.int32) - 1
            elif key == "Tetrahedra":
                T = rows[:, :4].astype(np.int32) - 1
        elif key == "End":
            break
        else:
            raise ValueError(f"{path}: unsupported section {key!r}")
    if T.size and (T.min() < 0 or T.max() >= len(V)):
        raise ValueError(f"{path}: tetrahedron vertex index out of range")
    return V, T, F


def write_mesh(path, V, T, F=None) -> None:
    """Writes V, T (zero-based) and optional triangles F in the MEDIT ASCII
    format read by read_mesh."""
    with open(path, "w") as fh:
        fh.write("MeshVersionFormatted 1\nDimension 3\n")
        fh.write(f"Vertices\n{len(V)}\n")
        for x, y, z in np.asarray(V, np.float64).tolist():
            fh.write(f"{x!r} {y!r} {z!r} 0\n")
        if F is not None and len(F):
            fh.write(f"Triangles\n{len(F)}\n")
            for a, b, c in (np.asarray(F) + 1).tolist():
                fh.write(f"{a} {b} {c} 0\n")
        fh.write(f"Tetrahedra\n{len(T)}\n")
        for a, b, c, d in (np.asarray(T) + 1).tolist():
            fh.write(f"{a} {b} {c} {d} 0\n")
        fh.write("End\n")


def uniform_pm1(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    check(lib.gs_uniform_pm1(seed, n, out))
    return out


def gaussian_field(nx: int, ny: int, nz: int, seed: int, sigma: float) -> np.ndarray:
    out = np.empty(nx * ny * nz, np.float64)
    check(lib.gs_gaussian_field(nx, ny, nz, seed, sigma, out))
    return out


# ---------------------------------------------------------------------------
# Setup artifacts
# ---------------------------------------------------------------------------
def spgemm(A: Csr, B: Csr, device: bool = False) -> Csr:
    """C = A B, rows sorted (gs_csr_spgemm); host or GPU, same bits."""
    ha, hb = A.handle(), B.handle()  # alive for the call
    return _from_handle_fn(lib.gs_csr_spgemm, ha.h, hb.h, int(device))


def transpose(A: Csr, device: bool = False) -> Csr:
    """A^T, rows sorted (gs_csr_transpose); host or GPU, same bits."""
    ha = A.handle()  # alive for the call
    return _from_handle_fn(lib.gs_csr_transpose, ha.h, int(device))


def sell_layout(A: Csr):
    """(slots, min column, max column) of A packed as SELL-64; raises when
    the packed layout would break a kernel's assumptions."""
    h = A.handle()
    slots, lo, hi = _i64(), _i32(), _i32()
    check(lib.gs_csr_sell_layout(h.h, C.byref(slots), C.byref(lo), C.byref(hi)))
    return slots.value, lo.value, hi.value


def graph_coloring(A: Csr):
    n = A.rows
    colors = np.empty(n, np.int32)
    verts = np.empty(n, np.int32)
    offs = np.empty(n + 1, np.int32)
    nc = _i32()
    check(lib.gs_graph_coloring(n, A.ptr, A.col, colors, verts, offs, C.byref(nc)))
    return colors, verts, offs[:nc.value + 1].copy(), nc.value


def rcm(A: Csr):
    n = A.rows
    o2n = np.empty(n, np.int32)
    n2o = np.empty(n, np.int32)
    check(lib.gs_rcm(n, A.ptr, A.col, o2n, n2o))
    return o2n, n2o


def permutation_check(o2n, n2o) -> None:
    check(lib.gs_permutation_check(len(o2n), np.ascontiguousarray(o2n, np.int32),
                                   np.ascontiguousarray(n2o, np.int32)))


def permute_keep_order(A: Csr, o2n) -> Csr:
    h = A.handle()
    return _from_handle_fn(lib.gs_csr_permute_keep_order, h.h, np.ascontiguousarray(o2n, np.int32))


def permute_sorted(A: Csr, row_o2n, col_o2n) -> Csr:
    h = A.handle()
    return _from_handle_fn(lib.gs_csr_permute_sorted, h.h, np.ascontiguousarray(row_o2n, np.int32),
                           np.ascontiguousarray(col_o2n, np.int32))


def ell_pack(A: Csr, width: int):
    h = A.handle()
    ov = _i64()
    check(lib.gs_ell_pack(h.h, width, None, None, None, None, None, C.byref(ov)))
    n = A.rows
    ec = np.empty(n * width, np.int32)
    ev = np.empty(n * width, np.float64)
    op = np.empty(n + 1, np.int32)
    oc = np.empty(max(ov.value, 1), np.int32)
    ovv = np.empty(max(ov.value, 1), np.float64)
    check(lib.gs_ell_pack(h.h, width, ec.ctypes.data_as(_vp), ev.ctypes.data_as(_vp),
                          op.ctypes.data_as(_vp), oc.ctypes.data_as(_vp), ovv.ctypes.data_as(_vp),
                          C.byref(ov)))
    return ec, ev, op, oc[:ov.value].copy(), ovv[:ov.value].copy()


def block_graph(A: Csr, bs: int = 3):
    h = A.handle()
    nnz = _i64()
    check(lib.gs_block_graph(h.h, bs, None, None, C.byref(nnz)))
    gp = np.empty(A.rows // bs + 1, np.int32)
    gc = np.empty(max(nnz.value, 1), np.int32)
    check(lib.gs_block_graph(h.h, bs, gp.ctypes.data_as(_vp), gc.ctypes.data_as(_vp), C.byref(nnz)))
    return gp, gc[:nnz.value].copy()


def expand_block_permutation(node_o2n, bs: int = 3):
    nodes = len(node_o2n)
    o2n = np.empty(nodes * bs, np.int32)
    n2o = np.empty(nodes * bs, np.int32)
    check(lib.gs_expand_block_permutation(nodes, bs, np.ascontiguousarray(node_o2n, np.int32),
                                          o2n, n2o))
    return o2n, n2o


def inverse_diag_blocks3(A: Csr) -> np.ndarray:
    h = A.handle()
    out = np.empty(9 * (A.rows // 3), np.float64)
    check(lib.gs_inverse_diag_blocks3(h.h, out))
    return out


def dense_spd_inverse(a: np.ndarray) -> np.ndarray:
    """FP64 inverse of a dense SPD matrix through its Cholesky factor (the
    mixed mode's coarsest-level direct solve)."""
    m = np.ascontiguousarray(a, np.float64).copy()
    n = m.shape[0]
    if m.shape != (n, n):
        raise ValueError("square matrix expected")
    check(lib.gs_dense_spd_inverse(n, m.reshape(-1)))
    return m


def dense_spd_inverse_device(a: np.ndarray) -> np.ndarray:
    """The same inverse built on the GPU (blocked Gauss-Jordan, FP64)."""
    m = np.ascontiguousarray(a, np.float64).copy()
    n = m.shape[0]
    if m.shape != (n, n):
        raise ValueError("square matrix expected")
    check(lib.gs_dense_spd_inverse_device(n, m.reshape(-1)))
    return m


# ---------------------------------------------------------------------------
# Hierarchy
# ---------------------------------------------------------------------------
class Hierarchy:
    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib.gs_hierarchy_destroy(self.h)
            self.h = None

    @classmethod
    def smoothed_aggregation(cls, A: Csr, block_size: int = 1, eps_strong: float = 0.1,
                             max_levels: int = 10, min_dofs: int = 1000,
                             nullspace: np.ndarray | None = None,
                             device: bool = False) -> "Hierarchy":
        """SA hierarchy (gs_hierarchy_build_sa); device=True forms the
        products of each coarsening step on the GPU (the same hierarchy bit
        for bit, gs_hierarchy_build_sa_device)."""
        ah = A.handle()
        h = _vp()
        ncols = 0
        nsp = None
        if nullspace is not None:
            ns = np.ascontiguousarray(nullspace, np.float64)
            ncols = ns.shape[1]
            nsp = ns.ctypes.data_as(_vp)
            keep = ns  # noqa: F841  (keeps the buffer alive during the call)
        build = lib.gs_hierarchy_build_sa_device if device else lib.gs_hierarchy_build_sa
        check(build(ah.h, block_size, eps_strong, max_levels, min_dofs, nsp, ncols, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path) -> "Hierarchy":
        """Reads a file written by save (gs_hierarchy_load)."""
        h = _vp()
        check(lib.gs_hierarchy_load(os.fsencode(path), C.byref(h)))
        return cls(h)

    def save(self, path) -> None:
        """Binary file of A, P, R and the aggregate maps (gs_hierarchy_save)."""
        check(lib.gs_hierarchy_save(self.h, os.fsencode(path)))

    @classmethod
    def from_levels(cls, A: list, P: list, R: list) -> "Hierarchy":
        h = _vp()
        check(lib.gs_hierarchy_create(C.byref(h)))
        obj = cls(h)
        for a in A:
            ah = a.handle()
            check(lib.gs_hierarchy_push_level(h, ah.h))
        for p, r in zip(P, R):
            ph, rh = p.handle(), r.handle()
            check(lib.gs_hierarchy_push_transfer(h, ph.h, rh.h))
        return obj

    @property
    def levels(self) -> int:
        n = _i32()
        check(lib.gs_hierarchy_levels(self.h, C.byref(n)))
        return n.value

    def get(self, which: str, k: int) -> Csr:
        w = {"A": 0, "P": 1, "R": 2}[which]
        return _from_handle_fn(lib.gs_hierarchy_get, self.h, w, k)

    def features(self, block_size: int = 1) -> np.ndarray:
        """Per-level features (levels x 5) for the learned smoother weights."""
        out = np.empty(5 * self.levels, np.float64)
        check(lib.gs_hierarchy_features(self.h, block_size, out))
        return out.reshape(self.levels, 5)

    def level_rows(self, k: int) -> int:
        n = _i64()
        check(lib.gs_hierarchy_level_rows(self.h, k, C.byref(n)))
        return n.value

    def aggregates(self, k: int) -> np.ndarray:
        n = self.level_rows(k)
        out = np.empty(n, np.int32)
        check(lib.gs_hierarchy_aggregates(self.h, k, out))
        return out


def hierarchical_reordering(H: Hierarchy):
    """HierarchicalReordering::compute
    (src/ReorderingStrategies/HierarchicalReordering.cc:217-278): returns the
    list of per-level (old_to_new, new_to_old) pairs, finest first, and the
    per-transfer counts of fine DOFs without a parent in R (which the
    reference prints)."""
    sizes = [H.level_rows(k) for k in range(H.levels)]
    total = int(sum(sizes))
    o2n = np.empty(total, np.int32)
    n2o = np.empty(total, np.int32)
    orph = np.zeros(max(len(sizes) - 1, 1), np.int64)
    check(lib.gs_hierarchical_reordering(H.h, o2n, n2o, orph.ctypes.data_as(_vp)))
    perms, off = [], 0
    for n in sizes:
        perms.append((o2n[off:off + n].copy(), n2o[off:off + n].copy()))
        off += n
    return perms, orph[:len(sizes) - 1].copy()


# ---------------------------------------------------------------------------
# Solver
# ---------------------------------------------------------------------------
@dataclass
class SolveInfo:
    iterations: int
    relres2: float
    true_relres: float
    solve_ms: float
    converged: bool
    loop_iterations: int
    breakdown: int
    wall_ms: float


_REAL_OPTIONS = ("omega", "group_min_scalar", "group_min_block", "full_group_min_scalar",
                 "full_group_min_block")


class Solver:
    """CGSolver + preconditioner on the GPU (gs_solver_* in include/gpusolver.h)."""

    def __init__(self, precision: int = PREC_DOUBLE, **opts):
        h = _vp()
        check(lib.gs_solver_create(precision, C.byref(h)))
        self.h = h
        self.precision = precision
        self.n = 0
        for k, v in opts.items():
            self.set(k, v)

    def __del__(self):
        if getattr(self, "h", None):
            lib.gs_solver_destroy(self.h)
            self.h = None

    def set(self, key: str, value) -> None:
        if key in _REAL_OPTIONS:
            check(lib.gs_solver_set_real(self.h, key.encode(), float(value)))
        else:
            check(lib.gs_solver_set_int(self.h, key.encode(), int(value)))

    def set_level_weights(self, omega=None, sweeps=None) -> None:
        nl = len(omega) if omega is not None else len(sweeps)
        om = np.ascontiguousarray(omega, np.float64) if omega is not None else None
        sw = np.ascontiguousarray(sweeps, np.int32) if sweeps is not None else None
        check(lib.gs_solver_set_level_weights(self.h, nl,
                                              om.ctypes.data_as(_vp) if om is not None else None,
                                              sw.ctypes.data_as(_vp) if sw is not None else None))

    def setup(self, hier: Hierarchy, perms: list | None = None) -> None:
        arr = None
        keep = []
        if perms is not None:
            arr = (_vp * len(perms))()
            for i, p in enumerate(perms):
                if p is None:
                    arr[i] = None
                else:
                    a = np.ascontiguousarray(p, np.int32)
                    keep.append(a)
                    arr[i] = a.ctypes.data_as(_vp)
        check(lib.gs_solver_setup(self.h, hier.h, C.cast(arr, _vp) if arr is not None else None))
        self.n = hier.level_rows(0)  # (get() would copy the whole level-0 matrix)

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None, tol: float = 1e-8,
              maxit: int = 1000):
        b = np.ascontiguousarray(b, np.float64).reshape(-1)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True).reshape(-1)
        it = _i64()
        rr, tr, ms = C.c_double(), C.c_double(), C.c_double()
        check(lib.gs_solver_solve(self.h, b.size, b, x.size, x, tol, maxit, C.byref(it),
                                  C.byref(rr), C.byref(tr), C.byref(ms)))
        res = {}
        for k in ("converged", "loop_iterations", "breakdown", "wall_ms"):
            v = C.c_double()
            check(lib.gs_solver_result(self.h, k.encode(), C.byref(v)))
            res[k] = v.value
        return x, SolveInfo(it.value, rr.value, tr.value, ms.value, bool(res["converged"]),
                            int(res["loop_iterations"]), int(res["breakdown"]), res["wall_ms"])

    def set_fallback(self, plain: "Solver", lag: float = 1.5) -> None:
        self._fallback = plain  # keep alive
        check(lib.gs_solver_set_fallback(self.h, plain.h if plain is not None else None, lag))

    def fell_back(self) -> bool:
        v = C.c_double()
        check(lib.gs_solver_result(self.h, b"fell_back", C.byref(v)))
        return bool(v.value)

    def history(self) -> np.ndarray:
        n = _i64()
        check(lib.gs_solver_history(self.h, np.empty(1, np.float64), 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.float64)
        check(lib.gs_solver_history(self.h, out, n.value, C.byref(n)))
        return out[:n.value]

    def vcycle(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float64).reshape(-1)
        z = np.empty_like(r)
        check(lib.gs_solver_vcycle(self.h, r.size, r, z))
        return z

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.empty_like(x)
        check(lib.gs_solver_spmv(self.h, x.size, x, y))
        return y

    def size(self) -> int:
        n = _i64()
        check(lib.gs_solver_size(self.h, C.byref(n)))
        return n.value

    def result(self, key: str) -> float:
        v = C.c_double()
        check(lib.gs_solver_result(self.h, key.encode(), C.byref(v)))
        return v.value

    def time_kernel(self, which: int, reps: int = 20):
        ms, by = C.c_double(), C.c_double()
        check(lib.gs_solver_time_kernel(self.h, which, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def describe(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        check(lib.gs_solver_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device query only)
        return bool(torch.cuda.is_available())
    except Exception:
        return os.path.exists("/dev/kfd")


# ---------------------------------------------------------------------------
# Backend plugin (include/GPUSolver/Backend.h) on HIP
# ---------------------------------------------------------------------------
class _Obj:
    _destroy = None

    def __init__(self, h, owner=None):
        self.h = h
        self.owner = owner  # keeps the backend alive

    def __del__(self):
        if getattr(self, "h", None) and self._destroy:
            getattr(lib, self._destroy)(self.h)
            self.h = None


class DeviceVector(_Obj):
    _destroy = "gs_dvec_destroy"

    def size(self) -> int:
        n = _i64()
        check(lib.gs_dvec_size(self.h, C.byref(n)))
        return n.value

    def update_values(self, values) -> None:
        a = np.ascontiguousarray(values, self.owner.np_dtype)
        check(lib.gs_dvec_update(self.h, a.ctypes.data_as(_vp), a.size))

    def download(self) -> np.ndarray:
        out = np.empty(self.size(), self.owner.np_dtype)
        check(lib.gs_dvec_download(self.h, out.ctypes.data_as(_vp)))
        return out

    def data(self) -> int:
        """Device address of the values (CudaDeviceVector::data)."""
        p = _vp()
        check(lib.gs_dvec_data(self.h, C.byref(p)))
        return p.value or 0


class DeviceIndexVector(_Obj):
    _destroy = "gs_divec_destroy"

    def __init__(self, h, n, owner=None):
        super().__init__(h, owner)
        self.n = n

    def download(self) -> np.ndarray:
        out = np.empty(self.n, np.int32)
        check(lib.gs_divec_download(self.h, out))
        return out

    def size(self) -> int:
        n = _i64()
        check(lib.gs_divec_size(self.h, C.byref(n)))
        return n.value

    def update_values(self, values) -> None:
        a = np.ascontiguousarray(values, np.int32)
        check(lib.gs_divec_update(self.h, a, a.size))


class DeviceScalar(_Obj):
    _destroy = "gs_dscalar_destroy"

    def download(self) -> float:
        v = C.c_double()
        check(lib.gs_dscalar_download(self.h, C.byref(v)))
        return v.value

    def data(self) -> int:
        """Device address of the scalar (CudaDeviceScalar::data)."""
        p = _vp()
        check(lib.gs_dscalar_data(self.h, C.byref(p)))
        return p.value or 0


class DeviceCSRMatrix(_Obj):
    _destroy = "gs_dcsr_destroy"

    def update_values(self, values) -> None:
        a = np.ascontiguousarray(values, np.float64)
        check(lib.gs_dcsr_update_values(self.h, a, a.size))

    def info(self) -> tuple:
        """(rows, cols, nnz)."""
        v = [_i64() for _ in range(3)]
        check(lib.gs_dcsr_info(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)


class DeviceELLMatrix(_Obj):
    _destroy = "gs_dell_destroy"

    def info(self) -> tuple:
        """(rows, cols, ellWidth, overflowNnz)."""
        v = [_i64() for _ in range(4)]
        check(lib.gs_dell_info(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)


class Encoder(_Obj):
    _destroy = "gs_encoder_destroy"

    def stream(self) -> int:
        """The encoder's hipStream_t (CudaEncoder::stream)."""
        p = _vp()
        check(lib.gs_encoder_stream(self.h, C.byref(p)))
        return p.value or 0


class Backend:
    """gpuSolver::Backend on HIP: create*, encode*, submit / submitAndWait."""

    def __init__(self, dtype=np.float32):
        self.np_dtype = np.dtype(dtype)
        code = 1 if self.np_dtype == np.float64 else 0
        h = _vp()
        check(lib.gs_backend_create(code, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.gs_backend_destroy(self.h)
            self.h = None

    # factories (Backend.h:19-41)
    def create_encoder(self) -> Encoder:
        h = _vp()
        check(lib.gs_backend_create_encoder(self.h, C.byref(h)))
        return Encoder(h, self)

    def submit(self, e: Encoder) -> None:
        check(lib.gs_backend_submit(self.h, e.h))

    def submit_and_wait(self, e: Encoder) -> None:
        check(lib.gs_backend_submit_and_wait(self.h, e.h))

    def create_vector(self, n_or_values) -> DeviceVector:
        h = _vp()
        if np.isscalar(n_or_values):
            check(lib.gs_backend_create_vector(self.h, int(n_or_values), None, C.byref(h)))
        else:
            a = np.ascontiguousarray(n_or_values, self.np_dtype)
            check(lib.gs_backend_create_vector(self.h, a.size, a.ctypes.data_as(_vp), C.byref(h)))
        return DeviceVector(h, self)

    def create_index_vector(self, values) -> DeviceIndexVector:
        a = np.ascontiguousarray(values, np.int32)
        h = _vp()
        check(lib.gs_backend_create_index_vector(self.h, a.size, a, C.byref(h)))
        return DeviceIndexVector(h, a.size, self)

    def create_scalar(self, value: float = 0.0) -> DeviceScalar:
        h = _vp()
        check(lib.gs_backend_create_scalar(self.h, float(value), C.byref(h)))
        return DeviceScalar(h, self)

    def create_csr_matrix(self, A: Csr) -> DeviceCSRMatrix:
        ah = A.handle()
        h = _vp()
        check(lib.gs_backend_create_csr(self.h, ah.h, C.byref(h)))
        return DeviceCSRMatrix(h, self)

    def create_ell_matrix(self, A: Csr, width: int) -> DeviceELLMatrix:
        ah = A.handle()
        h = _vp()
        check(lib.gs_backend_create_ell(self.h, ah.h, width, C.byref(h)))
        return DeviceELLMatrix(h, self)

    def create_ell_packed(self, rows: int, cols: int, width: int, ell_col, ell_val, ovf_ptr,
                          ovf_col, ovf_val) -> DeviceELLMatrix:
        """An ELL matrix from arrays packed on the host (slot-major ELL, then
        the overflow CSR), values in the backend's dtype."""
        ec = np.ascontiguousarray(ell_col, np.int32)
        ev = np.ascontiguousarray(ell_val, self.np_dtype)
        op = np.ascontiguousarray(ovf_ptr, np.int32)
        oc = np.ascontiguousarray(ovf_col, np.int32)
        ov = np.ascontiguousarray(ovf_val, self.np_dtype)
        h = _vp()
        check(lib.gs_backend_create_ell_packed(
            self.h, rows, cols, width, ec, ev.ctypes.data, oc.size, op,
            oc.ctypes.data if oc.size else None, ov.ctypes.data if ov.size else None,
            C.byref(h)))
        return DeviceELLMatrix(h, self)

    # encode* (Backend.h:47-114)
    def encode_scale(self, e, x, alpha):
        if isinstance(alpha, DeviceScalar):
            check(lib.gs_encode_scale_dev(e.h, x.h, alpha.h))
        else:
            check(lib.gs_encode_scale(e.h, x.h, float(alpha)))

    def encode_copy(self, e, x, y):
        check(lib.gs_encode_copy(e.h, x.h, y.h))

    def encode_scale_vector(self, e, s, out):
        check(lib.gs_encode_scale_vector(e.h, s.h, out.h))

    def encode_set_zero(self, e, x):
        check(lib.gs_encode_set_zero(e.h, x.h))

    def encode_axpy(self, e, alpha, x, y):
        if isinstance(alpha, DeviceScalar):
            check(lib.gs_encode_axpy_dev(e.h, alpha.h, x.h, y.h))
        else:
            check(lib.gs_encode_axpy(e.h, float(alpha), x.h, y.h))

    def encode_spmv(self, e, A, x, y):
        if isinstance(A, DeviceELLMatrix):
            check(lib.gs_encode_spmv_ell(e.h, A.h, x.h, y.h))
        else:
            check(lib.gs_encode_spmv(e.h, A.h, x.h, y.h))

    def encode_gauss_seidel_color(self, e, A, colors, start, count, rhs, x, omega):
        check(lib.gs_encode_gauss_seidel_color(e.h, A.h, colors.h, start, count, rhs.h, x.h,
                                               float(omega)))

    def encode_dot(self, e, x, y, out):
        check(lib.gs_encode_dot(e.h, x.h, y.h, out.h))

    def encode_scalar_divide(self, e, a, b, out):
        check(lib.gs_encode_scalar_divide(e.h, a.h, b.h, out.h))

    def encode_scalar_multiply(self, e, a, b, out):
        check(lib.gs_encode_scalar_multiply(e.h, a.h, b.h, out.h))

    def encode_scalar_set(self, e, value, out):
        check(lib.gs_encode_scalar_set(e.h, float(value), out.h))

    def encode_scalar_copy(self, e, a, out):
        check(lib.gs_encode_scalar_copy(e.h, a.h, out.h))

    def encode_scalar_negate(self, e, a, out):
        check(lib.gs_encode_scalar_negate(e.h, a.h, out.h))

    def encode_scalar_sqrt(self, e, a, out):
        check(lib.gs_encode_scalar_sqrt(e.h, a.h, out.h))

    def encode_apply_block_inverses3x3(self, e, inv, res, corr):
        check(lib.gs_encode_apply_block_inverses3x3(e.h, inv.h, res.h, corr.h))

    def encode_block_gauss_seidel_color3x3(self, e, A, blocks, dinv, begin, end, rhs, x, omega):
        check(lib.gs_encode_block_gauss_seidel_color3x3(e.h, A.h, blocks.h, dinv.h, begin, end,
                                                        rhs.h, x.h, float(omega)))

    # blocking convenience methods (Backend.h:122-155)
    def _blocking(self, fn, *args):
        e = self.create_encoder()
        fn(e, *args)
        self.submit_and_wait(e)

    def spmv(self, A, x, y):
        self._blocking(self.encode_spmv, A, x, y)

    def dot(self, x, y, out):
        self._blocking(self.encode_dot, x, y, out)

    def axpy(self, alpha, x, y):
        self._blocking(self.encode_axpy, alpha, x, y)

    def scale(self, x, alpha):
        self._blocking(self.encode_scale, x, alpha)

    def scale_vector_by_vector(self, s, out):
        self._blocking(self.encode_scale_vector, s, out)

    def scalar_set(self, value, out):
        self._blocking(self.encode_scalar_set, value, out)

    def scalar_copy(self, a, out):
        self._blocking(self.encode_scalar_copy, a, out)

    def scalar_divide(self, a, b, out):
        self._blocking(self.encode_scalar_divide, a, b, out)

    def scalar_multiply(self, a, b, out):
        self._blocking(self.encode_scalar_multiply, a, b, out)

    def scalar_negate(self, a, out):
        self._blocking(self.encode_scalar_negate, a, out)

    def scalar_sqrt(self, a, out):
        self._blocking(self.encode_scalar_sqrt, a, out)


class MLP:
    """Smoother-weight MLP on the GPU (gs_mlp_* in include/gpusolver.h)."""

    def __init__(self, inputs: int = 5, hidden: int = 64, outputs: int = 2, seed: int = 0):
        h = _vp()
        check(lib.gs_mlp_create(inputs, hidden, outputs, seed, C.byref(h)))
        self.h = h
        self.shape = (inputs, hidden, outputs)

    def __del__(self):
        if getattr(self, "h", None):
            lib.gs_mlp_destroy(self.h)
            self.h = None

    @property
    def num_params(self) -> int:
        n = _i64()
        check(lib.gs_mlp_num_params(self.h, C.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float64)
        check(lib.gs_mlp_get_params(self.h, out))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float64)
        assert p.size == self.num_params
        check(lib.gs_mlp_set_params(self.h, p))

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float64)
        Y = np.empty((X.shape[0], self.shape[2]), np.float64)
        check(lib.gs_mlp_forward(self.h, X, X.shape[0], Y))
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        l = C.c_double()
        g = np.empty(self.num_params, np.float64)
        check(lib.gs_mlp_loss_grad(self.h, X, T, X.shape[0], C.byref(l), g))
        return l.value, g

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        check(lib.gs_mlp_adam_step(self.h, lr, beta1, beta2, eps))

    def train(self, X, T, epochs: int, lr: float = 3e-3) -> float:
        X = np.ascontiguousarray(X, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        l = C.c_double()
        check(lib.gs_mlp_train(self.h, X, T, X.shape[0], epochs, lr, C.byref(l)))
        return l.value
