"""The C-ABI library loads without a GPU, exports every symbol declared in
include/*.h, and reports errors the way the reference throws them. CPU
only: no call here touches the device."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpusolver as gs

REPO = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(gs.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_every_declared_symbol_is_exported():
    decl = declared_symbols()
    assert len(decl) > 40
    missing = [d for d in decl if d not in exported_symbols()]
    assert not missing, missing


def test_python_binding_covers_header():
    bound = set(gs.exported_symbols())
    missing = [d for d in declared_symbols() if d not in bound]
    assert not missing, missing


def test_error_messages_follow_reference():
    with pytest.raises(gs.GpuSolverError, match="matrix dimensions must be nonzero"):
        gs.Csr(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), (0, 0)).handle()
    with pytest.raises(gs.GpuSolverError, match="invalid CSR row pointer"):
        gs.Csr(np.array([0, 3, 2], np.int32), np.array([0, 1], np.int32), np.ones(2), (2, 2)).handle()
    lib = gs.lib
    s = C.c_void_p()
    assert lib.gs_solver_create(7, C.byref(s)) == gs_err()
    assert b"unknown precision" in lib.gs_last_error()


def gs_err():
    return -1


def test_solver_requires_setup():
    s = gs.Solver(gs.PREC_DOUBLE)
    with pytest.raises(gs.GpuSolverError, match="not set up"):
        s.solve(np.ones(4))


def test_unknown_option_rejected():
    s = gs.Solver(gs.PREC_DOUBLE)
    with pytest.raises(gs.GpuSolverError, match="unknown integer option"):
        s.set("nonsense", 1)
    with pytest.raises(gs.GpuSolverError, match="unsupported smoother type"):
        s.set("smoother", 9)


def test_dense_spd_inverse_matches_numpy_and_oracle():
    """The product's coarsest-level inverse (host setup) against numpy and
    the oracle's independent restatement."""
    import oracle_api as orc
    rng = np.random.default_rng(3)
    for n in (1, 3, 17, 64):
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        a = (q * rng.uniform(0.5, 20.0, n)) @ q.T
        inv = gs.dense_spd_inverse(a)
        ref = np.linalg.inv(a)
        assert np.max(np.abs(inv - ref)) <= 1e-12 * np.max(np.abs(ref))
        assert np.max(np.abs(inv - orc.dense_spd_inverse(a))) <= 1e-12 * np.max(np.abs(ref))
        assert np.array_equal(inv, inv.T)
    with pytest.raises(gs.GpuSolverError):
        gs.dense_spd_inverse(np.array([[0.0, 1.0], [1.0, 0.0]]))
