"""gpuSolver::HipBackend on the GPU, driven through the reference's
interface only (gpuSolver::Backend &): the check program built by build()
from gpu-conjugate-gradient-solver_amd/adapter/check.cc compares every
method with the C function it forwards to (bit for bit), the element-wise,
SpMV, ELL and Gauss-Seidel results with FP32 host sums in the same order,
checks the interface's error behaviour, and runs a conjugate-gradient loop
in the order of src/CGSolver.cc:87-306.
"""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = (Path(__file__).resolve().parents[1] / "gpu-conjugate-gradient-solver_amd" / "adapter" /
       "_build" / "adapter_check")


@pytest.mark.skipif(not BIN.exists(),
                    reason="built only where the reference's interface headers are present")
def test_backend_adapter_through_reference_interface():
    print(f"running {BIN}", flush=True)
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "checks passed" in p.stdout
