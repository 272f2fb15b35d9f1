"""The C++ host surface (include/pt_tracer.hpp) on the GPU: the compiled
program tests/cpp/facade_main drives it like the reference application
drives CUDAPathTracer and checks its passes against the C-ABI bit for bit
(see the program's header for the list of checks)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_main")


def test_cpp_facade():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK")
    assert "setScreenBuffer failed" in r.stderr  # logged like CUDA_ERROR_CHECK, not thrown
