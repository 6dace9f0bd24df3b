"""HipBackend through the ExecutionBackend interface (the counterpart of the
reference's tests/test_backend.cpp + tests/test_cuda_consistency.cpp), run as
the C++ binary tests/cpp/hip_backend_test built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "hip_backend_test")


@pytest.mark.gpu
def test_hip_backend_dropin_binary():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    print("running", BIN, flush=True)
    proc = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(proc.stdout[-4000:], proc.stderr[-2000:])
    assert proc.returncode == 0
    assert "ALL HIPBACKEND TESTS PASSED" in proc.stdout
