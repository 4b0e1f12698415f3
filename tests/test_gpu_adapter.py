"""The QuantumGates adapter (qpanda_amd/cpp/HipQuantumGates.cpp) under the
reference back end's own known-answer tests, re-expressed in
tests/cpp/test_hip_quantum_gates.cpp and run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qpanda_amd", "_build", "test_hip_quantum_gates")


def test_quantum_gates_adapter_known_answers():
    assert os.path.exists(EXE), "adapter test driver not built (qpanda_amd/build.sh)"
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:], flush=True)
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-2000:]
    assert "0 failures" in p.stdout
