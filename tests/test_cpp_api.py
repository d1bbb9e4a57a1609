"""The C++20 drop-in API (include/qsl/qubits/hip*.hpp) through its own
typed-test program, tests/cpp/test_qubits_hip (built by
__graft_entry__.build()).  Its static_asserts check the reference's
concepts (include/qsl/concepts.hpp) at compile time."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_qubits_hip")


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built: run __graft_entry__.build()")


def test_cpp_api_compiles_and_satisfies_concepts():
    _need_exe()
    r = subprocess.run([EXE, "--compile"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "concept checks passed" in r.stdout


@pytest.mark.gpu
def test_cpp_api_typed_tests():
    _need_exe()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout
