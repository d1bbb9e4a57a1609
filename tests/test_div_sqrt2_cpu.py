"""The Hadamard's division by sqrt(2) runs on the device as q0 = x*r,
e = fma(-q0, d, x), q = fma(e, r, q0) with r = RN(1/d) (gate_math.h).
The reference divides (src/qubits/default/gates.cpp:94-100), so the two must
agree bit for bit: the host build of the same header is checked against
x / d for every FP32 value and 2^26 sampled FP64 values, including the
magnitudes where the guard falls back to the division itself."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_div_sqrt2_matches_division():
    exe = os.path.join(CPP, "div_sqrt2_check")
    r = subprocess.run(["make", "-s", "div_sqrt2_check"], cwd=CPP, capture_output=True,
                       text=True)
    if r.returncode != 0:
        pytest.fail("build failed: " + r.stderr[-2000:])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
