"""Pin the oracle's Resize restatement (addQubit, collapseOut via
postselectOut, measureOut) bit for bit against the reference library's own
Resize backend (src/qubits/resize/, built into oracle/_ref/libqslref.so),
and against what the mathematics fixes: addQubit(t) then postselectOut(t, 0)
is the identity, and removing a qubit keeps the order of the others."""
import os

import numpy as np
import pytest

from oracle import pyoracle as P

O = P.Oracle()
HAVE_REF = os.path.exists(P.REF_PATH)
DTYPES = [np.complex128, np.complex64]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("dtype", DTYPES)
def test_add_qubit_closed_form(dtype):
    rng = np.random.default_rng(3)
    for n in range(1, 7):
        st = P.random_state(rng, n, dtype)
        for t in range(n + 1):
            got = O.add_qubit(st, t)
            want = np.zeros(2 * st.size, dtype)
            for i in range(st.size):
                lo, hi = i & ((1 << t) - 1), i >> t
                want[(hi << (t + 1)) | lo] = st[i]
            assert np.array_equal(_bits(got), _bits(want)), (n, t)
            # removing the new qubit (outcome 0, factor 1) is the identity
            back = O.collapse_out(got, t, 0, 1.0)
            assert np.array_equal(_bits(back), _bits(st))
            # postselecting it keeps the state up to the norm's rounding
            p, back = O.postselect_out(got, t, 0)
            tol = 1e-15 if dtype == np.complex128 else 1e-6
            assert abs(p - 1) < 4 * tol and np.abs(back - st).max() < 4 * tol


@pytest.mark.parametrize("dtype", DTYPES)
def test_collapse_out_closed_form(dtype):
    rng = np.random.default_rng(4)
    rt = np.float64 if dtype == np.complex128 else np.float32
    for n in range(1, 7):
        st = P.random_state(rng, n, dtype)
        for t in range(n):
            for o in (0, 1):
                f = rt(1.5)
                got = O.collapse_out(st, t, o, f)
                keep = [i for i in range(st.size) if ((i >> t) & 1) == o]
                want = np.array([st[i] for i in keep], dtype) * f
                assert np.array_equal(_bits(got), _bits(want.astype(dtype))), (n, t, o)


ref_only = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs /root/reference)")


@ref_only
@pytest.mark.parametrize("dtype", DTYPES)
def test_resize_vs_reference(dtype):
    ref = P.RefLib()
    rng = np.random.default_rng(5)
    for n in range(1, 7):
        st = P.random_state(rng, n, dtype)
        for t in range(n + 1):
            r = P.RefSim(ref, ref.RESIZE, dtype, n)
            r.set_state_any(st)
            r.add_qubit(t)
            assert r.nqubits() == n + 1
            assert np.array_equal(_bits(r.get_state_any()), _bits(O.add_qubit(st, t)))
        for t in range(n if n > 1 else 0):
            for o in (0, 1):
                r = P.RefSim(ref, ref.RESIZE, dtype, n)
                r.set_state_any(st)
                p = r.postselect_out(t, o)
                want_p, want = O.postselect_out(st, t, o)
                assert p == float(want_p)
                assert r.nqubits() == n - 1
                assert np.array_equal(_bits(r.get_state_any()), _bits(want))


@ref_only
@pytest.mark.parametrize("dtype", DTYPES)
def test_resize_measure_out_vs_reference(dtype):
    ref = P.RefLib()
    # qubit 0 has prob(0) = 0.09: outcome frequencies pin the rule's
    # direction, and each outcome's state pins the factor
    st = np.zeros(8, dtype)
    st[0b000] = 0.3
    st[0b101] = np.sqrt(1 - 0.09)
    after = {o: O.measure_out(st, 0, 0.0 if o == 0 else 0.999)[1] for o in (0, 1)}
    zeros = 0
    for _ in range(2000):
        r = P.RefSim(ref, ref.RESIZE, dtype, 3)
        r.set_state_any(st)
        o = r.measure_out(0)
        zeros += o == 0
        assert np.array_equal(_bits(r.get_state_any()), _bits(after[o]))
    assert 90 < zeros < 280, zeros  # binomial(2000, 0.09): 180 +- 13
    # shrinking and growing again within the allocation, then trimming it
    r = P.RefSim(ref, ref.RESIZE, dtype, 5)
    big = P.random_state(np.random.default_rng(8), 5, dtype)
    r.set_state_any(big)
    p, small = O.postselect_out(big, 3, 1)
    assert r.postselect_out(3, 1) == float(p)
    r.add_qubit(1)
    r.reallocate()
    assert r.nqubits() == 5
    assert np.array_equal(_bits(r.get_state_any()), _bits(O.add_qubit(small, 1)))
    # (Resize::setState with a larger vector keeps the old capacity,
    # resize.cpp:70-81, so a following addQubit writes past the vector,
    # resize.cpp:110-117: not exercised here)
