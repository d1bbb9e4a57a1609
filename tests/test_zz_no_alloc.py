"""qslhip_sample_all_noalloc: the sampler whose buffers live in a static
device array (no allocation), called through the C ABI, against
qslhip_sample_all on the same handle and uniforms: outcomes equal index for
index.  Two calls in a row, the second on a smaller state with fewer
uniforms (the static array keeps the first call's contents), in both
precisions; the largest accepted sizes; the refusals beyond them."""
import ctypes as C

import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2

pytestmark = pytest.mark.gpu
O = Oracle()


def _fn():
    f = L.lib().qslhip_sample_all_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    return f


def _noalloc(q, u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.full(len(u), 0xDEADBEEF, np.uint64)
    rc = _fn()(q._h, L.ptr(u), len(u), L.ptr(out))
    return rc, out


def _reference(q, u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(len(u), np.uint64)
    L.check(L.lib().qslhip_sample_all(q._h, L.ptr(u), len(u), L.ptr(out)))
    return out


def _circuit(rng, n, count, dtype):
    ops = np.zeros(count, OP_DTYPE)
    kinds = GATE1 + GATE2
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ang = float(rng.uniform(-2 * np.pi, 2 * np.pi))
        if dtype == np.complex64:
            ang = float(np.float32(ang))
        ops[i] = (k, a, b if k >= 16 else 0, 0, ang)
    return ops


def _state(n, dtype, seed):
    rng = np.random.default_rng(seed)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.apply_ops(_circuit(rng, n, 60, dtype), mode=qsl_amd.MODE_FUSED)
    return q


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_matches_sample_all_twice(dtype):
    # first call: 16 qubits, 20000 uniforms
    q = _state(16, dtype, 1)
    u = qsl_amd.uniforms(11, 20000)
    u[:3] = [0.0, 0.5, 1.0 - 2 ** -53]
    rc, got = _noalloc(q, u)
    assert rc == L.OK, L.lib().qslhip_last_error()
    assert np.array_equal(got, _reference(q, u))
    # second call: a smaller state and fewer uniforms; the array still holds
    # the first call's sums and outcomes beyond the new sizes
    q2 = _state(9, dtype, 2)
    u2 = qsl_amd.uniforms(12, 700)
    rc, got2 = _noalloc(q2, u2)
    assert rc == L.OK
    assert np.array_equal(got2, _reference(q2, u2))
    assert got2.max() < (1 << 9)
    # the existing entry point is unchanged by the calls above
    assert np.array_equal(_reference(q, u), got)


def test_noalloc_largest_sizes_and_oracle():
    n = 24  # the largest state it takes, with the most uniforms per call
    q = _state(n, np.complex128, 3)
    u = qsl_amd.uniforms(13, 1 << 18)
    rc, got = _noalloc(q, u)
    assert rc == L.OK, L.lib().qslhip_last_error()
    assert np.array_equal(got, _reference(q, u))
    # and on a small state, against the oracle itself (exact outside 1e-12
    # of a CDF value)
    q3 = _state(12, np.complex128, 4)
    st = q3.getState()
    u3 = qsl_amd.uniforms(14, 5000)
    rc, got3 = _noalloc(q3, u3)
    assert rc == L.OK
    want = O.sample_all(st, u3)
    _, cdf = O.gen_dist(st)
    from oracle.pyoracle import near_cdf_boundary
    assert np.all((got3 == want) | near_cdf_boundary(cdf, u3, 1e-12))


def test_noalloc_refusals():
    q = qsl_amd.Qubits(25)  # above 2^24 amplitudes
    rc, _ = _noalloc(q, qsl_amd.uniforms(1, 10))
    assert rc == L.EINVAL
    assert b"2^24" in L.lib().qslhip_last_error()
    q = qsl_amd.Qubits(10)
    rc, _ = _noalloc(q, np.zeros((1 << 18) + 1))
    assert rc == L.EINVAL
    assert b"2^18" in L.lib().qslhip_last_error()
    qn = qsl_amd.QubitsNP(10, nones=4)
    rc, _ = _noalloc(qn, qsl_amd.uniforms(1, 10))
    assert rc == L.EKIND
    # zero uniforms: nothing to do
    rc, out = _noalloc(q, np.zeros(0))
    assert rc == L.OK and out.size == 0
