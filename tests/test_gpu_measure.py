"""GPU parity of prob / collapse / measure / postselect / sampling against the
oracle (src/qubits/default/measure.cpp semantics with injected uniforms).

prob is accumulated in FP64 on the device in a different order from the
reference's serial Fp sum, so it is compared within a tolerance: 1e-12 for
FP64 states, 1e-5 for FP32 states (whose reference accumulator is float).
Sampled outcomes must be identical except for uniforms within 1e-12 (FP64)
of a CDF boundary of the oracle, where summation order may flip the draw.
"""
import numpy as np
import pytest

import qsl_amd
from oracle.pyoracle import Oracle, random_state, near_cdf_boundary

pytestmark = pytest.mark.gpu
O = Oracle()
PTOL = {np.complex128: 1e-12, np.complex64: 1e-5}


_near_boundary = near_cdf_boundary


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_prob_all_targets(dtype):
    n = 13
    rng = np.random.default_rng(1)
    st = random_state(rng, n, dtype)
    q = qsl_amd.Qubits(st)
    for t in range(n):
        for o in (0, 1):
            assert abs(q.prob(t, o) - O.prob(st, t, o)) <= PTOL[dtype]
    assert abs(q.norm() - 1) <= PTOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_postselect_and_collapse(dtype):
    n = 10
    rng = np.random.default_rng(2)
    st = random_state(rng, n, dtype)
    for t in (0, 4, n - 1):
        for o in (0, 1):
            q = qsl_amd.Qubits(st)
            p = q.postselect(t, o)
            want = st.copy()
            pw = O.postselect(want, t, o)
            assert abs(p - pw) <= PTOL[dtype]
            assert np.abs(q.getState() - want).max() <= 10 * PTOL[dtype]
            assert abs(q.prob(t, o) - 1) <= 10 * PTOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_measure_with_injected_uniform(dtype):
    n = 9
    rng = np.random.default_rng(3)
    st = random_state(rng, n, dtype)
    for t in range(n):
        p0 = O.prob(st, t, 0)
        for u in (0.0, p0 * 0.5, min(0.999, p0 + 0.01), 0.999999):
            if abs(u - p0) < 1e-6:
                continue
            q = qsl_amd.Qubits(st)
            got = q.measure(t, u)
            want = st.copy()
            w = O.measure(want, t, u)
            assert got == w
            assert np.abs(q.getState() - want).max() <= 10 * PTOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 6, 14, 20])
def test_sample_all_bit_exact(dtype, n):
    rng = np.random.default_rng(4 + n)
    st = random_state(rng, n, dtype)
    # a few exact zeros, as generateDist skips them
    st[rng.choice(st.size, st.size // 4, replace=False)] = 0
    st /= np.sqrt(np.sum(np.abs(st.astype(np.complex128)) ** 2))
    st = st.astype(dtype)
    q = qsl_amd.Qubits(st)
    u = qsl_amd.uniforms(99 + n, 20000)
    u[:4] = [0.0, 1e-300, 0.5, 1.0 - 2 ** -53]
    got = q.sampleAllOutcomes(u)
    want = O.sample_all(st, u)
    idx, cdf = O.gen_dist(st)
    uf = u if dtype == np.complex128 else u.astype(np.float32).astype(np.float64)
    eps = 1e-12 if dtype == np.complex128 else 1e-5
    ok = (got == want) | _near_boundary(cdf, uf, eps)
    assert ok.all(), np.nonzero(~ok)[0][:10]
    # outcomes never land on zero-probability entries
    assert np.all(np.abs(st[got.astype(np.int64)]) > 0)


def test_sample_above_total_returns_zero():
    q = qsl_amd.Qubits(3)
    st = np.zeros(8, np.complex128)
    st[5] = np.sqrt(0.5)
    st[6] = np.sqrt(0.5) * 0.999  # total mass < 1
    q.setState(st)
    out = q.sampleAllOutcomes(np.array([0.25, 0.6, 0.9999]))
    assert list(out) == [5, 6, 0]


def test_dyadic_distribution_exact():
    # probabilities 1/4, 0, 1/2, 1/4: every CDF value is exact
    st = np.array([0.5, 0, np.sqrt(0.5), 0.5], np.complex128)
    q = qsl_amd.Qubits(st)
    u = np.array([1e-9, 0.25, 0.25 + 1e-12, 0.5, 0.75, 0.75 + 1e-12, 1.0])
    assert list(q.sampleAllOutcomes(u)) == [0, 0, 2, 2, 2, 3, 3]
    assert list(O.sample_all(st, u)) == [0, 0, 2, 2, 2, 3, 3]


def test_measure_all_collapses():
    st = np.array([0.5, 0, np.sqrt(0.5), 0.5], np.complex128)
    q = qsl_amd.Qubits(st)
    assert q.measureAll(0.6) == 2
    s = q.getState()
    assert s[2] == 1 and np.count_nonzero(s) == 1


def test_sample_counts():
    n = 8
    rng = np.random.default_rng(5)
    st = random_state(rng, n)
    q = qsl_amd.Qubits(st)
    u = qsl_amd.uniforms(5, 5000)
    for t in (0, 7):
        assert q.sample(t, len(u), u) == list(O.sample(st, t, u))
