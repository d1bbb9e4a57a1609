"""Pin the oracle's measurement decision rules, independently of the GPU.

The reference draws its uniforms from an unseedable generator
(src/utils/random.cpp:32-45), so its measure / sample / sampleAll2 cannot
be fed a chosen uniform.  Two kinds of pins are used instead.

1. Known answers that the mathematics fixes.  States are chosen so that the
   probabilities are exact binary fractions (amplitudes 1/2 and (1+i)/2) or
   far from 1/2, and uniforms sit on, just below and just above a CDF value:
   - measure(t):  u < prob0 -> 0 with factor 1/sqrt(prob0), else 1 with
     1/sqrt(1 - prob0)               (src/qubits/default/measure.cpp:94-101)
   - sample(t, N): counts of u < prob0, the same rule      (measure.cpp:204-212)
   - drawSample: the first m with cdf[m] < u <= cdf[m+1], over the nonzero
     probabilities only; u > total -> 0           (measure.cpp:60-78,164-188)
   - sampleAll2: sorted uniforms merged against the running sum with the
     strict rule u < sum                                   (measure.cpp:237-267)
   A reversed or non-strict rule fails at least one case (checked below by
   planting such a rule in a copy of the oracle's decision, on the CPU).

2. The reference library itself (oracle/_ref/libqslref.so, built from qsl's
   sources), where its own generator draws: outcome frequencies on a qubit
   with prob0 = 0.09 (a reversed rule gives 0.91), and, for every outcome it
   draws, its collapsed state bit for bit against the oracle's collapse for
   that outcome (pins the factor 1/sqrt(prob0) resp. 1/sqrt(1 - prob0) in Fp).
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as P

O = P.Oracle()
HAVE_REF = os.path.exists(P.REF_PATH)
DTYPES = [np.complex128, np.complex64]


def _rt(dtype):
    return np.float64 if np.dtype(dtype) == np.complex128 else np.float32


def _below(x, dtype):
    rt = _rt(dtype)
    return rt(np.nextafter(rt(x), rt(0)))


def _above(x, dtype):
    rt = _rt(dtype)
    return rt(np.nextafter(rt(x), rt(2)))


def dyadic_half(dtype):
    """3 qubits, amplitude 1/2 on |000>, |010>, |001>, |011>: prob(0, 0) =
    prob(0, 1) = 1/2 exactly, prob(1, .) = 1/2, prob(2, 0) = 1."""
    st = np.zeros(8, dtype)
    st[[0b000, 0b010, 0b001, 0b011]] = 0.5
    return st


def skewed(dtype):
    """2 qubits: 0.3|00> + sqrt(1 - 0.09)|01>: prob(0, 0) = 0.09, far from
    1/2, so a reversed rule gives the other outcome for most uniforms."""
    st = np.zeros(4, dtype)
    st[0b00] = 0.3
    st[0b01] = math.sqrt(1 - 0.09)
    return st


def cdf_state(dtype, scale=1.0):
    """p = [1/4, 0, 1/2, 1/4] exactly: amplitudes 1/2, 0, (1+i)/2, 1/2
    (|(1+i)/2|^2 = 1/4 + 1/4).  scale < 1 leaves total mass scale^2."""
    st = np.array([0.5, 0, 0.5 + 0.5j, 0.5], dtype) * _rt(dtype)(scale)
    return st


# ---- 1. known answers -------------------------------------------------------


@pytest.mark.parametrize("dtype", DTYPES)
def test_measure_rule_on_exact_half(dtype):
    st = dyadic_half(dtype)
    assert O.prob(st, 0, 0) == 0.5
    # u == prob0: not below it, so outcome 1 (strict <)
    a = st.copy()
    assert O.measure(a, 0, _rt(dtype)(0.5)) == 1
    # the factor is 1/sqrt(1 - prob0) = sqrt(2): kept amplitudes 1/sqrt(2)
    want = np.zeros(8, dtype)
    f = _rt(dtype)(1) / np.sqrt(_rt(dtype)(1) - _rt(dtype)(0.5))
    want[[0b001, 0b011]] = _rt(dtype)(0.5) * f
    assert np.array_equal(a, want)
    # just below prob0: outcome 0
    b = st.copy()
    assert O.measure(b, 0, _below(0.5, dtype)) == 0
    want = np.zeros(8, dtype)
    want[[0b000, 0b010]] = _rt(dtype)(0.5) * (_rt(dtype)(1) / np.sqrt(_rt(dtype)(0.5)))
    assert np.array_equal(b, want)
    # prob0 == 1 on qubit 2: every u in [0, 1) gives 0 and the state is unchanged
    for u in (0.0, 0.5, _below(1.0, dtype)):
        c = st.copy()
        assert O.measure(c, 2, _rt(dtype)(u)) == 0
        assert np.array_equal(c, st)


@pytest.mark.parametrize("dtype", DTYPES)
def test_measure_rule_far_from_half(dtype):
    st = skewed(dtype)
    p0 = O.prob(st, 0, 0)
    assert abs(p0 - 0.09) < 1e-6
    cases = [(0.0, 0), (0.05, 0), (0.0899, 0), (0.0901, 1), (0.5, 1), (0.95, 1)]
    for u, want in cases:
        a = st.copy()
        got = O.measure(a, 0, _rt(dtype)(u))
        assert got == want, (u, got)
        # the collapsed state is the normalised projection
        proj = st.copy()
        proj[[i for i in range(4) if ((i >> 0) & 1) != want]] = 0
        proj = proj / np.sqrt(np.sum(np.abs(proj.astype(np.complex128)) ** 2))
        tol = 1e-15 if dtype == np.complex128 else 1e-6
        assert np.max(np.abs(a - proj)) < tol
        # the other qubit keeps prob 1 on |0>
        assert abs(O.prob(a, 1, 0) - 1) < tol


@pytest.mark.parametrize("dtype", DTYPES)
def test_sample_counts(dtype):
    st = dyadic_half(dtype)
    u = [0.0, 0.25, _below(0.5, dtype), 0.5, _above(0.5, dtype), 0.75]
    # u < 1/2: 0.0, 0.25 and the one just below -> 3 zeros, 3 ones
    assert list(O.sample(st, 0, u)) == [3, 3]
    # prob0 == 1: all zeros
    assert list(O.sample(st, 2, u)) == [6, 0]
    sk = skewed(dtype)
    assert list(O.sample(sk, 0, [0.05, 0.5, 0.95, 0.0899])) == [2, 2]


@pytest.mark.parametrize("dtype", DTYPES)
def test_draw_sample_rule(dtype):
    st = cdf_state(dtype)
    idx, cdf = O.gen_dist(st)
    # generateDist: {0, 0} then the nonzero entries only, running FP64 sum
    assert list(idx) == [0, 0, 2, 3]
    assert list(cdf) == [0.0, 0.25, 0.75, 1.0]
    cases = [
        (0.1, 0),
        (0.25, 0),                   # cdf[0] < u <= cdf[1]: the boundary belongs below
        (_above(0.25, dtype), 2),    # index 1 has probability 0 and is skipped
        (0.5, 2),
        (0.75, 2),
        (_above(0.75, dtype), 3),
        (_below(1.0, dtype), 3),
    ]
    u = np.array([c[0] for c in cases], _rt(dtype))
    assert list(O.sample_all(st, u)) == [c[1] for c in cases]
    for uu, want in cases:
        a = st.copy()
        assert O.measure_all(a, _rt(dtype)(uu)) == want
    # u == 0 reads dist[-1] in the reference (out of bounds); the oracle
    # returns the first outcome of nonzero probability
    assert list(O.sample_all(st, [0.0])) == [0]
    # u beyond the total mass -> 0 (drawSample's initial value)
    low = cdf_state(dtype, scale=0.5)  # total 1/4: cdf [1/16, 1/16, 3/16, 1/4]
    assert list(O.sample_all(low, [0.1, 0.2, 0.3, 0.9])) == [2, 3, 0, 0]


@pytest.mark.parametrize("dtype", DTYPES)
def test_sample_all2_strict_merge(dtype):
    st = cdf_state(dtype)
    # running sums 1/4, 1/4, 3/4, 1: u = 1/4 is not < 1/4, so it falls to
    # index 2 together with the one just above; 3/4 falls to index 3
    u = [0.25, _above(0.25, dtype), 0.75]
    assert list(O.sample_all2(st, u)) == [0, 0, 2, 1]
    assert list(O.sample_all2(st, [0.0, _below(0.25, dtype)])) == [2, 0, 0, 0]
    # unsorted input is sorted first
    assert list(O.sample_all2(st, [0.75, 0.1, 0.5])) == [1, 0, 1, 1]
    # uniforms at or beyond the total mass are not counted
    low = cdf_state(dtype, scale=0.5)  # running sums 1/16, 1/16, 3/16, 1/4
    assert list(O.sample_all2(low, [0.05, 0.2, 0.25, 0.9])) == [1, 0, 0, 1]


def _planted_measure(p0, u, reverse=False, nonstrict=False):
    """A copy of the decision with a planted mistake (CPU only), used to show
    that the cases above tell the rules apart."""
    if reverse:
        return 0 if u >= p0 else 1
    if nonstrict:
        return 0 if u <= p0 else 1
    return 0 if u < p0 else 1


def test_planted_mistakes_are_caught():
    # the exact-half boundary case separates strict from non-strict
    assert _planted_measure(0.5, 0.5) == 1
    assert _planted_measure(0.5, 0.5, nonstrict=True) != 1
    # the skewed cases separate the rule from its reverse
    for u, want in [(0.05, 0), (0.5, 1), (0.95, 1)]:
        assert _planted_measure(0.09, u) == want
    assert any(_planted_measure(0.09, u, reverse=True) != w
               for u, w in [(0.05, 0), (0.5, 1), (0.95, 1)])
    # sampleAll2 with <= instead of < puts u = 1/4 at index 0
    sums = [0.25, 0.25, 0.75, 1.0]

    def merge(us, strict):
        us, m, out = sorted(us), 0, []
        for s in sums:
            c = 0
            while m < len(us) and ((us[m] < s) if strict else (us[m] <= s)):
                c, m = c + 1, m + 1
            out.append(c)
        return out
    us = [0.25, float(np.nextafter(0.25, 1)), 0.75]
    assert merge(us, True) == [0, 0, 2, 1]
    assert merge(us, False) != [0, 0, 2, 1]


# ---- 2. the reference library ------------------------------------------------

ref_only = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs /root/reference)")


@pytest.fixture(scope="module")
def ref():
    return P.RefLib()


@ref_only
@pytest.mark.parametrize("dtype", DTYPES)
def test_reference_measure_frequency_and_collapse(ref, dtype):
    st = skewed(dtype)
    after = {}
    for o in (0, 1):
        a = st.copy()
        assert O.measure(a, 0, _rt(dtype)(0.0 if o == 0 else 0.999)) == o
        after[o] = a
    n, zeros = 3000, 0
    for _ in range(n):
        r = P.RefSim(ref, ref.DEFAULT, dtype, 2)
        r.set_state(st)
        o = r.measure(0)
        zeros += o == 0
        # the reference's collapse for the outcome it drew is the oracle's,
        # bit for bit
        assert np.array_equal(r.get_state().view(np.uint8), after[o].view(np.uint8))
    # binomial(3000, 0.09): sd 15.7; a reversed rule gives about 2730
    assert 150 < zeros < 390, zeros


@ref_only
@pytest.mark.parametrize("dtype", DTYPES)
def test_reference_sample_frequency(ref, dtype):
    r = P.RefSim(ref, ref.DEFAULT, dtype, 2)
    r.set_state(skewed(dtype))
    c = r.sample(0, 40000)
    assert int(c.sum()) == 40000
    assert abs(c[0] / 40000 - 0.09) < 0.008  # sd 0.0014


@ref_only
@pytest.mark.parametrize("dtype", DTYPES)
def test_reference_sample_all_and_all2(ref, dtype):
    n = 40000
    r = P.RefSim(ref, ref.DEFAULT, dtype, 2)
    r.set_state(cdf_state(dtype))
    for counts in (r.sample_all(n), r.sample_all2(n)):
        assert counts[1] == 0  # probability 0 is never drawn
        assert int(counts.sum()) == n
        assert np.allclose(counts / n, [0.25, 0, 0.5, 0.25], atol=0.012)
    # total mass 1/4: sampleAll2 counts only the uniforms below it
    r.set_state(cdf_state(dtype, scale=0.5))
    c2 = r.sample_all2(n)
    assert abs(c2.sum() / n - 0.25) < 0.012
    assert np.allclose(c2 / n, [1 / 16, 0, 1 / 8, 1 / 16], atol=0.008)


@ref_only
def test_reference_measure_all_collapses_to_basis(ref):
    for _ in range(200):
        r = P.RefSim(ref, ref.DEFAULT, np.complex128, 2)
        r.set_state(cdf_state(np.complex128))
        x = r.measure_all()
        assert x in (0, 2, 3)
        want = np.zeros(4, np.complex128)
        want[x] = 1
        assert np.array_equal(r.get_state(), want)
