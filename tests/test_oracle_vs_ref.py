"""Pin the oracle bit for bit against the reference library itself
(oracle/_ref/libqslref.so, compiled by `make -C oracle ref` from qsl's own
sources under /root/reference with the reference's -O2).

Gates: every kind, every target (pair) at n = 1..7, random states drawn
with the recipe of makeRandomState, many random angles (GCC's sincos
folding in the reference build changes a cosine by an ulp for some angles,
so many angles are tried).  prob / postselect on the Default, Omp, NP and
OmpNP backends.  Utilities: next, hammingWeight, generateLookup, and the
reference's 32-bit choose overflow.  Skipped when the reference library is
not built (it is not shipped to the GPU box)."""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as P

if not os.path.exists(P.REF_PATH):
    pytest.skip("oracle/_ref not built (needs /root/reference)", allow_module_level=True)

O = P.Oracle()


class _LazyRef:
    """The reference library, loaded by the first test that uses it (so a
    run that deselects these tests never maps it)."""

    _lib = None

    def __getattr__(self, name):
        if _LazyRef._lib is None:
            _LazyRef._lib = P.RefLib()
        return getattr(_LazyRef._lib, name)


REF = _LazyRef()


def _bits_equal(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_default_gates_bit_exact(dtype):
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 5, 7):
        for kind in P.GATE1:
            for t in range(n):
                for _ in range(3):
                    st = P.random_state(rng, n, dtype)
                    ang = float(rng.uniform(-8, 8))
                    if dtype == np.complex64:
                        ang = float(np.float32(ang))
                    a = st.copy()
                    O.gate1(a, kind, t, ang)
                    r = P.RefSim(REF, REF.DEFAULT, dtype, n)
                    r.set_state(st)
                    r.gate1(kind, t, ang)
                    assert _bits_equal(a, r.get_state()), (n, kind, t, ang)
        if n < 2:
            continue
        for kind in P.GATE2:
            for x in range(n):
                for y in range(n):
                    if x == y:
                        continue
                    st = P.random_state(rng, n, dtype)
                    ang = float(rng.uniform(-8, 8))
                    if dtype == np.complex64:
                        ang = float(np.float32(ang))
                    a = st.copy()
                    O.gate2(a, kind, x, y, ang)
                    r = P.RefSim(REF, REF.DEFAULT, dtype, n)
                    r.set_state(st)
                    r.gate2(kind, x, y, ang)
                    assert _bits_equal(a, r.get_state()), (n, kind, x, y, ang)


def test_sincos_angles_sweep():
    # 2000 angles through rotateY / phase at n = 3: the coefficient path
    rng = np.random.default_rng(9)
    st = P.random_state(rng, 3)
    for ang in rng.uniform(-20, 20, 2000):
        for kind in (P.ROT_Y, P.PHASE):
            a = st.copy()
            O.gate1(a, kind, 1, float(ang))
            r = P.RefSim(REF, REF.DEFAULT, np.complex128, 3)
            r.set_state(st)
            r.gate1(kind, 1, float(ang))
            assert _bits_equal(a, r.get_state()), (kind, ang)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_omp_backend_matches(dtype):
    rng = np.random.default_rng(2)
    n = 9
    st = P.random_state(rng, n, dtype)
    for kind in P.GATE1:
        for t in (0, 4, n - 2, n - 1):
            a = st.copy()
            O.omp_gate1(a, kind, t, 0.77, 3)
            r = P.RefSim(REF, REF.OMP, dtype, n, 3)
            r.set_state(st)
            r.gate1(kind, t, 0.77)
            assert _bits_equal(a, r.get_state())


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_prob_postselect(dtype):
    rng = np.random.default_rng(3)
    n = 6
    st = P.random_state(rng, n, dtype)
    r = P.RefSim(REF, REF.DEFAULT, dtype, n)
    r.set_state(st)
    for t in range(n):
        for o in (0, 1):
            assert O.prob(st, t, o) == r.prob(t, o)
    for t in (0, 3, 5):
        for o in (0, 1):
            a = st.copy()
            pa = O.postselect(a, t, o)
            rr = P.RefSim(REF, REF.DEFAULT, dtype, n)
            rr.set_state(st)
            assert pa == rr.postselect(t, o)
            assert _bits_equal(a, rr.get_state())


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_np_backend_bit_exact(dtype):
    rng = np.random.default_rng(4)
    for n, k in ((4, 2), (6, 3), (7, 2), (8, 5)):
        for kind in P.NP_GATE1:
            for t in range(n):
                st = P.random_np_state(rng, n, k, dtype)
                ang = float(rng.uniform(-8, 8))
                if dtype == np.complex64:
                    ang = float(np.float32(ang))
                a = st.copy()
                O.np_gate1(a, k, kind, t, ang)
                r = P.RefSim(REF, REF.NP, dtype, n, k)
                r.set_state(st)
                r.gate1(kind, t, ang)
                assert _bits_equal(a, r.get_state()), (n, k, kind, t)
        for kind in P.NP_GATE2:
            for x in range(n):
                for y in range(n):
                    if x == y:
                        continue
                    st = P.random_np_state(rng, n, k, dtype)
                    ang = float(rng.uniform(-8, 8))
                    if dtype == np.complex64:
                        ang = float(np.float32(ang))
                    a = st.copy()
                    O.np_gate2(a, k, kind, x, y, ang)
                    r = P.RefSim(REF, REF.NP, dtype, n, k)
                    r.set_state(st)
                    r.gate2(kind, x, y, ang)
                    assert _bits_equal(a, r.get_state()), (n, k, kind, x, y)
        st = P.random_np_state(rng, n, k, dtype)
        r = P.RefSim(REF, REF.OMPNP, dtype, n, k)
        r.set_state(st)
        for t in range(n):
            for o in (0, 1):
                assert O.np_prob(st, k, t, o) == pytest.approx(r.prob(t, o), rel=1e-5)


def test_np_compact_oracle_matches_reference_np():
    # the compact colex layout applied through orc_npc_* equals the
    # reference NP backend on the full vector
    rng = np.random.default_rng(5)
    n, k = 9, 4
    st = P.random_np_state(rng, n, k)
    r = P.RefSim(REF, REF.NP, np.complex128, n, k)
    r.set_state(st)
    amp = O.npc_from_full(st, n, k)
    ops = []
    for i in range(60):
        kind = int(rng.choice(P.NP_GATE1 + P.NP_GATE2))
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        ang = float(rng.uniform(0, 6.3))
        ops.append((kind, a, b if kind >= 16 else 0, 0, ang))
        if kind < 16:
            r.gate1(kind, a, ang)
        else:
            r.gate2(kind, a, b, ang)
    O.npc_apply_ops(amp, n, k, np.array(ops, P.OP_DTYPE), nthreads=2)
    assert _bits_equal(O.npc_to_full(amp, n, k), r.get_state())


def test_utilities_match_reference():
    L = REF.lib
    x = 0b111
    for _ in range(200):
        assert L.ref_next(x) == O.next(x)
        x = O.next(x)
    for v in (0, 1, 0b1010, 0b11011010, 2 ** 40 - 1):
        assert L.ref_hamming_weight(v) == O.hamming_weight(v)
    for n in range(1, 12):
        for k in range(1, n + 1):
            out = np.zeros(math.comb(n, k), np.uint64)
            m = L.ref_generate_lookup(n, k, out.ctypes.data, len(out))
            assert m == len(out)
            assert np.array_equal(out, O.generate_lookup(n, k))
    # the reference's 32-bit choose overflows at C(32,16); the oracle's does not
    assert L.ref_choose(10, 3) == 120
    assert L.ref_choose(32, 16) != math.comb(32, 16)
    assert O.choose(32, 16) == math.comb(32, 16)
