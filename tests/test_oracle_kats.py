"""Pin the oracle's integer combinatorics against the reference's own
known-answer tests and against closed forms.

KATs restated from the reference's test suite:
  next() golden sequence      src/gtest/utils.cpp:194-215
  hammingWeight               src/gtest/utils.cpp:218-224
  choose                      src/gtest/utils.cpp:60-66
  NP reset index / print      src/gtest/sim-constructors.cpp:272-293
Closed forms: the colex rank is sum_i C(p_i, i) (math.comb), and
generateLookup(n, k) lists the weight-k strings in increasing order
(brute force over range(2^n)).
"""
import itertools
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle

O = Oracle()


def test_next_golden_sequence():
    x = 0b111
    seq = []
    for _ in range(9):
        x = O.next(x)
        seq.append(x)
    assert seq == [0b1011, 0b1101, 0b1110, 0b10011, 0b10101, 0b10110,
                   0b11001, 0b11010, 0b11100]


def test_hamming_weight_kats():
    assert O.hamming_weight(0b11011010) == 5
    assert O.hamming_weight(0b1010) == 2
    assert O.hamming_weight(0b1) == 1
    assert O.hamming_weight(0b0) == 0
    assert O.hamming_weight((1 << 64) - 1) == 64


def test_choose_kats_and_u64():
    assert O.choose(1, 1) == 1
    assert O.choose(4, 3) == 4
    assert O.choose(10, 3) == 120
    assert O.choose(10, 10) == 1
    # the reference's 32-bit choose overflows here (returns 28,418,082)
    assert O.choose(32, 16) == 601080390 == math.comb(32, 16)
    for n in range(0, 63):
        for k in range(0, n + 1):
            assert O.choose(n, k) == math.comb(n, k)
    # the table rank/unrank read holds the same values (k > n gives 0;
    # n <= 62, the largest NP register, where the u64 product is exact)
    for n in range(0, 63):
        for k in range(0, 65):
            assert O.lib.orc_choose_table(n, k) == math.comb(n, k)


@pytest.mark.parametrize("n", range(1, 13))
def test_generate_lookup_is_increasing_fixed_weight(n):
    for k in range(1, n + 1):
        got = [int(v) for v in O.generate_lookup(n, k)]
        want = [x for x in range(1 << n) if bin(x).count("1") == k]
        assert got == want


def _rank_closed_form(x):
    r, i = 0, 0
    for p in range(64):
        if (x >> p) & 1:
            i += 1
            r += math.comb(p, i)
    return r


@pytest.mark.parametrize("n", [4, 8, 12, 14])
def test_rank_unrank_exhaustive(n):
    for k in range(1, n):
        lst = O.generate_lookup(n, k)
        for r, x in enumerate(int(v) for v in lst):
            assert O.rank(x) == r
            assert O.unrank(r, k) == x
        assert len(lst) == math.comb(n, k)


def test_rank_unrank_n32_k16_sampled():
    rng = np.random.default_rng(32)
    total = math.comb(32, 16)
    assert O.unrank(0, 16) == (1 << 16) - 1
    assert O.unrank(total - 1, 16) == ((1 << 16) - 1) << 16
    for _ in range(3000):
        bits = rng.choice(32, 16, replace=False)
        x = int(sum(1 << int(b) for b in bits))
        r = O.rank(x)
        assert r == _rank_closed_form(x)
        assert 0 <= r < total
        assert O.unrank(r, 16) == x
    # consecutive ranks are consecutive weight-16 strings (Gosper order)
    r0 = int(rng.integers(0, total - 200))
    x = O.unrank(r0, 16)
    for r in range(r0, r0 + 200):
        assert O.unrank(r, 16) == x
        x = O.next(x)


def test_adjacent_bond_rank_offset():
    # moving the j-th one from q to q+1 adds C(q, j-1) to the rank
    rng = np.random.default_rng(7)
    for _ in range(2000):
        bits = rng.choice(20, 9, replace=False)
        x = int(sum(1 << int(b) for b in bits))
        for q in range(19):
            if (x >> q) & 1 and not (x >> (q + 1)) & 1:
                j = bin(x & ((1 << q) - 1)).count("1") + 1
                y = x ^ (1 << q) ^ (1 << (q + 1))
                assert O.rank(y) == O.rank(x) + math.comb(q, j - 1)


def test_np_reset_index():
    # NP reset puts the amplitude at 2^nones - 1 (sim-constructors.cpp:272-293:
    # n=3, nones=2 prints (1,0) on line 4)
    st = np.zeros(8, np.complex128)
    O.np_reset(st, 2)
    assert np.nonzero(st)[0].tolist() == [3]
    assert O.rank(3) == 0


def test_check_state_np_kats():
    # src/gtest/utils.cpp:298-331
    s0 = np.array([0, 1j, 1, 0], np.complex128)
    O.normalise(s0)
    assert O.check_state_np(s0) == 1
    s1 = np.array([1, 0, 0, 0], np.complex128)
    assert O.check_state_np(s1) == 0
    s2 = np.array([0, 0, 0, 1], np.complex128)
    assert O.check_state_np(s2) == 2
    s3 = np.array([1, 1j, 1, 0], np.complex128)
    O.normalise(s3)
    with pytest.raises(ValueError):
        O.check_state_np(s3)


@pytest.mark.parametrize("n,k", [(6, 3), (8, 2), (9, 5)])
def test_compact_roundtrip(n, k):
    rng = np.random.default_rng(n * 10 + k)
    full = np.zeros(1 << n, np.complex128)
    lst = O.generate_lookup(n, k).astype(np.int64)
    full[lst] = rng.normal(size=len(lst)) + 1j * rng.normal(size=len(lst))
    amp = O.npc_from_full(full, n, k)
    assert amp.size == math.comb(n, k)
    np.testing.assert_array_equal(amp, full[lst])
    np.testing.assert_array_equal(O.npc_to_full(amp, n, k), full)


def test_combinations_order_matches_colex():
    # independent enumeration: itertools.combinations sorted by value
    n, k = 10, 4
    xs = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(n), k))
    assert [O.unrank(r, k) for r in range(len(xs))] == xs
