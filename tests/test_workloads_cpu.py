"""Host-side workload helpers (no GPU): the basis string a number-preserving
workload starts from is the weight-k string of a given colex rank, the order
of the reference's generateLookup (src/utils/misc.cpp:75-90)."""
import math

import numpy as np

from qsl_amd import workloads as W
from oracle.pyoracle import Oracle


def test_colex_unrank_endpoints():
    # misc.cpp:79-81: the first weight-k string is 2^k - 1, the last one
    # (2^k - 1) << (n - k)
    for n, k in ((32, 16), (10, 3), (7, 7), (6, 1)):
        assert W.colex_unrank(0, k) == (1 << k) - 1
        assert W.colex_unrank(math.comb(n, k) - 1, k) == ((1 << k) - 1) << (n - k)


def test_colex_unrank_is_increasing_numeric_order():
    # brute force: the rank-r string is the r-th weight-k integer
    for n, k in ((10, 4), (9, 1), (8, 8)):
        strings = [x for x in range(1 << n) if bin(x).count("1") == k]
        assert [W.colex_unrank(r, k) for r in range(len(strings))] == strings


def test_colex_unrank_matches_the_oracle():
    o = Oracle()
    rng = np.random.default_rng(3)
    n, k = 32, 16
    total = math.comb(n, k)
    for r in [0, 1, total - 2, total - 1] + [int(x) for x in rng.integers(0, total, 40)]:
        assert W.colex_unrank(r, k) == o.unrank(r, k), r
