"""bench.py's byte count for number-preserving passes (np_pass_bytes),
checked against brute-force enumeration of the weight-k strings."""
import itertools
import math

import numpy as np

import bench
from qsl_amd import _lib as L


def _ops(rows):
    return np.array([(k, a, b, 0, 0.3) for k, a, b in rows], L.OP_DTYPE)


def _strings(n, k):
    for ones in itertools.combinations(range(n), k):
        yield sum(1 << q for q in ones)


def test_high_bond_pass_counts_the_strings_of_active_orbits():
    # n = 20, L = 16: pairs in the high part (qubits 16..19); the pass reads
    # and writes every string whose high part holds some pair with its two
    # bits different
    n, k = 20, 8
    pairs = [(16, 17), (19, 18)]
    ops = _ops([(L.NP_ROT_X, a, b) for a, b in pairs])
    want = sum(32 for x in _strings(n, k)
               if any(((x >> a) & 1) != ((x >> b) & 1) for a, b in pairs))
    assert bench.np_pass_bytes(ops, n, k) == want


def test_low_mid_and_phase_passes():
    n, k = 20, 8
    namps = math.comb(n, k)
    full = 2 * namps * 16
    pair = 2 * 2 * math.comb(n - 2, k - 1) * 16
    # a run of phases: one pass
    assert bench.np_pass_bytes(_ops([(L.PHASE, q, 0) for q in range(5)]), n, k) == full
    # disjoint adjacent bonds below qubit 12: one pass
    assert bench.np_pass_bytes(_ops([(L.NP_ROT_Y, 0, 1), (L.NP_ROT_X, 3, 2)]), n, k) == full
    # a single bond between 12 and 16: its pairs only
    assert bench.np_pass_bytes(_ops([(L.NP_ROT_X, 13, 14)]), n, k) == pair
    # the brute-force pair count of that bond agrees
    pairs = sum(2 for x in _strings(n, k) if ((x >> 13) & 1) != ((x >> 14) & 1))
    assert pair == pairs * 16
