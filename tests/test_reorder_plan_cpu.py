"""Host-side checks of the commutation-aware planner (QSLHIP_MODE_REORDER).

qslhip_plan_mode returns the passes and the order in which the gates run.
The order must be a permutation of the gate list that keeps program order
for every pair of gates that share a qubit unless both are diagonal; each
pass must hold the non-diagonal targets of its gates.  Replaying the order
through the oracle must give the program-order state to FP64 rounding.
A plan that swapped two non-commuting gates (say a Hadamard past a CNOT on
the same qubit) fails both the order check and the replay by O(1).
"""
import numpy as np
import pytest

from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2, random_state

O = Oracle()
DIAG = {2, 4, 7, 18, 21, 22}  # pauliZ, phase, rotateZ, CZ, CRZ, CPhase
NONDIAG1 = {0, 1, 3, 5, 6}


def qubits(op):
    k = int(op["kind"])
    return {int(op["q0"])} if k < 16 else {int(op["q0"]), int(op["q1"])}


def targets(op):
    k, a, b = int(op["kind"]), int(op["q0"]), int(op["q1"])
    if k < 16:
        return {a} if k in NONDIAG1 else set()
    if k in (16, 17, 19, 20, 23):
        return {b}
    if k in (24, 25, 26, 27, 28):
        return {a, b}
    return set()


def random_ops(rng, n, count, kinds):
    ops = np.zeros(count, OP_DTYPE)
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops[i] = (k, a, b if k >= 16 else 0, 0, float(rng.uniform(-6, 6)))
    return ops


@pytest.fixture
def widths():
    saved = {k: L.tuning_get(k) for k in ("logt_f64", "low_f64", "reg_tile")}
    L.tuning_set("reg_tile", 1)
    yield
    try:
        L.tuning_set("logt_f64", saved["logt_f64"])
        L.tuning_set("low_f64", saved["low_f64"])
    except L.QslHipError:
        L.tuning_set("low_f64", saved["low_f64"])
        L.tuning_set("logt_f64", saved["logt_f64"])
    L.tuning_set("reg_tile", saved["reg_tile"])


@pytest.mark.parametrize("n,logt,low", [(9, 10, 3), (12, 10, 3), (14, 11, 3), (16, 12, 4)])
def test_reorder_plan_is_a_valid_schedule(widths, n, logt, low):
    try:
        L.tuning_set("logt_f64", logt)
        L.tuning_set("low_f64", low)
    except L.QslHipError:
        L.tuning_set("low_f64", low)
        L.tuning_set("logt_f64", logt)
    rng = np.random.default_rng(900 + n)
    ops = random_ops(rng, n, 300, GATE1 + GATE2)
    masks, counts, order = L.plan_order(n, 8, ops, L.MODE_REORDER)
    order = [int(x) for x in order]
    assert sorted(order) == list(range(len(ops)))
    assert int(counts.sum()) == len(ops)
    pos = np.empty(len(ops), int)
    pos[order] = np.arange(len(ops))
    for i in range(len(ops)):
        for j in range(i + 1, len(ops)):
            if qubits(ops[i]) & qubits(ops[j]):
                both_diag = int(ops[i]["kind"]) in DIAG and int(ops[j]["kind"]) in DIAG
                if not both_diag:
                    assert pos[i] < pos[j], (i, j)
    # the tile width narrows on small states (at least 64 workgroups), not
    # below 2^10 while register tiles are on
    K = min(logt, max(10, n - 6), n)
    first = 0
    for m, c in zip(masks, counts):
        m = int(m)
        assert bin(m).count("1") == K
        assert m & ((1 << min(low, K)) - 1) == (1 << min(low, K)) - 1
        for g in order[first:first + int(c)]:
            for q in targets(ops[g]):
                assert (m >> q) & 1, (g, q)
        first += int(c)
    # the same product: replay the execution order through the oracle
    st0 = random_state(rng, n)
    want = st0.copy()
    O.apply_ops(want, ops)
    got = st0.copy()
    O.apply_ops(got, ops[order])
    assert np.abs(got - want).max() <= 1e-12


def test_reorder_packs_fewer_passes_than_program_order(widths):
    from qsl_amd import workloads as W
    L.tuning_set("logt_f64", 11)
    L.tuning_set("low_f64", 3)
    ops, _ = W.c2_random(30, 200, W.Stream(2))
    fused = len(L.plan_order(30, 8, ops, L.MODE_FUSED)[0])
    reord = len(L.plan_order(30, 8, ops, L.MODE_REORDER)[0])
    assert reord <= fused // 2 + 1, (fused, reord)
