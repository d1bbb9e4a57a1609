"""Host-side check of the dense-fusion planner (no GPU needed).

qslhip_plan_dense returns the plan QSLHIP_MODE_DENSE would execute: passes
of ordinary tile ops and fused dense 2^F x 2^F unitaries.  Replaying that
plan on the CPU -- tile ops through the oracle, dense ops as the returned
FP32 matrices applied to their qubits -- must reproduce the oracle's
gate-by-gate FP64 state to FP32 matrix rounding.  A wrong commutation
(a gate moved past one it does not commute with), a wrong product order or
a transposed matrix fails this by O(1).
"""
import numpy as np
import pytest

from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2, random_state

O = Oracle()


def apply_dense(st, qubits, M):
    n = int(st.size).bit_length() - 1
    F = len(qubits)
    mask = sum(1 << q for q in qubits)
    rest = np.array([i for i in range(1 << n) if not (i & mask)], np.int64)
    off = np.zeros(1 << F, np.int64)
    for t in range(1 << F):
        off[t] = sum(1 << qubits[j] for j in range(F) if (t >> j) & 1)
    idx = rest[None, :] | off[:, None]
    st[idx] = M @ st[idx]


def replay(st, ops, pi, oi, mats):
    for p in range(len(pi)):
        first, count = int(pi[p, 1]), int(pi[p, 2])
        for r in oi[first:first + count]:
            kind, gidx, nq, moff = (int(v) for v in r[:4])
            if kind == 0:
                op = ops[gidx]
                k = int(op["kind"])
                if k < 16:
                    O.gate1(st, k, int(op["q0"]), float(op["angle"]))
                else:
                    O.gate2(st, k, int(op["q0"]), int(op["q1"]), float(op["angle"]))
            else:
                d = 1 << nq
                m = mats[moff:moff + 2 * d * d].astype(np.float64)
                M = (m[0::2] + 1j * m[1::2]).reshape(d, d)
                apply_dense(st, [int(q) for q in r[4:4 + nq]], M)


@pytest.fixture
def dense_reorder():
    saved = L.tuning_get("dense_reorder")
    yield lambda v: L.tuning_set("dense_reorder", v)
    L.tuning_set("dense_reorder", saved)


@pytest.mark.parametrize("reorder", [0, 1])
@pytest.mark.parametrize("F", [4, 5])
@pytest.mark.parametrize("n", [12, 13, 14])
def test_dense_plan_replays_to_oracle(dense_reorder, reorder, F, n):
    # program-order tile passes and commutation-aware ones (tuning
    # dense_reorder): either plan replays to the gate-by-gate state
    dense_reorder(reorder)
    rng = np.random.default_rng(10 * n + F)
    ops = np.zeros(250, OP_DTYPE)
    for i in range(len(ops)):
        k = int(rng.choice(GATE1 + GATE2))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops[i] = (k, a, b if k >= 16 else 0, 0, float(rng.uniform(0, 2 * np.pi)))
    pi, oi, mats = L.plan_dense(n, ops, F)
    assert (oi[:, 0] == 1).sum() > 0, "no dense op was formed"
    st0 = random_state(rng, n)
    got = st0.copy()
    replay(got, ops, pi, oi, mats)
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.abs(got - want).max() < 2e-6


def test_haar_layer_fusion_is_dense():
    from qsl_amd import workloads as W
    n = 16
    ops, _ = W.c5_haar_layers(n, 4, W.Stream(5))
    pi, oi, mats = L.plan_dense(n, ops, 5)
    ndense = int((oi[:, 0] == 1).sum())
    assert ndense >= 4 * 3  # every layer's rotations fuse into 5-qubit blocks
    st0 = np.zeros(1 << n, np.complex128)
    st0[0] = 1
    got = st0.copy()
    replay(got, ops, pi, oi, mats)
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.abs(got - want).max() < 2e-6
