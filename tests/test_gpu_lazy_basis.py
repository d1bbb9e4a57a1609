"""Basis states written lazily: reset(), setBasisState() and measureAll()
record the basis index; the next reader writes the vector, and the next
register-tile gate pass synthesizes the amplitudes in place of its first
load.  Every observable must be bit-identical to a state written eagerly
(setState of the same basis vector) -- default.cpp:57-65 (reset),
default.cpp:76-82 (setBasisState) -- except the sign of zero amplitudes in
the reordered mode, whose first pass stores the groups without the basis
index as +0 without running their ops.
"""
import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2

pytestmark = pytest.mark.gpu

O = Oracle()


def _ops(rng, n, count, dtype):
    kinds = GATE1 + GATE2
    ops = np.zeros(count, OP_DTYPE)
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ang = float(rng.uniform(-2 * np.pi, 2 * np.pi))
        if dtype == np.complex64:
            ang = float(np.float32(ang))
        ops[i] = (k, a, b if k >= 16 else 0, 0, ang)
    return ops


def _basis(n, idx, dtype):
    v = np.zeros(1 << n, dtype)
    v[idx] = 1
    return v


def _same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.fixture
def jit_mode():
    saved = L.tuning_get("jit")
    yield lambda v: L.tuning_set("jit", v)
    L.tuning_set("jit", saved)


MODES = [("fused", qsl_amd.MODE_FUSED), ("reorder", qsl_amd.MODE_REORDER),
         ("per_gate", qsl_amd.MODE_PER_GATE)]


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("mode", MODES, ids=lambda m: m[0])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [12, 20])
def test_gates_on_a_recorded_basis_state(jit_mode, jit, mode, dtype, n):
    # n = 20 takes the register-tile passes (synthesized first load), n = 12
    # the LDS tile kernel (the basis state is written first)
    jit_mode(jit)
    rng = np.random.default_rng(n * 10 + jit + (mode[1] << 4))
    ops = _ops(rng, n, 200, dtype)
    for idx in (0, int(rng.integers(1, 1 << n))):
        lazy = qsl_amd.Qubits(n, dtype=dtype)
        if idx:
            lazy.setBasisState(idx)
        lazy.apply_ops(ops, mode=mode[1])
        eager = qsl_amd.Qubits(n, dtype=dtype)
        eager.setState(_basis(n, idx, dtype))
        eager.apply_ops(ops, mode=mode[1])
        got = lazy.getState()
        if mode[1] == qsl_amd.MODE_REORDER:
            # reordered mode stores the groups without the basis index as
            # they are (+0) instead of running the ops on zeros, which may
            # leave -0: equal values, zero signs aside
            assert np.array_equal(got, eager.getState()), (mode[0], idx)
        else:
            assert _same(got, eager.getState()), (mode[0], idx)
        want = _basis(n, idx, dtype)
        O.apply_ops(want, ops)
        if mode[1] != qsl_amd.MODE_REORDER:
            assert _same(got, want), (mode[0], idx)
        else:
            assert np.abs(got - want).max() <= (1e-12 if dtype == np.complex128 else 1e-5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_readers_see_the_recorded_state(dtype):
    n = 20
    q = qsl_amd.Qubits(n, dtype=dtype)
    rng = np.random.default_rng(7)
    q.apply_ops(_ops(rng, n, 50, dtype), mode=qsl_amd.MODE_FUSED)
    # reset after gates: every reader sees |0...0>
    q.reset()
    assert q.prob(0, 1) == 0.0 and q.prob(5, 0) == 1.0
    assert q.norm() == 1.0
    assert _same(q.getState(), _basis(n, 0, dtype))
    idx = 0b1011_0000_0000_0110_1001
    q.setBasisState(idx)
    amps = q.getAmps(idx - 2, 5)
    assert _same(amps, np.array([0, 0, 1, 0, 0], dtype))
    q.setBasisState(idx)
    assert all(q.prob(t, (idx >> t) & 1) == 1.0 for t in range(n))
    q.setBasisState(idx)
    c = q.copy()
    assert _same(c.getState(), _basis(n, idx, dtype))
    q.setBasisState(3)
    q.setAmps(np.array([0.5, 0.5], dtype), first=1)  # a partial write keeps the rest
    want = _basis(n, 3, dtype)
    want[1:3] = 0.5
    assert _same(q.getState(), want)
    # assign from a handle holding a recorded state, and onto one
    a = qsl_amd.Qubits(n, dtype=dtype)
    a.setBasisState(9)
    b = qsl_amd.Qubits(n, dtype=dtype)
    b.assign(a)
    assert _same(b.getState(), _basis(n, 9, dtype))
    b.setBasisState(4)
    b.assign(a)
    assert _same(b.getState(), _basis(n, 9, dtype))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_refused_assign_leaves_recorded_state(dtype):
    # operator= throws on a qubit-count mismatch and leaves the object as it
    # was (default.cpp:94-105); a fresh handle only records |0...0>, so the
    # refusal must not drop that record
    a = qsl_amd.Qubits(10, dtype=dtype)
    b = qsl_amd.Qubits(12, dtype=dtype)
    with pytest.raises(qsl_amd.QslHipError):
        a.assign(b)
    assert _same(a.getState(), _basis(10, 0, dtype))
    a.setBasisState(5)
    with pytest.raises(qsl_amd.QslHipError):
        a.assign(b)
    assert _same(a.getState(), _basis(10, 5, dtype))
    # precision and type mismatches are refused the same way
    other = np.complex64 if dtype == np.complex128 else np.complex128
    c = qsl_amd.Qubits(10, dtype=other)
    a.setBasisState(7)
    with pytest.raises(qsl_amd.QslHipError):
        a.assign(c)
    assert _same(a.getState(), _basis(10, 7, dtype))
    d = qsl_amd.QubitsNP(10, nones=3, dtype=dtype)
    a.reset()
    with pytest.raises(qsl_amd.QslHipError):
        a.assign(d)
    assert _same(a.getState(), _basis(10, 0, dtype))


def test_measure_all_then_gates():
    # measureAll collapses onto the outcome (measure.cpp:110-123); the
    # collapsed basis state is recorded and synthesized by the next pass
    n = 20
    rng = np.random.default_rng(11)
    q = qsl_amd.Qubits(n)
    q.apply_ops(_ops(rng, n, 100, np.complex128), mode=qsl_amd.MODE_FUSED)
    o = q.measureAll(0.37)
    assert _same(q.getState(), _basis(n, o, np.complex128))
    o2 = q.measureAll(0.61)
    assert o2 == o
    ops = _ops(rng, n, 150, np.complex128)
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    want = _basis(n, o, np.complex128)
    O.apply_ops(want, ops)
    assert _same(q.getState(), want)


def test_empty_op_list_keeps_the_recorded_state():
    n = 20
    q = qsl_amd.Qubits(n)
    q.setBasisState(5)
    q.apply_ops(np.zeros(0, OP_DTYPE), mode=qsl_amd.MODE_FUSED)
    q.apply_ops(np.array([(0, 3, 0, 0, 0.0)], OP_DTYPE), mode=qsl_amd.MODE_FUSED)  # pauliX
    assert _same(q.getState(), _basis(n, 5 ^ 8, np.complex128))


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_reordered_passes_over_the_nonzero_groups_only(jit_mode, jit, dtype):
    # From a basis state the reordered passes run only the groups that can
    # be nonzero: those agreeing with the basis on every qubit no earlier
    # tile held.  A circuit whose qubits join slowly (gates on the low
    # qubits first, then higher ones, with controls and diagonal gates
    # reaching back) leaves most groups of the early passes out.
    jit_mode(jit)
    n = 22
    rng = np.random.default_rng(2200 + jit)
    parts = []
    for lo, hi in ((0, 8), (5, 14), (12, 22), (0, 22)):
        sub = _ops(rng, hi - lo, 60, dtype)
        sub["q0"] += lo
        sub["q1"] = np.where(sub["kind"] >= 16, sub["q1"] + lo, 0)
        parts.append(sub)
    ops = np.concatenate(parts)
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    for idx in (0, int(rng.integers(1, 1 << n))):
        q = qsl_amd.Qubits(n, dtype=dtype)
        q.setBasisState(idx)
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        want = _basis(n, idx, dtype)
        O.apply_ops(want, ops)
        got = q.getState()
        assert np.abs(got - want).max() <= tol, idx
        # and the same circuit applied again (every group, nothing recorded)
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        O.apply_ops(want, ops)
        assert np.abs(q.getState() - want).max() <= 2 * tol, idx


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_reordered_from_a_basis_state_on_a_used_vector(jit_mode, jit, dtype):
    # The vector still holds the previous state.  When every qubit is in some
    # tile the passes store nothing before computing it, reading only what an
    # earlier pass wrote; otherwise the first pass stores every group.  Both
    # must leave no trace of the old contents.
    jit_mode(jit)
    n = 20
    rng = np.random.default_rng(3300 + jit)
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    every = _ops(rng, n, 150, dtype)             # touches every qubit
    some = _ops(rng, 12, 80, dtype)              # qubits 0..11 only: 8 never in a tile
    for ops in (every, some):
        q = qsl_amd.Qubits(n, dtype=dtype)
        q.apply_ops(_ops(rng, n, 60, dtype), mode=qsl_amd.MODE_FUSED)  # old contents
        for idx in (0, int(rng.integers(1, 1 << n))):
            q.setBasisState(idx)
            q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
            want = _basis(n, idx, dtype)
            O.apply_ops(want, ops)
            got = q.getState()
            assert np.abs(got - want).max() <= tol, (len(ops), idx)
            assert abs(np.vdot(got, got).real - 1) <= 1e3 * tol
