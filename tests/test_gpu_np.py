"""GPU parity of the compact number-preserving backend.

Checked against (a) the oracle's NP semantics on the full 2^n vector (the
reference's NP backend restated, src/qubits/np/*.cpp) at n <= 14, bit for
bit, and (b) the oracle's compact colex layout (the same amplitude order the
device uses) at n = 20..24, and (c) size-independent properties at n = 32,
k = 16 (config 3): norm preservation, a gate sequence followed by its
inverse returns the exact start state, and rank/unrank round trips via
set_basis / get_amps.
"""
import math

import numpy as np
import pytest

import qsl_amd
from oracle.pyoracle import (Oracle, OP_DTYPE, NP_GATE1, NP_GATE2, HAS_ANGLE,
                             near_cdf_boundary, random_np_state)

pytestmark = pytest.mark.gpu
O = Oracle()


def _bits_equal(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _angle(rng, dtype):
    a = float(rng.uniform(-2 * np.pi, 2 * np.pi))
    return float(np.float32(a)) if dtype == np.complex64 else a


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(4, 2), (7, 3), (10, 5), (13, 4)])
def test_np_gates_vs_full_vector_semantics(dtype, n, k):
    rng = np.random.default_rng(n * 100 + k)
    st0 = random_np_state(rng, n, k, dtype)
    q = qsl_amd.QubitsNP(n, k, dtype=dtype)
    for kind in NP_GATE1:
        for t in range(n):
            ang = _angle(rng, dtype) if kind in HAS_ANGLE else 0.0
            q.setState(st0)
            q._g1(kind, t, ang)
            want = st0.copy()
            O.np_gate1(want, k, kind, t, ang)
            assert _bits_equal(q.getState(), want), (kind, t)
    pairs = [(a, b) for a in range(n) for b in range(n) if a != b]
    if len(pairs) > 30:
        pairs = [pairs[i] for i in rng.choice(len(pairs), 30, replace=False)]
    for kind in NP_GATE2:
        for a, b in pairs:
            ang = _angle(rng, dtype) if kind in HAS_ANGLE else 0.0
            q.setState(st0)
            q._g2(kind, a, b, ang)
            want = st0.copy()
            O.np_gate2(want, k, kind, a, b, ang)
            assert _bits_equal(q.getState(), want), (kind, a, b)


def _np_ops(rng, n, count):
    ops = np.zeros(count, OP_DTYPE)
    kinds = NP_GATE1 + NP_GATE2
    for i in range(count):
        kind = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops[i] = (kind, a, b if kind >= 16 else 0, 0, rng.uniform(0, 2 * np.pi))
    return ops


@pytest.mark.parametrize("n,k", [(20, 10), (22, 7), (24, 12)])
def test_np_compact_vs_compact_oracle(n, k):
    rng = np.random.default_rng(n + k)
    total = math.comb(n, k)
    amp = (rng.normal(size=total) + 1j * rng.normal(size=total)).astype(np.complex128)
    amp /= np.linalg.norm(amp)
    q = qsl_amd.QubitsNP(n, k)
    q.setAmps(amp)
    ops = _np_ops(rng, n, 40)
    # adjacent bonds of both kinds of segments (low part, high part, across)
    for b in (0, 7, 14, 15, 16, n - 2):
        ops = np.concatenate([ops, np.array([(26, b, b + 1, 0, 0.7), (27, b + 1, b, 0, 1.1)],
                                            OP_DTYPE)])
    q.apply_ops(ops)
    want = amp.copy()
    O.npc_apply_ops(want, n, k, ops, nthreads=8)
    got = q.getAmps()
    assert np.abs(got - want).max() <= 1e-12
    assert _bits_equal(got, want)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_np_measurement(dtype):
    n, k = 11, 5
    rng = np.random.default_rng(11)
    st = random_np_state(rng, n, k, dtype)
    q = qsl_amd.QubitsNP(st)
    assert q.getNumOnes() == k
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    for t in range(n):
        for o in (0, 1):
            assert abs(q.prob(t, o) - O.np_prob(st, k, t, o)) <= tol
    u = qsl_amd.uniforms(3, 5000)
    got = q.sampleAllOutcomes(u)
    uf = u.astype(np.float32) if dtype == np.complex64 else u
    want = O.sample_all(st, uf)
    # exact outside draws within eps of a CDF value (FP64 1e-12, FP32 1e-5)
    _, cdf = O.gen_dist(st)
    eps = 1e-12 if dtype == np.complex128 else 1e-5
    ok = (got == want) | near_cdf_boundary(cdf, np.asarray(uf, np.float64), eps)
    assert ok.all(), np.nonzero(~ok)[0][:10]
    assert all(bin(int(x)).count("1") == k for x in got)
    # measure / postselect against the full-vector oracle
    q2 = qsl_amd.QubitsNP(st)
    out = q2.measure(3, 0.4)
    w = st.copy()
    assert out == O.measure(w, 3, 0.4)
    assert np.abs(q2.getState() - w).max() <= 10 * tol


def test_np_state_management():
    q = qsl_amd.QubitsNP(6, 2)
    st = q.getState()
    assert np.nonzero(st)[0].tolist() == [3]  # np.cpp:77-78
    q.setBasisState(0b101100)
    assert q.getNumOnes() == 3
    assert np.nonzero(q.getState())[0].tolist() == [0b101100]
    q.setNumOnes(4)
    assert np.nonzero(q.getState())[0].tolist() == [0b1111]
    with pytest.raises(qsl_amd.NotNPError):
        q.setNumOnes(6)
    with pytest.raises(qsl_amd.NotNPError):
        qsl_amd.QubitsNP(4, 5)
    bad = np.zeros(64, np.complex128)
    bad[1] = bad[3] = 1 / np.sqrt(2)
    with pytest.raises(qsl_amd.NotNPError):
        q.setState(bad)
    with pytest.raises(qsl_amd.SizeError):
        q.setState(np.zeros(32, np.complex128))
    with pytest.raises(qsl_amd.QslHipError):
        q.hadamard(0)  # not number preserving


def test_np_n32_k16_properties():
    n, k = 32, 16
    q = qsl_amd.QubitsNP(n, k)
    total = math.comb(n, k)
    assert q.numAmps() == total
    rng = np.random.default_rng(32)
    # rank/unrank: set a random basis string, read back its compact index
    for _ in range(5):
        bits = rng.choice(n, k, replace=False)
        x = int(sum(1 << int(b) for b in bits))
        q.setBasisState(x)
        r = O.rank(x)
        assert q.getAmps(r, 1)[0] == 1
        assert abs(q.norm() - 1) < 1e-15
        out = q.sampleAllOutcomes(np.array([0.5]))
        assert int(out[0]) == x
    # forward layers then exact inverses return to the start string
    x0 = O.unrank(total // 3, k)
    q.setBasisState(x0)
    ops = []
    for layer in range(2):
        for b in range(layer % 2, n - 1, 2):
            ops.append((26 + (b % 2), b, b + 1, 0, 0.3 + 0.01 * b))
        for t in range(n):
            ops.append((4, t, 0, 0, 0.1 * t))
    ops = np.array(ops, OP_DTYPE)
    q.apply_ops(ops)
    assert abs(q.norm() - 1) < 1e-12
    inv = ops[::-1].copy()
    inv["angle"] = -inv["angle"]
    q.apply_ops(inv)
    amp = q.getAmps(O.rank(x0), 1)[0]
    assert abs(amp - 1) < 1e-12
    assert abs(q.norm() - 1) < 1e-12


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(10, 4), (20, 10)])
def test_np_diagonal_runs_in_one_pass(dtype, n, k):
    # runs of diagonal gates (phase, pauliZ, rotateZ, controlZ, controlPhase,
    # controlRotateZ) go through one pass each, every amplitude taking the
    # run's factors in program order: bit-identical to one kernel per gate
    rng = np.random.default_rng(500 + n + k)
    diag1, diag2 = [2, 4, 7], [18, 21, 22]
    seq = []
    for layer in range(4):
        for q0 in range(layer % 2, n - 1, 2):  # adjacent bonds, as in config 3
            seq.append((int(rng.choice([26, 27])), q0, q0 + 1, 0, _angle(rng, dtype)))
        for _ in range(3 * n):
            if rng.random() < 0.5:
                seq.append((int(rng.choice(diag1)), int(rng.integers(n)), 0, 0, _angle(rng, dtype)))
            else:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                seq.append((int(rng.choice(diag2)), a, b, 0, _angle(rng, dtype)))
    ops = np.array(seq, OP_DTYPE)
    st0 = random_np_state(rng, n, k, dtype)
    fused = qsl_amd.QubitsNP(n, k, dtype=dtype)
    fused.setState(st0)
    fused.apply_ops(ops)
    per_gate = qsl_amd.QubitsNP(n, k, dtype=dtype)
    per_gate.setState(st0)
    per_gate.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    assert _bits_equal(fused.getAmps(), per_gate.getAmps())
    if n <= 10:
        want = st0.copy()
        for o in ops:
            kind, a, b, ang = int(o["kind"]), int(o["q0"]), int(o["q1"]), float(o["angle"])
            if kind >= 16:
                O.np_gate2(want, k, kind, a, b, ang)
            else:
                O.np_gate1(want, k, kind, a, ang)
        assert _bits_equal(fused.getState(), want)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(10, 4), (20, 10), (24, 12), (34, 3)])
def test_np_phase_runs_in_one_pass(dtype, n, k):
    # runs made only of phase gates take the masked pass (per string, the
    # run indices acting on it; gates on segment bits applied to the whole
    # workgroup): runs of 2 to 70 gates (split into passes of 32), qubits
    # repeated, low and segment qubits interleaved, and the config-3 shape
    # (every qubit once, in order).  n = 34 has a 3-byte low part.  The
    # result is bit-identical to one kernel per gate and to the generic
    # diagonal pass.
    rng = np.random.default_rng(700 + n + k)
    seq = []
    for length in (2, 5, 31, 32, 33, 70):
        for q0 in range(length % 2, n - 1, 2):  # bonds between the runs
            seq.append((int(rng.choice([26, 27])), q0, q0 + 1, 0, _angle(rng, dtype)))
        qs = np.arange(length) % n if length == 32 else rng.integers(n, size=length)
        for q in qs:
            seq.append((4, int(q), 0, 0, _angle(rng, dtype)))
    ops = np.array(seq, OP_DTYPE)
    total = math.comb(n, k)
    amp = (rng.normal(size=total) + 1j * rng.normal(size=total))
    amp = (amp / np.linalg.norm(amp)).astype(dtype)
    lib = qsl_amd._lib
    saved = lib.tuning_get("np_phase_lut")
    got = {}
    try:
        for lut in (1, 0):
            lib.tuning_set("np_phase_lut", lut)
            q = qsl_amd.QubitsNP(n, k, dtype=dtype)
            q.setAmps(amp)
            q.apply_ops(ops)
            got[lut] = q.getAmps()
    finally:
        lib.tuning_set("np_phase_lut", saved)
    per_gate = qsl_amd.QubitsNP(n, k, dtype=dtype)
    per_gate.setAmps(amp)
    per_gate.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    assert _bits_equal(got[1], per_gate.getAmps())
    assert _bits_equal(got[0], got[1])


@pytest.fixture
def bond_groups():
    saved = qsl_amd._lib.tuning_get("np_bonds")
    qsl_amd._lib.tuning_set("np_bonds", 4)
    yield
    qsl_amd._lib.tuning_set("np_bonds", saved)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(12, 5), (20, 10), (24, 12)])
def test_np_disjoint_bond_groups_in_one_pass(bond_groups, dtype, n, k):
    # consecutive swap-like gates (npRotateX/Y, npHadamard, swap) on disjoint
    # qubit pairs run in groups of up to four per pass, each orbit of strings
    # gathered once: bit-identical to one kernel per gate, for pairs inside
    # the low part, the high part and across
    rng = np.random.default_rng(700 + n + k)
    seq = []
    for _ in range(40):
        qs = rng.permutation(n)
        for j in range(int(rng.integers(1, 6))):  # 5 pairs: a group of 4 and a single
            a, b = int(qs[2 * j]), int(qs[2 * j + 1])
            seq.append((int(rng.choice([25, 26, 27, 28])), a, b, 0, _angle(rng, dtype)))
        if rng.random() < 0.3:
            seq.append((4, int(rng.integers(n)), 0, 0, _angle(rng, dtype)))  # breaks a group
    ops = np.array(seq, OP_DTYPE)
    st0 = random_np_state(rng, n, k, dtype) if n <= 14 else None
    fused = qsl_amd.QubitsNP(n, k, dtype=dtype)
    per_gate = qsl_amd.QubitsNP(n, k, dtype=dtype)
    if st0 is not None:
        fused.setState(st0)
        per_gate.setState(st0)
    else:
        total = math.comb(n, k)
        amp = (rng.normal(size=total) + 1j * rng.normal(size=total)).astype(dtype)
        amp /= np.linalg.norm(amp)
        fused.setAmps(amp)
        per_gate.setAmps(amp)
    fused.apply_ops(ops)
    per_gate.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    assert _bits_equal(fused.getAmps(), per_gate.getAmps())
    if st0 is not None:
        want = st0.copy()
        for o in ops:
            kind, a, b, ang = int(o["kind"]), int(o["q0"]), int(o["q1"]), float(o["angle"])
            if kind >= 16:
                O.np_gate2(want, k, kind, a, b, ang)
            else:
                O.np_gate1(want, k, kind, a, ang)
        assert _bits_equal(fused.getState(), want)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(14, 6), (20, 10), (22, 4), (26, 13)])
def test_np_fused_low_and_high_bond_passes(dtype, n, k):
    # groups of swap-like gates on disjoint adjacent pairs below qubit 12
    # (one pass over runs of strings that share the bits above, staged in
    # LDS) and groups of up to four on disjoint pairs in the high part of
    # the strings (one pass over orbits of segments): bit-identical to one
    # kernel per gate, for every swap-like kind and both orientations
    rng = np.random.default_rng(900 + n + k)
    L = min(n, max(16, n - 16))
    seq = []
    for rep in range(12):
        lows = list(range(rep % 2, min(L, 12) - 1, 2))
        for q in rng.permutation(lows)[: int(rng.integers(2, 7))]:
            a, b = (int(q), int(q) + 1) if rng.random() < 0.5 else (int(q) + 1, int(q))
            seq.append((int(rng.choice([24, 26, 27, 28])), a, b, 0, _angle(rng, dtype)))
        if n - L >= 4:
            hs = rng.permutation(np.arange(L, n))
            for j in range(min(4, (n - L) // 2)):
                seq.append((int(rng.choice([24, 26, 27, 28])), int(hs[2 * j]), int(hs[2 * j + 1]),
                            0, _angle(rng, dtype)))
        seq.append((4, int(rng.integers(n)), 0, 0, _angle(rng, dtype)))
    ops = np.array(seq, OP_DTYPE)
    total = math.comb(n, k)
    amp = (rng.normal(size=total) + 1j * rng.normal(size=total)).astype(dtype)
    amp /= np.linalg.norm(amp)
    fused = qsl_amd.QubitsNP(n, k, dtype=dtype)
    per_gate = qsl_amd.QubitsNP(n, k, dtype=dtype)
    fused.setAmps(amp)
    per_gate.setAmps(amp)
    fused.apply_ops(ops)
    per_gate.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    assert _bits_equal(fused.getAmps(), per_gate.getAmps())
    if dtype == np.complex128 and n <= 22:
        want = amp.astype(np.complex128)
        O.npc_apply_ops(want, n, k, ops, nthreads=8)
        assert _bits_equal(fused.getAmps(), want)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k", [(32, 4), (32, 6), (30, 8), (28, 14)])
def test_np_high_bond_orbits_staged_in_lds(dtype, n, k):
    # groups of five to eight swap-like gates on disjoint pairs in the high
    # part of the strings (tuning np_hibonds 2: one pass over orbits of up
    # to 2^8 segments staged through LDS): bit-identical to one kernel per
    # gate, for every swap-like kind and both orientations
    rng = np.random.default_rng(1100 + n + k)
    L = min(n, max(16, n - 16))
    npairs = (n - L) // 2
    seq = []
    for rep in range(6):
        hs = rng.permutation(np.arange(L, n))
        for j in range(int(rng.integers(5, npairs + 1))):
            seq.append((int(rng.choice([24, 26, 27, 28])), int(hs[2 * j]), int(hs[2 * j + 1]),
                        0, _angle(rng, dtype)))
        seq.append((4, int(rng.integers(n)), 0, 0, _angle(rng, dtype)))
    ops = np.array(seq, OP_DTYPE)
    total = math.comb(n, k)
    amp = (rng.normal(size=total) + 1j * rng.normal(size=total)).astype(dtype)
    amp /= np.linalg.norm(amp)
    saved = qsl_amd._lib.tuning_get("np_hibonds")
    try:
        qsl_amd._lib.tuning_set("np_hibonds", 2)
        fused = qsl_amd.QubitsNP(n, k, dtype=dtype)
        fused.setAmps(amp)
        fused.apply_ops(ops)
        got = fused.getAmps()
        del fused
    finally:
        qsl_amd._lib.tuning_set("np_hibonds", saved)
    per_gate = qsl_amd.QubitsNP(n, k, dtype=dtype)
    per_gate.setAmps(amp)
    per_gate.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    assert _bits_equal(got, per_gate.getAmps())
