"""Expectation values and adjoint gradients on the GPU.

Pins (the reference has no gradients, so these are mathematical):
  * E = <psi|H|psi> against numpy on the oracle's FP64 state;
  * plain (per-gate) FP64 adjoint gradients against the parameter-shift
    rule dE/dt = [E(t + pi/2) - E(t - pi/2)] / 2, exact for rotateX/Y/Z and
    phase, evaluated on the oracle (tolerance 1e-10);
  * controlled / number-preserving rotations against central finite
    differences, h = 1e-3, on the oracle (tolerance 1e-6);
  * the batched FP32 kernel against the FP64 plain path and against FD,
    max gradient error <= 1e-4 (the stated bound), energies <= 1e-5;
  * hipGraph replays with new angles, several in a row, each checked.
"""
import math

import numpy as np
import pytest

import qsl_amd
from qsl_amd import workloads as W
from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2

pytestmark = pytest.mark.gpu
O = Oracle()
PAULI = qsl_amd.PAULI_DTYPE


def pauli_apply(psi, terms):
    n = int(psi.size).bit_length() - 1
    idx = np.arange(psi.size, dtype=np.int64)
    out = np.zeros_like(psi, dtype=np.complex128)
    for c, x, z in terms:
        x, z = int(x), int(z)
        j = idx ^ x
        ny = bin(x & z).count("1") % 4
        parity = np.zeros(psi.size, np.int64)
        jz = j & z
        for q in range(n):
            parity ^= (jz >> q) & 1
        out += c * (1j ** ny) * np.where(parity == 1, -1.0, 1.0) * psi[j]
    return out


def energy(psi, terms):
    psi = psi.astype(np.complex128)
    return float(np.real(np.vdot(psi, pauli_apply(psi, terms))))


def oracle_energy(n, ops, terms):
    st = np.zeros(1 << n, np.complex128)
    st[0] = 1
    O.apply_ops(st, ops)
    return energy(st, terms)


def random_terms(rng, n, count):
    t = []
    for _ in range(count):
        x = int(rng.integers(0, 1 << n))
        z = int(rng.integers(0, 1 << n))
        t.append((float(rng.normal()), x, z))
    return np.array(t, PAULI)


def random_param_circuit(rng, n, count, kinds):
    ops = np.zeros(count, OP_DTYPE)
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        ops[i] = (k, a, b if k >= 16 else 0, 0, float(rng.uniform(-np.pi, np.pi)))
    return ops


def test_expval_matches_numpy():
    rng = np.random.default_rng(0)
    n = 7
    st = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n))
    st /= np.linalg.norm(st)
    terms = random_terms(rng, n, 6)
    q = qsl_amd.Qubits(st)
    assert abs(q.expval(terms) - energy(st, terms)) < 1e-12


@pytest.mark.parametrize("n", [3, 6, 9])
def test_plain_fp64_vs_parameter_shift(n):
    rng = np.random.default_rng(n)
    ops = random_param_circuit(rng, n, 30, [5, 6, 7, 4, 16, 18, 3])
    terms = random_terms(rng, n, 5)
    q = qsl_amd.Qubits(n)
    e, g = q.expvalGrad(ops, terms)
    assert abs(e - oracle_energy(n, ops, terms)) < 1e-12
    for i, op in enumerate(ops):
        if int(op["kind"]) in (4, 5, 6, 7):
            p, m = ops.copy(), ops.copy()
            p[i]["angle"] += math.pi / 2
            m[i]["angle"] -= math.pi / 2
            want = (oracle_energy(n, p, terms) - oracle_energy(n, m, terms)) / 2
            assert abs(g[i] - want) < 1e-10, (i, op, g[i], want)
        else:
            assert g[i] == 0
    # the handle keeps psi = U|0>
    st = np.zeros(1 << n, np.complex128)
    st[0] = 1
    O.apply_ops(st, ops)
    assert np.abs(q.getState() - st).max() < 1e-12


def test_plain_fp64_controlled_and_np_vs_fd():
    rng = np.random.default_rng(11)
    n = 5
    ops = random_param_circuit(rng, n, 24, [19, 20, 21, 22, 26, 27, 5, 16])
    terms = random_terms(rng, n, 4)
    q = qsl_amd.Qubits(n)
    e, g = q.expvalGrad(ops, terms)
    h = 1e-3
    for i in range(len(ops)):
        p, m = ops.copy(), ops.copy()
        p[i]["angle"] += h
        m[i]["angle"] -= h
        want = (oracle_energy(n, p, terms) - oracle_energy(n, m, terms)) / (2 * h)
        assert abs(g[i] - want) < 1e-6, (i, ops[i], g[i], want)


def test_plain_fp32_close_to_fp64():
    rng = np.random.default_rng(12)
    n = 8
    ops, _ = W.c4_ansatz(n, 8, W.Stream(40))
    terms = W.c4_observable(n)
    e64, g64 = qsl_amd.Qubits(n).expvalGrad(ops, terms)
    e32, g32 = qsl_amd.Qubits(n, dtype=np.complex64).expvalGrad(ops, terms)
    assert abs(e32 - e64) < 1e-5
    assert np.abs(g32 - g64).max() < 1e-4


@pytest.mark.parametrize("sizes", [[4, 7, 10, 12], [13, 14], [15, 16, 11]])
def test_batch_vs_fp64_plain(sizes):
    circuits = []
    s = W.Stream(4)
    for n in sizes:
        ops, _ = W.c4_ansatz(n, 8, s)
        circuits.append((n, ops, W.c4_observable(n)))
    b = qsl_amd.Batch(circuits)
    e, g = b.run()
    off = 0
    for c, (n, ops, terms) in enumerate(circuits):
        e64, g64 = qsl_amd.Qubits(n).expvalGrad(ops, terms)
        assert abs(e[c] - e64) < 1e-5, (n, e[c], e64)
        gc = g[off:off + len(ops)]
        assert np.abs(gc - g64).max() < 1e-4, (n, np.abs(gc - g64).max())
        off += len(ops)


@pytest.mark.parametrize("sizes", [[1, 2, 3, 5], [8, 12], [13, 15]])
def test_batch_every_gate_kind_vs_fp64_plain(sizes):
    # the batch kernel applies each pass's ops in register blocks of three
    # tile positions: every one- and two-qubit kind (controls, swaps, the
    # number-preserving rotations, diagonal gates on qubits outside the
    # tile), against the FP64 per-circuit path, and the energy against the
    # oracle
    rng = np.random.default_rng(sum(sizes))
    circuits = []
    for n in sizes:
        if n == 1:
            ops = np.zeros(40, OP_DTYPE)
            ops["kind"] = rng.choice(GATE1, 40)
            ops["angle"] = rng.uniform(-np.pi, np.pi, 40)
        else:
            ops = random_param_circuit(rng, n, 40 + 6 * n, GATE1 + GATE2)
        circuits.append((n, ops, random_terms(rng, n, 4)))
    e, g = qsl_amd.Batch(circuits).run()
    off = 0
    for c, (n, ops, terms) in enumerate(circuits):
        e64, g64 = qsl_amd.Qubits(n).expvalGrad(ops, terms)
        assert abs(e[c] - e64) < 1e-5, (n, e[c], e64)
        assert abs(e64 - oracle_energy(n, ops, terms)) < 1e-12, n
        gc = g[off:off + len(ops)]
        assert np.abs(gc - g64).max() < 1e-4, (n, np.abs(gc - g64).max())
        off += len(ops)


def test_batch_each_kind_at_each_placement():
    # every kind on every qubit (pair) of a 3-qubit circuit, between
    # rotations on all three qubits: the whole state is one 8-amplitude
    # register block, so each kind runs at each of its block-position
    # patterns, forward and inverse
    rng = np.random.default_rng(7)
    n = 3
    terms = np.array([(0.7, 0b011, 0b101), (0.4, 0b100, 0b010), (0.3, 0b001, 0b110)], PAULI)
    circuits = []
    for k in GATE1 + GATE2:
        places = ([(q, 0) for q in range(n)] if k < 16 else
                  [(a, b) for a in range(n) for b in range(n) if a != b])
        for a, b in places:
            seq = [(L.ROT_X, 0, 0), (L.ROT_Y, 1, 0), (L.ROT_X, 2, 0), (k, a, b),
                   (L.ROT_Y, 0, 0), (L.ROT_X, 1, 0), (L.ROT_Y, 2, 0)]
            ops = np.zeros(len(seq), OP_DTYPE)
            for i, (kk, q0, q1) in enumerate(seq):
                ops[i] = (kk, q0, q1, 0, float(rng.uniform(-3, 3)))
            circuits.append((n, ops, terms))
    e, g = qsl_amd.Batch(circuits).run()
    off = 0
    for c, (nn, ops, t) in enumerate(circuits):
        e64, g64 = qsl_amd.Qubits(nn).expvalGrad(ops, t)
        assert abs(e[c] - e64) < 1e-5, (c, ops[3])
        assert np.abs(g[off:off + len(ops)] - g64).max() < 1e-4, (c, ops[3])
        off += len(ops)


def test_batch_circuits_without_ops():
    # a circuit with no ops keeps |0...0>: energy <0|H|0>, no gradients --
    # on the whole-state-in-LDS path (5 qubits) and the tile path (13),
    # next to circuits with ops that use the same workgroup slots first
    rng = np.random.default_rng(21)
    circuits = []
    for n in (13, 5, 14, 13, 6):
        ops = np.zeros(0, OP_DTYPE) if n in (13, 5) else W.c4_ansatz(n, 2, W.Stream(n))[0]
        circuits.append((n, ops, random_terms(rng, n, 5)))
    e, g = qsl_amd.Batch(circuits).run()
    for c, (n, ops, terms) in enumerate(circuits):
        assert abs(e[c] - oracle_energy(n, ops, terms)) < 1e-5, (c, n)
    assert len(g) == sum(len(c[1]) for c in circuits)


def test_batch_circuits_on_a_few_qubits():
    # gates on qubits 0..9 only (plus diagonal gates on high qubits): every
    # forward group with a high bit set stays zero and is stored without its
    # ops, including the last groups of each pass; each pass must finish
    # those stores before the next one restages its offsets
    rng = np.random.default_rng(33)
    circuits = []
    for n in (16, 14, 17):
        ops = random_param_circuit(rng, 10, 120, GATE1 + GATE2)
        diag = np.zeros(6, OP_DTYPE)
        diag["kind"] = [4, 7, 4, 7, 4, 7]  # phase / rotateZ on high qubits
        diag["q0"] = [n - 1, n - 2, n - 1, n - 3, n - 2, n - 1]
        diag["angle"] = rng.uniform(-np.pi, np.pi, 6)
        ops = np.concatenate([ops[:60], diag, ops[60:]])
        circuits.append((n, ops, random_terms(rng, n, 4)))
    e, g = qsl_amd.Batch(circuits).run()
    off = 0
    for c, (n, ops, terms) in enumerate(circuits):
        e64, g64 = qsl_amd.Qubits(n).expvalGrad(ops, terms)
        assert abs(e[c] - e64) < 1e-5, (n, e[c], e64)
        gc = g[off:off + len(ops)]
        assert np.abs(gc - g64).max() < 1e-4, (n, np.abs(gc - g64).max())
        off += len(ops)


def test_batch_vs_central_fd():
    n = 13  # tile-pass path (more than 12 qubits)
    ops, _ = W.c4_ansatz(n, 2, W.Stream(41))
    terms = W.c4_observable(n)
    b = qsl_amd.Batch([(n, ops, terms)])
    e, g = b.run()
    assert abs(e[0] - oracle_energy(n, ops, terms)) < 1e-5
    rng = np.random.default_rng(3)
    h = 1e-3
    for i in rng.choice(len(ops), 8, replace=False):
        if int(ops[i]["kind"]) == 16:
            assert g[i] == 0
            continue
        p, m = ops.copy(), ops.copy()
        p[i]["angle"] += h
        m[i]["angle"] -= h
        want = (oracle_energy(n, p, terms) - oracle_energy(n, m, terms)) / (2 * h)
        assert abs(g[i] - want) < 1e-4, (i, g[i], want)


def test_batch_graph_replays_with_new_angles():
    circuits = []
    s = W.Stream(42)
    for n in (6, 11, 14):
        ops, _ = W.c4_ansatz(n, 3, s)
        circuits.append((n, ops, W.c4_observable(n)))
    b = qsl_amd.Batch(circuits)
    e0, g0 = b.run()
    rng = np.random.default_rng(5)
    for rep in range(4):
        angles = rng.uniform(-np.pi, np.pi, len(b.ops))
        e, g = b.run(angles)
        # eager reference: a fresh batch built with those angles
        fresh = []
        off = 0
        for n, ops, terms in circuits:
            o = ops.copy()
            o["angle"] = angles[off:off + len(o)]
            off += len(o)
            fresh.append((n, o, terms))
        ef, gf = qsl_amd.Batch(fresh).run()
        assert np.abs(e - ef).max() == 0 and np.abs(g - gf).max() == 0, rep
        # and the FP64 plain path agrees
        off = 0
        for c, (n, o, terms) in enumerate(fresh):
            e64, g64 = qsl_amd.Qubits(n).expvalGrad(o, terms)
            assert abs(e[c] - e64) < 1e-5
            assert np.abs(g[off:off + len(o)] - g64).max() < 1e-4
            off += len(o)
    e1, g1 = b.run(b.ops["angle"])
    assert np.array_equal(e1, e0) and np.array_equal(g1, g0)
