"""Parity at the configured (full) sizes.

* Config 2, 30 qubits FP64 (16 GiB): a dense random state, then 40 gates
  of the config-2 mix on the qubits S = {0..5} u {22..29}, so both the
  lowest (register/LDS tile) and the highest targets are hit and amplitude
  byte offsets run past 4 GiB.  Gates on S never mix amplitudes that
  differ outside S, so for each fixed value m of qubits 6..21 the 2^14
  amplitudes with those bits evolve as an independent 14-qubit state under
  the same gates.  The oracle runs 64 such states (m drawn at random, all
  8 high-bit blocks of each) and the device result is compared there:
  per-gate and program-order passes bit for bit (the latter applied twice:
  interpreted, then compiled), reordered passes within 1e-12.
* Config 3, 32 qubits weight 16 FP64 (9.62 GB compact): a dense random
  compact state, two layers of the config-3 circuit (31 bonds, 64 phases),
  against the oracle's compact restatement, bit for bit over all
  601,080,390 amplitudes.
* Config 4: a batch holding 18-, 19- and 20-qubit ansatz circuits (and
  smaller ones) against the FP64 per-gate adjoint, max gradient error
  <= 1e-4.
* Config 5, 34 qubits FP32 (128 GiB) on the dense matrix-core path: two
  Haar layers, then their inverse; the state must return to |0...0>: every
  prob(q, 0) within 1e-4 of 1 and the norm within 1e-5 of 1.
"""
import gc
import math

import numpy as np
import pytest

import qsl_amd
from qsl_amd import workloads as W
from oracle.pyoracle import Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
O = Oracle()


def _free():
    gc.collect()


# ---- config 2 at 30 qubits -------------------------------------------------------

N2 = 30
LOW = 6            # qubits 0..5
HIGH0 = 22         # qubits 22..29
CH = 22            # amplitudes per chunk: one value of the 8 high bits
NMID = 64          # sampled values of the 16 middle bits


def _c2_reduced_ops():
    """40 config-2 gates drawn over 14 slots; slot i is qubit S[i]."""
    red, _ = W.c2_random(n=14, ngates=40)
    slots = np.array(list(range(LOW)) + list(range(HIGH0, N2)), np.uint32)
    full = red.copy()
    full["q0"] = slots[red["q0"]]
    two = red["kind"] >= 16
    full["q1"] = np.where(two, slots[red["q1"]], 0)
    return red, full


def _c2_chunk(c):
    rng = np.random.default_rng([20261202, 2, c])
    return (rng.random(2 << CH) * 2 - 1).view(np.complex128)


def _gather(q, mids):
    """Amplitudes of the sampled middle values: [m, high, low] -> reduced
    states [m, high * 64 + low] (reduced qubit j < 6 is qubit j, reduced
    qubit 6 + h is qubit 22 + h)."""
    out = np.empty((len(mids), 1 << (N2 - HIGH0), 1 << LOW), np.complex128)
    for c in range(1 << (N2 - HIGH0)):
        for i, m in enumerate(mids):
            out[i, c] = q.getAmps((c << CH) | (int(m) << LOW), 1 << LOW)
    return out.reshape(len(mids), -1)


def test_c2_30q_fullsize_parity():
    red_ops, ops = _c2_reduced_ops()
    mids = np.sort(np.random.default_rng(7).choice(1 << (HIGH0 - LOW), NMID, replace=False))
    q = qsl_amd.Qubits(N2)
    init = np.empty((NMID, 1 << (N2 - HIGH0), 1 << LOW), np.complex128)
    for c in range(1 << (N2 - HIGH0)):
        chunk = _c2_chunk(c)
        q.setAmps(chunk, first=c << CH)
        init[:, c] = chunk.reshape(1 << (HIGH0 - LOW), 1 << LOW)[mids]
    init = init.reshape(NMID, -1)
    once = init.copy()
    twice = init.copy()
    for i in range(NMID):
        O.apply_ops(once[i], red_ops)
        O.apply_ops(twice[i], red_ops)
        O.apply_ops(twice[i], red_ops)
    assert np.abs(once - init).max() > 0.1  # the circuit changes the sampled states

    per = q.copy()
    per.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    got = _gather(per, mids)
    assert np.array_equal(got.view(np.uint8), once.view(np.uint8)), "per-gate"
    del per
    _free()

    fused = q.copy()
    fused.apply_ops(ops, mode=qsl_amd.MODE_FUSED)   # interpreted
    fused.apply_ops(ops, mode=qsl_amd.MODE_FUSED)   # compiled on repeat
    got = _gather(fused, mids)
    assert np.array_equal(got.view(np.uint8), twice.view(np.uint8)), "program-order passes"
    del fused
    _free()

    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    got = _gather(q, mids)
    err = np.abs(got - once).max()
    assert err <= 1e-12, f"reordered passes: max |diff| {err}"
    # prob over the whole 2^30 vector agrees with its two halves
    p0, p1 = q.prob(29, 0), q.prob(29, 1)
    assert abs((p0 + p1) - q.norm() ** 2) <= 1e-9 * (p0 + p1)
    del q
    _free()


# ---- config 3 at 32 qubits, weight 16 -------------------------------------------

def test_c3_32q_w16_fullsize_bit_exact():
    n, k = 32, 16
    total = math.comb(n, k)
    ops, _, _ = W.c3_np_layers(n, 2)
    assert len(ops) == 31 + 64
    q = qsl_amd.QubitsNP(n, nones=k)
    assert q.numAmps() == total == 601080390
    amp = np.empty(total, np.complex128)
    step = 1 << 25
    for f in range(0, total, step):
        m = min(step, total - f)
        rng = np.random.default_rng([20261202, 3, f])
        amp[f:f + m] = (rng.random(2 * m) * 2 - 1).view(np.complex128)
        q.setAmps(amp[f:f + m], first=f)
    q.apply_ops(ops)
    O.npc_apply_ops(amp, n, k, ops, nthreads=16)
    for f in range(0, total, step):
        m = min(step, total - f)
        got = q.getAmps(f, m)
        same = np.array_equal(got.view(np.uint8), amp[f:f + m].view(np.uint8))
        assert same, f"compact amplitudes differ in [{f}, {f + m})"
    del q, amp
    _free()


# ---- config 4 with its largest circuits ------------------------------------------

def test_c4_batch_with_20q_circuits_vs_fp64_plain():
    s = W.Stream(4)
    circuits = []
    for n in (20, 19, 18, 20, 16, 13, 12, 10):
        ops, s = W.c4_ansatz(n, stream=s)
        circuits.append((n, ops, W.c4_observable(n)))
    b = qsl_amd.Batch(circuits)
    e, g = b.run()
    for c, (n, ops, terms) in enumerate(circuits):
        q = qsl_amd.Qubits(n)
        e64, g64 = q.expvalGrad(ops, terms)
        lo, hi = int(b.op_off[c]), int(b.op_off[c + 1])
        gerr = np.abs(g[lo:hi] - g64).max()
        assert abs(e[c] - e64) <= 1e-5, (n, e[c], e64)
        assert gerr <= 1e-4, (n, gerr)
    del b
    _free()


# ---- config 5 at 34 qubits on the dense matrix-core path ------------------------

def _inverse(ops):
    inv = ops[::-1].copy()
    inv["angle"] = -inv["angle"]  # RZ, RY and controlPhase: R(-t) = R(t)^-1
    return inv


def test_c5_34q_dense_mfma_round_trip():
    n = 34
    ops, _ = W.c5_haar_layers(n, 2)
    q = qsl_amd.Qubits(n, dtype=np.complex64)
    q.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
    # the circuit spreads the state: qubit 0 is no longer |0>
    assert q.prob(0, 0) < 0.99
    q.apply_ops(_inverse(ops), mode=qsl_amd.MODE_DENSE)
    nrm = q.norm()
    assert abs(nrm - 1) <= 1e-5, nrm
    worst = min(q.prob(t, 0) for t in range(n))
    assert worst >= 1 - 1e-4, worst
    del q
    _free()
