"""Seeded synthetic workloads (gate lists) for the five benchmark configs.

Every config draws from one uniform stream: mt19937_64 seeded with
splitmix64(20261202 + c) for config c = 1..5, u = (x >> 11) * 2^-53
(qslhip_uniforms).  The reference's own generators are unseedable
(src/utils/random.cpp:32-45), so these streams feed both the GPU and the
oracle.  Angles: 2*pi*u on [0, 2*pi), pi*(2u-1) on [-pi, pi].
"""
import math

import numpy as np

from . import _lib as L
from ._lib import OP_DTYPE

MASTER_SEED = 20261202


class Stream:
    """Sequential reader over a config's uniform stream."""

    def __init__(self, config, block=1 << 16):
        self.seed = MASTER_SEED + config
        self.block = block
        self.buf = L.uniforms(self.seed, block)
        self.pos = 0
        self.used = 0

    def take(self, count=None):
        one = count is None
        count = 1 if one else count
        out = np.empty(count)
        got = 0
        while got < count:
            if self.pos == len(self.buf):
                self._refill(max(self.block, count - got))
            m = min(count - got, len(self.buf) - self.pos)
            out[got:got + m] = self.buf[self.pos:self.pos + m]
            self.pos += m
            got += m
        self.used += count
        return float(out[0]) if one else out

    def _refill(self, size):
        # regenerate the stream prefix up to the new end (the stream is a
        # single mt19937_64 sequence, so the prefix is identical)
        total = self.used + size
        full = L.uniforms(self.seed, total)
        self.buf = full[self.used:]
        self.pos = 0

    def below(self, n):
        return min(n - 1, int(self.take() * n))


def _op(kind, q0, q1=0, angle=0.0):
    return (kind, q0, q1, 0, angle)


def c1_layers(n=16, layers=20, stream=None):
    """Config 1: 20 layers of {H, RX, RY, RZ} per qubit (1/4 each, angle
    2*pi*u), then CNOT on a Fisher-Yates perfect matching."""
    s = stream or Stream(1)
    ops = []
    for _ in range(layers):
        for q in range(n):
            kind = [L.HADAMARD, L.ROT_X, L.ROT_Y, L.ROT_Z][s.below(4)]
            ang = 2 * math.pi * s.take() if kind != L.HADAMARD else 0.0
            ops.append(_op(kind, q, 0, ang))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = s.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for i in range(n // 2):
            ops.append(_op(L.CNOT, perm[2 * i], perm[2 * i + 1]))
    return np.array(ops, OP_DTYPE), s


def c2_random(n=30, ngates=200, stream=None):
    """Config 2: r = u; r < 0.4 -> RX/RY/RZ (1/3 each, angle 2*pi*u);
    < 0.6 -> H; < 0.8 -> CNOT; else CPhase (phase 2*pi*u).  Targets uniform
    over 0..n-1; two-qubit gates take a uniform ordered distinct pair."""
    s = stream or Stream(2)
    ops = []
    for _ in range(ngates):
        r = s.take()
        if r < 0.4:
            kind = [L.ROT_X, L.ROT_Y, L.ROT_Z][s.below(3)]
            t = s.below(n)
            ops.append(_op(kind, t, 0, 2 * math.pi * s.take()))
        elif r < 0.6:
            ops.append(_op(L.HADAMARD, s.below(n)))
        else:
            a = s.below(n)
            b = s.below(n - 1)
            b += b >= a
            if r < 0.8:
                ops.append(_op(L.CNOT, a, b))
            else:
                ops.append(_op(L.CPHASE, a, b, 2 * math.pi * s.take()))
    return np.array(ops, OP_DTYPE), s


def colex_unrank(r, k):
    """The weight-k basis string of rank r among weight-k strings in
    increasing numeric order, the order of the reference's generateLookup
    (src/utils/misc.cpp:75-90): rank(x) = sum_i C(p_i, i) over the set bits
    p_1 < ... < p_k (the combinatorial number system)."""
    x = 0
    for i in range(k, 0, -1):
        p = i - 1
        while math.comb(p + 1, i) <= r:
            p += 1
        r -= math.comb(p, i)
        x |= 1 << p
    return x


def c3_np_layers(n=32, layers=100, stream=None):
    """Config 3: layer l uses bonds (q, q+1) with q = l mod 2 (step 2), each
    npRotateX or npRotateY (1/2 each, angle 2*pi*u); then phase(q, 2*pi*u)
    for every qubit.  Returns (ops, start basis index of weight n/2)."""
    s = stream or Stream(3)
    k = n // 2
    total = math.comb(n, k)
    r0 = min(total - 1, int(s.take() * total))
    ops = []
    for layer in range(layers):
        for q in range(layer % 2, n - 1, 2):
            kind = [L.NP_ROT_X, L.NP_ROT_Y][s.below(2)]
            ops.append(_op(kind, q, q + 1, 2 * math.pi * s.take()))
        for q in range(n):
            ops.append(_op(L.PHASE, q, 0, 2 * math.pi * s.take()))
    return np.array(ops, OP_DTYPE), r0, s


def c4_ansatz(n, layers=8, stream=None):
    """Hardware-efficient ansatz of config 4: per layer RY(q) for every
    qubit, RZ(q) for every qubit (angles pi*(2u-1)), then the CNOT ladder
    0->1->...->n-1."""
    s = stream or Stream(4)
    ops = []
    for _ in range(layers):
        for q in range(n):
            ops.append(_op(L.ROT_Y, q, 0, math.pi * (2 * s.take() - 1)))
        for q in range(n):
            ops.append(_op(L.ROT_Z, q, 0, math.pi * (2 * s.take() - 1)))
        for q in range(n - 1):
            ops.append(_op(L.CNOT, q, q + 1))
    return np.array(ops, OP_DTYPE), s


def c4_observable(n):
    """H = Z0 Z1 + 0.5 * sum_i X_i as Pauli terms."""
    t = [(1.0, 0, 0b11)] + [(0.5, 1 << i, 0) for i in range(n)]
    return np.array(t, L.PAULI_DTYPE)


def c4_batch(ncirc=4096, layers=8, stream=None):
    """Config 4: ncirc circuits with n_c = 10 + floor(11 u) qubits, each the
    ansatz of c4_ansatz drawn from the same stream in sequence (one u for
    n_c, then 2 n_c angles per layer).  Vectorised: same ops as calling
    c4_ansatz circuit by circuit."""
    s = stream or Stream(4)
    circuits = []
    for _ in range(ncirc):
        n = 10 + min(10, int(11 * s.take()))
        ang = np.pi * (2 * s.take(2 * n * layers) - 1)
        per = 3 * n - 1
        ops = np.zeros(layers * per, OP_DTYPE)
        q = np.arange(n)
        for l in range(layers):
            base = l * per
            ops["kind"][base:base + n] = L.ROT_Y
            ops["q0"][base:base + n] = q
            ops["angle"][base:base + n] = ang[2 * n * l:2 * n * l + n]
            ops["kind"][base + n:base + 2 * n] = L.ROT_Z
            ops["q0"][base + n:base + 2 * n] = q
            ops["angle"][base + n:base + 2 * n] = ang[2 * n * l + n:2 * n * (l + 1)]
            ops["kind"][base + 2 * n:base + per] = L.CNOT
            ops["q0"][base + 2 * n:base + per] = q[:-1]
            ops["q1"][base + 2 * n:base + per] = q[1:]
        circuits.append((n, ops, c4_observable(n)))
    return circuits


def c5_haar_layers(n, layers=30, stream=None):
    """Config 5: per layer, RZ(a) RY(b) RZ(g) on every qubit with a, g =
    2*pi*u and b = acos(1 - 2u) (Haar on SU(2)), then controlPhase on a
    Fisher-Yates perfect matching (phase 2*pi*u)."""
    s = stream or Stream(5)
    ops = []
    for _ in range(layers):
        for q in range(n):
            a = 2 * math.pi * s.take()
            b = math.acos(1 - 2 * s.take())
            g = 2 * math.pi * s.take()
            ops.append(_op(L.ROT_Z, q, 0, g))
            ops.append(_op(L.ROT_Y, q, 0, b))
            ops.append(_op(L.ROT_Z, q, 0, a))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = s.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for i in range(n // 2):
            ops.append(_op(L.CPHASE, perm[2 * i], perm[2 * i + 1],
                           2 * math.pi * s.take()))
    return np.array(ops, OP_DTYPE), s


# Algorithmic bytes per gate (read + write of the touched amplitudes only):
# dense one-qubit and rotateZ 2*2^n*A, phase/pauliZ/CNOT 2*2^(n-1)*A,
# controlPhase/controlZ 2*2^(n-2)*A; A = bytes per amplitude.
def algorithmic_bytes(ops, n, amp_bytes):
    full = 2 * (1 << n) * amp_bytes
    tot = 0
    for k in ops["kind"]:
        k = int(k)
        if k in (L.PAULI_Z, L.PHASE, L.CNOT, L.CY, L.CROT_X, L.CROT_Y,
                 L.CROT_Z, L.CHADAMARD):
            tot += full // 2
        elif k in (L.CZ, L.CPHASE):
            tot += full // 4
        elif k in (L.SWAP, L.NP_ROT_X, L.NP_ROT_Y, L.NP_HADAMARD):
            tot += full // 2
        elif k == L.FSWAP:
            tot += full * 3 // 4
        else:
            tot += full
    return tot
