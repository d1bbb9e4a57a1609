"""Python mirror of qsl::Qubits<Type::Hip, Fp> / Qubits<Type::HipNP, Fp>.

Same member names, argument meaning and error behaviour as the reference's
class (include/qsl/qubits/default.hpp, include/qsl/qubits/np.hpp), over the
C ABI of libqslhip.so.  Tests and bench.py drive the backend through it;
C++ callers use include/qsl/qubits/hip.hpp instead.

Where the reference draws a uniform from its unseedable RNG
(src/utils/random.cpp:32-45), the methods here take the uniform(s) as an
extra argument, or draw them from a seeded mt19937_64 stream owned by the
object (seed argument of the constructor).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, ptr

_vp = C.c_void_p


class Qubits:
    """Full state vector on one MI355X: 2^n amplitudes of complex64/128."""

    np_backend = False

    def __init__(self, nqubits_or_state, dtype=np.complex128, seed=0):
        self._h = None
        lib = L.lib()
        if isinstance(nqubits_or_state, np.ndarray):
            st = nqubits_or_state
            dtype = st.dtype
            n = int(st.size).bit_length() - 1
            if st.size != 1 << n:  # checkStateSize, src/utils/quantum.cpp:108-131
                raise L.SizeError(L.ESIZE, "State vector size is not a power of two")
        else:
            n, st = int(nqubits_or_state), None
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.complex64, np.complex128):
            raise TypeError("dtype must be complex64 or complex128")
        self.prec = 8 if self.dtype == np.complex128 else 4
        self.name = "Qub<hip,%s>" % ("double" if self.prec == 8 else "float")
        h = _vp()
        check(self._create(lib, n, h))
        self._h = h
        self.nqubits = n
        self._seed = seed
        self._ustream = None
        if st is not None:
            self.setState(st)

    def _create(self, lib, n, h):
        return lib.qslhip_create(n, self.prec, C.byref(h))

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib().qslhip_destroy(self._h)
            self._h = None

    # -- copies ------------------------------------------------------------------
    def copy(self):
        """Copy constructor (default.hpp:137): device-to-device clone."""
        new = object.__new__(type(self))
        new.__dict__.update(self.__dict__)
        h = _vp()
        check(L.lib().qslhip_clone(self._h, C.byref(h)))
        new._h = h
        return new

    def assign(self, other):
        """operator= (default.cpp:94-105); SizeError when nqubits differ."""
        check(L.lib().qslhip_assign(self._h, other._h))

    # -- uniforms ------------------------------------------------------------------
    def _draw(self, count):
        if self._ustream is None:
            self._ustream = (L.uniforms(self._seed, 1 << 20), 0)
        buf, pos = self._ustream
        if pos + count > len(buf):
            self._seed += 1
            buf, pos = L.uniforms(self._seed, max(1 << 20, count)), 0
        self._ustream = (buf, pos + count)
        return buf[pos:pos + count]

    # -- state -----------------------------------------------------------------------
    def getNumQubits(self):
        return self.nqubits

    def reset(self):
        check(L.lib().qslhip_reset(self._h))

    def getState(self):
        out = np.empty(1 << self.nqubits, self.dtype)
        check(L.lib().qslhip_get_state(self._h, ptr(out), out.size))
        return out

    def setState(self, st):
        st = np.ascontiguousarray(st, dtype=self.dtype)
        check(L.lib().qslhip_set_state(self._h, ptr(st), st.size))

    def setBasisState(self, index):
        check(L.lib().qslhip_set_basis(self._h, int(index)))

    def getAmps(self, first=0, count=None):
        namps = self.numAmps()
        count = namps - first if count is None else count
        out = np.empty(count, self.dtype)
        check(L.lib().qslhip_get_amps(self._h, ptr(out), first, count))
        return out

    def setAmps(self, amps, first=0):
        amps = np.ascontiguousarray(amps, dtype=self.dtype)
        check(L.lib().qslhip_set_amps(self._h, ptr(amps), first, amps.size))

    def numAmps(self):
        v = C.c_uint64(0)
        check(L.lib().qslhip_info(self._h, None, None, None, C.byref(v)))
        return v.value

    def print(self, file=None):
        """print() (default.cpp:117-122): header line then one "(re, im)" per
        amplitude in the default ostream format of Fp."""
        import sys
        f = file or sys.stdout
        f.write(self._header())
        for a in self.getState():
            f.write("(%s, %s)\n" % (_fmt(a.real), _fmt(a.imag)))
        f.write("\n")

    def _header(self):
        return "Number of qubits = %d\n" % self.nqubits

    def sync(self):
        check(L.lib().qslhip_sync(self._h))

    def set_stream(self, stream_ptr):
        check(L.lib().qslhip_set_stream(self._h, stream_ptr))

    # -- gates -----------------------------------------------------------------------
    def _g1(self, kind, t, angle=0.0):
        check(L.lib().qslhip_gate1(self._h, kind, t, float(angle)))

    def _g2(self, kind, a, b, angle=0.0):
        check(L.lib().qslhip_gate2(self._h, kind, a, b, float(angle)))

    def pauliX(self, t): self._g1(L.PAULI_X, t)
    def pauliY(self, t): self._g1(L.PAULI_Y, t)
    def pauliZ(self, t): self._g1(L.PAULI_Z, t)
    def hadamard(self, t): self._g1(L.HADAMARD, t)
    def phase(self, t, angle): self._g1(L.PHASE, t, angle)
    def rotateX(self, t, angle): self._g1(L.ROT_X, t, angle)
    def rotateY(self, t, angle): self._g1(L.ROT_Y, t, angle)
    def rotateZ(self, t, angle): self._g1(L.ROT_Z, t, angle)
    def controlNot(self, c, t): self._g2(L.CNOT, c, t)
    def controlY(self, c, t): self._g2(L.CY, c, t)
    def controlZ(self, c, t): self._g2(L.CZ, c, t)
    def controlRotateX(self, c, t, angle): self._g2(L.CROT_X, c, t, angle)
    def controlRotateY(self, c, t, angle): self._g2(L.CROT_Y, c, t, angle)
    def controlRotateZ(self, c, t, angle): self._g2(L.CROT_Z, c, t, angle)
    def controlPhase(self, c, t, angle): self._g2(L.CPHASE, c, t, angle)
    def controlHadamard(self, c, t): self._g2(L.CHADAMARD, c, t)
    def swap(self, a, b): self._g2(L.SWAP, a, b)
    def fswap(self, a, b): self._g2(L.FSWAP, a, b)
    def npRotateX(self, a, b, angle): self._g2(L.NP_ROT_X, a, b, angle)
    def npRotateY(self, a, b, angle): self._g2(L.NP_ROT_Y, a, b, angle)
    def npHadamard(self, a, b): self._g2(L.NP_HADAMARD, a, b)

    def apply_ops(self, ops, mode=L.MODE_FUSED):
        ops = np.ascontiguousarray(ops, dtype=L.OP_DTYPE)
        check(L.lib().qslhip_apply_ops(self._h, ptr(ops), len(ops), mode))

    # -- observables and gradients (no reference counterpart) -------------------------
    def expval(self, terms):
        """<psi|H|psi> for H = sum of Pauli terms (PAULI_DTYPE records)."""
        terms = np.ascontiguousarray(terms, dtype=L.PAULI_DTYPE)
        v = C.c_double()
        check(L.lib().qslhip_expval(self._h, ptr(terms), len(terms), C.byref(v)))
        return v.value

    def expvalGrad(self, ops, terms):
        """Apply ops (the object keeps the result) and return
        (E, dE/dangle per op) by the adjoint reverse sweep."""
        ops = np.ascontiguousarray(ops, dtype=L.OP_DTYPE)
        terms = np.ascontiguousarray(terms, dtype=L.PAULI_DTYPE)
        e = C.c_double()
        g = np.zeros(len(ops), np.float64)
        check(L.lib().qslhip_expval_grad(self._h, ptr(ops), len(ops), ptr(terms),
                                         len(terms), C.byref(e), ptr(g)))
        return e.value, g

    # -- measurement -------------------------------------------------------------------
    def prob(self, t, outcome):
        p = C.c_double()
        check(L.lib().qslhip_prob(self._h, t, outcome, C.byref(p)))
        return p.value

    def norm(self):
        p = C.c_double()
        check(L.lib().qslhip_norm(self._h, C.byref(p)))
        return p.value

    def measure(self, t, u=None):
        u = float(self._draw(1)[0]) if u is None else float(u)
        o = C.c_int()
        check(L.lib().qslhip_measure(self._h, t, u, C.byref(o)))
        return o.value

    def postselect(self, t, outcome):
        p = C.c_double()
        check(L.lib().qslhip_postselect(self._h, t, outcome, C.byref(p)))
        return p.value

    def measureAll(self, u=None):
        u = float(self._draw(1)[0]) if u is None else float(u)
        o = C.c_uint64()
        check(L.lib().qslhip_measure_all(self._h, u, C.byref(o)))
        return o.value

    def sampleAllOutcomes(self, u):
        """One outcome per uniform (the draws behind sampleAll)."""
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(u.size, np.uint64)
        check(L.lib().qslhip_sample_all(self._h, ptr(u), u.size, ptr(out)))
        return out

    def sampleAll(self, nsamples, u=None):
        """sampleAll(N) (measure.cpp:220-234): {outcome: count}."""
        u = self._draw(nsamples) if u is None else u
        out = self.sampleAllOutcomes(u)
        keys, counts = np.unique(out, return_counts=True)
        return {int(k): int(c) for k, c in zip(keys, counts)}

    def sample(self, t, nsamples, u=None):
        """sample(targ, N) (measure.cpp:191-216): [count of 0, count of 1]."""
        u = self._draw(nsamples) if u is None else np.ascontiguousarray(u, np.float64)
        counts = np.zeros(2, np.uint64)
        check(L.lib().qslhip_sample(self._h, t, ptr(u), len(u), ptr(counts)))
        return [int(counts[0]), int(counts[1])]


class QubitsResize(Qubits):
    """Resizeable backend (include/qsl/qubits/resize.hpp): qubits are added
    and measured or postselected out on the device; the allocation doubles
    when a qubit no longer fits and shrinks only in reallocateState."""

    def __init__(self, nqubits_or_state, dtype=np.complex128, seed=0):
        super().__init__(nqubits_or_state, dtype=dtype, seed=seed)
        self.name = "Qub<hipresize,%s>" % ("double" if self.prec == 8 else "float")

    def _create(self, lib, n, h):
        return lib.qslhip_create_resize(n, self.prec, C.byref(h))

    def _refresh(self):
        n = C.c_uint()
        check(L.lib().qslhip_info(self._h, C.byref(n), None, None, None))
        self.nqubits = n.value

    def setState(self, st):
        """resize.cpp:70-81: any power-of-two size; the qubit count follows."""
        st = np.ascontiguousarray(st, dtype=self.dtype)
        if st.size < 2 or st.size & (st.size - 1):
            raise L.SizeError(L.ESIZE, "State vector size is not a power of two")
        check(L.lib().qslhip_resize_set_state(self._h, ptr(st), st.size))
        self._refresh()

    def addQubit(self, t):
        check(L.lib().qslhip_add_qubit(self._h, t))
        self._refresh()

    def appendQubit(self):
        self.addQubit(self.nqubits)

    def reallocateState(self):
        check(L.lib().qslhip_reallocate(self._h))

    def collapseOut(self, t, outcome, factor):
        check(L.lib().qslhip_collapse_out(self._h, t, outcome, factor))
        self._refresh()

    def measureOut(self, t, u=None):
        if u is None:
            u = float(self._draw(1)[0])
        o = C.c_int()
        check(L.lib().qslhip_measure_out(self._h, t, u, C.byref(o)))
        self._refresh()
        return o.value

    def postselectOut(self, t, outcome):
        p = C.c_double()
        check(L.lib().qslhip_postselect_out(self._h, t, outcome, C.byref(p)))
        self._refresh()
        return p.value

    def capacity(self):
        """Amplitudes the device allocation holds (>= 2^nqubits)."""
        v = C.c_uint64()
        check(L.lib().qslhip_capacity(self._h, C.byref(v)))
        return v.value


class QubitsNP(Qubits):
    """Number-preserving backend: only the C(n, nones) amplitudes of weight
    nones live on the device (compact colex-rank layout)."""

    np_backend = True

    def __init__(self, nqubits_or_state, nones=1, dtype=np.complex128, seed=0):
        self._nones_init = nones
        super().__init__(nqubits_or_state, dtype=dtype, seed=seed)
        self.name = "Qub<hipnp,%s>" % ("double" if self.prec == 8 else "float")

    def _create(self, lib, n, h):
        return lib.qslhip_create_np(n, self._nones_init, self.prec, C.byref(h))

    def getNumOnes(self):
        k = C.c_uint()
        check(L.lib().qslhip_info(self._h, None, C.byref(k), None, None))
        return k.value

    def setNumOnes(self, nones):
        check(L.lib().qslhip_set_num_ones(self._h, nones))

    def _header(self):
        return "Number of qubits = %d\nNumber of ones = %d\n" % (
            self.nqubits, self.getNumOnes())


def _fmt(x):
    """std::ostream default formatting of a float/double (6 significant
    digits, %g-like, no trailing zeros)."""
    s = "%g" % float(x)
    return s


class Batch:
    """Many variational circuits, one persistent FP32 kernel in a hipGraph
    (qslhip_batch_*): circuits = list of (nqubits, ops, terms)."""

    def __init__(self, circuits):
        nq = np.array([c[0] for c in circuits], np.uint32)
        ops = [np.ascontiguousarray(c[1], dtype=L.OP_DTYPE) for c in circuits]
        terms = [np.ascontiguousarray(c[2], dtype=L.PAULI_DTYPE) for c in circuits]
        self.op_off = np.zeros(len(circuits) + 1, np.uint64)
        self.op_off[1:] = np.cumsum([len(o) for o in ops])
        t_off = np.zeros(len(circuits) + 1, np.uint64)
        t_off[1:] = np.cumsum([len(t) for t in terms])
        self.ops = np.concatenate(ops) if ops else np.zeros(0, L.OP_DTYPE)
        allt = np.concatenate(terms) if terms else np.zeros(0, L.PAULI_DTYPE)
        self.ncirc = len(circuits)
        self._b = None
        b = _vp()
        check(L.lib().qslhip_batch_create(4, self.ncirc, ptr(nq), ptr(self.op_off),
                                          ptr(self.ops), ptr(t_off), ptr(allt),
                                          C.byref(b)))
        self._b = b

    def __del__(self):
        if getattr(self, "_b", None):
            L.lib().qslhip_batch_destroy(self._b)
            self._b = None

    def run(self, angles=None):
        """Energies per circuit and gradients per op (split per circuit)."""
        e = np.zeros(self.ncirc, np.float64)
        g = np.zeros(len(self.ops), np.float64)
        a = None if angles is None else np.ascontiguousarray(angles, np.float64)
        if a is not None and (a.ndim != 1 or len(a) != len(self.ops)):
            raise ValueError("angles must be a 1-D array of %d values (one per op), got shape %s"
                             % (len(self.ops), a.shape))
        check(L.lib().qslhip_batch_run(self._b, None if a is None else ptr(a), ptr(e),
                                       ptr(g)))
        return e, g

    def stream(self):
        s = _vp()
        check(L.lib().qslhip_batch_stream(self._b, C.byref(s)))
        return s
