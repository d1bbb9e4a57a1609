"""ctypes view of the CPU oracle (liboracle.so) and of the reference library
built from qsl's own sources (_ref/libqslref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libqslref.so")

# Gate kinds (qsl_oracle.h)
PAULI_X, PAULI_Y, PAULI_Z, HADAMARD, PHASE, ROT_X, ROT_Y, ROT_Z = range(8)
CNOT, CY, CZ, CROT_X, CROT_Y, CROT_Z, CPHASE, CHADAMARD = range(16, 24)
SWAP, FSWAP, NP_ROT_X, NP_ROT_Y, NP_HADAMARD = range(24, 29)
GATE1 = [PAULI_X, PAULI_Y, PAULI_Z, HADAMARD, PHASE, ROT_X, ROT_Y, ROT_Z]
GATE2 = list(range(16, 29))
NP_GATE1 = [PAULI_Z, PHASE, ROT_Z]
NP_GATE2 = [CZ, CROT_Z, CPHASE, SWAP, FSWAP, NP_ROT_X, NP_ROT_Y, NP_HADAMARD]
HAS_ANGLE = {PHASE, ROT_X, ROT_Y, ROT_Z, CROT_X, CROT_Y, CROT_Z, CPHASE,
             NP_ROT_X, NP_ROT_Y}

OP_DTYPE = np.dtype([("kind", "<i4"), ("q0", "<u4"), ("q1", "<u4"),
                     ("pad", "<u4"), ("angle", "<f8")])

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_vp)


class Oracle:
    """Thin typed wrapper around liboracle.so.  Arrays are numpy complex128
    (double) or complex64 (float) state vectors; layout is the interleaved
    {re, im} of qsl::complex<Fp>."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        self.lib = L = C.CDLL(path)
        U, I, S = C.c_uint, C.c_int, C.c_size_t
        u64 = C.c_uint64
        for sfx, R in (("d", C.c_double), ("f", C.c_float)):
            def f(name, res, *args):
                fn = getattr(L, f"{name}_{sfx}")
                fn.restype = res
                fn.argtypes = list(args)
            f("orc_reset", None, _vp, U)
            f("orc_set_basis", None, _vp, U, u64)
            f("orc_gate1", None, _vp, U, I, U, R)
            f("orc_gate2", None, _vp, U, I, U, U, R)
            f("orc_apply_ops", None, _vp, U, _vp, S)
            f("orc_prob", R, _vp, U, U, U)
            f("orc_collapse", None, _vp, U, U, U, R)
            f("orc_measure", I, _vp, U, U, R)
            f("orc_postselect", R, _vp, U, U, U)
            f("orc_gen_dist", S, _vp, U, _vp, _vp)
            f("orc_draw_sample", u64, _vp, _vp, S, R)
            f("orc_measure_all", u64, _vp, U, R)
            f("orc_sample_all", None, _vp, U, _vp, S, _vp)
            f("orc_sample_all2", None, _vp, U, _vp, S, _vp)
            f("orc_sample", None, _vp, U, U, _vp, S, _vp)
            f("orc_norm", R, _vp, u64)
            f("orc_add_qubit", None, _vp, U, U)
            f("orc_collapse_out", None, _vp, U, U, U, R)
            f("orc_measure_out", I, _vp, U, U, R)
            f("orc_postselect_out", R, _vp, U, U, U)
            f("orc_normalise", R, _vp, u64)
            f("orc_fubini_study", R, _vp, _vp, u64)
            f("orc_check_state_np", I, _vp, u64, C.POINTER(C.c_uint))
            f("orc_omp_gate1", None, _vp, U, I, U, R, I)
            f("orc_omp_gate2", None, _vp, U, I, U, U, R, I)
            f("orc_omp_apply_ops", None, _vp, U, _vp, S, I)
            f("orc_omp_prob", R, _vp, U, U, U, I)
            f("orc_np_reset", None, _vp, U, U)
            f("orc_np_gate1", None, _vp, U, U, I, U, R)
            f("orc_np_gate2", None, _vp, U, U, I, U, U, R)
            f("orc_np_prob", R, _vp, U, U, U, U)
            f("orc_np_gen_dist", S, _vp, U, U, _vp, _vp)
            f("orc_np_draw_sample", u64, _vp, _vp, S, R)
            f("orc_np_sample_all", None, _vp, U, U, _vp, S, _vp)
            f("orc_npc_gate1", None, _vp, U, U, I, U, R, I)
            f("orc_npc_gate2", None, _vp, U, U, I, U, U, R, I)
            f("orc_npc_apply_ops", None, _vp, U, U, _vp, S, I)
            f("orc_npc_to_full", None, _vp, U, U, _vp)
            f("orc_npc_from_full", None, _vp, U, U, _vp)
        L.orc_next.restype = None
        L.orc_next.argtypes = [_u64p]
        L.orc_hamming_weight.restype = C.c_uint
        L.orc_hamming_weight.argtypes = [u64]
        L.orc_choose.restype = u64
        L.orc_choose.argtypes = [U, U]
        L.orc_choose_table.restype = u64
        L.orc_choose_table.argtypes = [U, U]
        L.orc_generate_lookup.restype = S
        L.orc_generate_lookup.argtypes = [U, U, _vp]
        L.orc_np_rank.restype = u64
        L.orc_np_rank.argtypes = [u64]
        L.orc_np_unrank.restype = u64
        L.orc_np_unrank.argtypes = [u64, U]

    # -- helpers -------------------------------------------------------------
    @staticmethod
    def sfx(st):
        if st.dtype == np.complex128:
            return "d"
        if st.dtype == np.complex64:
            return "f"
        raise TypeError(st.dtype)

    def _fn(self, name, st):
        return getattr(self.lib, f"{name}_{self.sfx(st)}")

    @staticmethod
    def nqubits(st):
        n = int(st.size).bit_length() - 1
        assert st.size == 1 << n
        return n

    # -- Default semantics ---------------------------------------------------
    def reset(self, st):
        self._fn("orc_reset", st)(_ptr(st), self.nqubits(st))

    def set_basis(self, st, idx):
        self._fn("orc_set_basis", st)(_ptr(st), self.nqubits(st), idx)

    def gate1(self, st, kind, t, angle=0.0):
        self._fn("orc_gate1", st)(_ptr(st), self.nqubits(st), kind, t, angle)

    def gate2(self, st, kind, a, b, angle=0.0):
        self._fn("orc_gate2", st)(_ptr(st), self.nqubits(st), kind, a, b, angle)

    def apply_ops(self, st, ops):
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        self._fn("orc_apply_ops", st)(_ptr(st), self.nqubits(st), _ptr(ops),
                                      len(ops))

    def omp_apply_ops(self, st, ops, nthreads):
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        self._fn("orc_omp_apply_ops", st)(_ptr(st), self.nqubits(st),
                                          _ptr(ops), len(ops), nthreads)

    def omp_gate1(self, st, kind, t, angle, nthreads):
        self._fn("orc_omp_gate1", st)(_ptr(st), self.nqubits(st), kind, t,
                                      angle, nthreads)

    def omp_gate2(self, st, kind, a, b, angle, nthreads):
        self._fn("orc_omp_gate2", st)(_ptr(st), self.nqubits(st), kind, a, b,
                                      angle, nthreads)

    def prob(self, st, t, o):
        return self._fn("orc_prob", st)(_ptr(st), self.nqubits(st), t, o)

    def omp_prob(self, st, t, o, nthreads):
        return self._fn("orc_omp_prob", st)(_ptr(st), self.nqubits(st), t, o,
                                            nthreads)

    def collapse(self, st, t, o, factor):
        self._fn("orc_collapse", st)(_ptr(st), self.nqubits(st), t, o, factor)

    def measure(self, st, t, u):
        return self._fn("orc_measure", st)(_ptr(st), self.nqubits(st), t, u)

    def postselect(self, st, t, o):
        return self._fn("orc_postselect", st)(_ptr(st), self.nqubits(st), t, o)

    def gen_dist(self, st):
        n = self.nqubits(st)
        idx = np.zeros((1 << n) + 1, np.uint64)
        cdf = np.zeros((1 << n) + 1, np.float64)
        m = self._fn("orc_gen_dist", st)(_ptr(st), n, _ptr(idx), _ptr(cdf))
        return idx[:m].copy(), cdf[:m].copy()

    def draw_sample(self, idx, cdf, u, single=False):
        fn = self.lib.orc_draw_sample_f if single else self.lib.orc_draw_sample_d
        idx = np.ascontiguousarray(idx, np.uint64)
        cdf = np.ascontiguousarray(cdf, np.float64)
        return fn(_ptr(idx), _ptr(cdf), len(idx), u)

    def measure_all(self, st, u):
        return self._fn("orc_measure_all", st)(_ptr(st), self.nqubits(st), u)

    def _real(self, st, u):
        rt = np.float64 if st.dtype == np.complex128 else np.float32
        return np.ascontiguousarray(u, dtype=rt)

    def sample_all(self, st, u):
        u = self._real(st, u)
        out = np.zeros(len(u), np.uint64)
        self._fn("orc_sample_all", st)(_ptr(st), self.nqubits(st), _ptr(u),
                                       len(u), _ptr(out))
        return out

    def sample_all2(self, st, u):
        u = self._real(st, u)
        counts = np.zeros(st.size, np.uint64)
        self._fn("orc_sample_all2", st)(_ptr(st), self.nqubits(st), _ptr(u),
                                        len(u), _ptr(counts))
        return counts

    def sample(self, st, t, u):
        u = self._real(st, u)
        counts = np.zeros(2, np.uint64)
        self._fn("orc_sample", st)(_ptr(st), self.nqubits(st), t, _ptr(u),
                                   len(u), _ptr(counts))
        return counts

    def norm(self, st):
        return self._fn("orc_norm", st)(_ptr(st), st.size)

    # -- Resize semantics (src/qubits/resize/): each returns the new vector --
    def add_qubit(self, st, t):
        n = self.nqubits(st)
        buf = np.zeros(2 * st.size, st.dtype)
        buf[:st.size] = st
        self._fn("orc_add_qubit", st)(_ptr(buf), n, t)
        return buf

    def collapse_out(self, st, t, o, factor):
        buf = np.ascontiguousarray(st).copy()
        self._fn("orc_collapse_out", st)(_ptr(buf), self.nqubits(st), t, o, factor)
        return buf[:st.size // 2].copy()

    def measure_out(self, st, t, u):
        buf = np.ascontiguousarray(st).copy()
        out = self._fn("orc_measure_out", st)(_ptr(buf), self.nqubits(st), t, u)
        return out, buf[:st.size // 2].copy()

    def postselect_out(self, st, t, o):
        buf = np.ascontiguousarray(st).copy()
        p = self._fn("orc_postselect_out", st)(_ptr(buf), self.nqubits(st), t, o)
        return p, buf[:st.size // 2].copy()

    def normalise(self, st):
        return self._fn("orc_normalise", st)(_ptr(st), st.size)

    def fubini_study(self, v, w):
        return self._fn("orc_fubini_study", v)(_ptr(v), _ptr(w), v.size)

    def check_state_np(self, st):
        k = C.c_uint(0)
        rc = self._fn("orc_check_state_np", st)(_ptr(st), st.size, C.byref(k))
        if rc != 0:
            raise ValueError("Input state is not number preserving.")
        return k.value

    # -- NP semantics (full vector) --------------------------------------------
    def np_reset(self, st, nones):
        self._fn("orc_np_reset", st)(_ptr(st), self.nqubits(st), nones)

    def np_gate1(self, st, nones, kind, t, angle=0.0):
        self._fn("orc_np_gate1", st)(_ptr(st), self.nqubits(st), nones, kind, t,
                                     angle)

    def np_gate2(self, st, nones, kind, a, b, angle=0.0):
        self._fn("orc_np_gate2", st)(_ptr(st), self.nqubits(st), nones, kind, a,
                                     b, angle)

    def np_prob(self, st, nones, t, o):
        return self._fn("orc_np_prob", st)(_ptr(st), self.nqubits(st), nones, t,
                                           o)

    def np_sample_all(self, st, nones, u):
        u = self._real(st, u)
        out = np.zeros(len(u), np.uint64)
        self._fn("orc_np_sample_all", st)(_ptr(st), self.nqubits(st), nones,
                                          _ptr(u), len(u), _ptr(out))
        return out

    # -- NP compact layout ------------------------------------------------------
    def npc_gate1(self, amp, n, k, kind, t, angle=0.0, nthreads=1):
        self._fn("orc_npc_gate1", amp)(_ptr(amp), n, k, kind, t, angle, nthreads)

    def npc_gate2(self, amp, n, k, kind, a, b, angle=0.0, nthreads=1):
        self._fn("orc_npc_gate2", amp)(_ptr(amp), n, k, kind, a, b, angle,
                                       nthreads)

    def npc_apply_ops(self, amp, n, k, ops, nthreads=1):
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        self._fn("orc_npc_apply_ops", amp)(_ptr(amp), n, k, _ptr(ops), len(ops),
                                           nthreads)

    def npc_to_full(self, amp, n, k):
        full = np.zeros(1 << n, amp.dtype)
        self._fn("orc_npc_to_full", amp)(_ptr(amp), n, k, _ptr(full))
        return full

    def npc_from_full(self, full, n, k):
        amp = np.zeros(self.choose(n, k), full.dtype)
        self._fn("orc_npc_from_full", full)(_ptr(full), n, k, _ptr(amp))
        return amp

    # -- integers ---------------------------------------------------------------
    def next(self, x):
        v = C.c_uint64(x)
        self.lib.orc_next(C.byref(v))
        return v.value

    def hamming_weight(self, x):
        return self.lib.orc_hamming_weight(x)

    def choose(self, n, k):
        return self.lib.orc_choose(n, k)

    def generate_lookup(self, length, ones):
        out = np.zeros(max(1, self.choose(length, ones)), np.uint64)
        m = self.lib.orc_generate_lookup(length, ones, _ptr(out))
        return out[:m]

    def rank(self, x):
        return self.lib.orc_np_rank(x)

    def unrank(self, r, k):
        return self.lib.orc_np_unrank(r, k)


class RefLib:
    """The reference library itself, compiled from qsl's sources by
    `make -C oracle ref`.  Types: 0 Default, 1 Omp, 2 NP, 3 OmpNP, 4 Resize."""

    DEFAULT, OMP, NP, OMPNP, RESIZE = 0, 1, 2, 3, 4

    def __init__(self, path=REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = L = C.CDLL(path)
        U, I = C.c_uint, C.c_int
        u64 = C.c_uint64
        L.ref_create.restype = _vp
        L.ref_create.argtypes = [I, I, U, U]
        L.ref_destroy.argtypes = [_vp]
        L.ref_set_state.restype = I
        L.ref_set_state.argtypes = [_vp, _vp]
        L.ref_get_state.argtypes = [_vp, _vp]
        L.ref_gate1.restype = I
        L.ref_gate1.argtypes = [_vp, I, U, C.c_double]
        L.ref_gate2.restype = I
        L.ref_gate2.argtypes = [_vp, I, U, U, C.c_double]
        L.ref_prob.restype = C.c_double
        L.ref_prob.argtypes = [_vp, U, U]
        L.ref_postselect.restype = C.c_double
        L.ref_postselect.argtypes = [_vp, U, U]
        L.ref_measure.restype = I
        L.ref_measure.argtypes = [_vp, U]
        L.ref_measure_all.restype = u64
        L.ref_measure_all.argtypes = [_vp]
        L.ref_sample_all.argtypes = [_vp, u64, _vp]
        L.ref_sample.argtypes = [_vp, U, u64, _vp]
        L.ref_sample_all2.restype = I
        L.ref_sample_all2.argtypes = [_vp, u64, _vp]
        L.ref_nqubits.restype = U
        L.ref_nqubits.argtypes = [_vp]
        L.ref_set_state_n.restype = I
        L.ref_set_state_n.argtypes = [_vp, _vp, U]
        L.ref_add_qubit.restype = I
        L.ref_add_qubit.argtypes = [_vp, U]
        L.ref_measure_out.restype = I
        L.ref_measure_out.argtypes = [_vp, U]
        L.ref_postselect_out.restype = C.c_double
        L.ref_postselect_out.argtypes = [_vp, U, U]
        L.ref_reallocate.restype = I
        L.ref_reallocate.argtypes = [_vp]
        L.ref_choose.restype = U
        L.ref_choose.argtypes = [U, U]
        L.ref_hamming_weight.restype = U
        L.ref_hamming_weight.argtypes = [u64]
        L.ref_next.restype = u64
        L.ref_next.argtypes = [u64]
        L.ref_generate_lookup.restype = u64
        L.ref_generate_lookup.argtypes = [U, U, _vp, u64]


class RefSim:
    """One qsl::Qubits<Type, Fp> object of the reference library."""

    def __init__(self, ref, typ, dtype, n, extra=0):
        self.ref, self.n, self.dtype = ref, n, np.dtype(dtype)
        prec = 8 if self.dtype == np.complex128 else 4
        self.h = ref.lib.ref_create(typ, prec, n, extra)
        if not self.h:
            raise ValueError("reference constructor threw")

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.lib.ref_destroy(self.h)
            self.h = None

    def set_state(self, st):
        st = np.ascontiguousarray(st, self.dtype)
        if self.ref.lib.ref_set_state(self.h, _ptr(st)) != 0:
            raise ValueError("reference setState threw")

    def get_state(self):
        out = np.zeros(1 << self.n, self.dtype)
        self.ref.lib.ref_get_state(self.h, _ptr(out))
        return out

    def gate1(self, kind, t, angle=0.0):
        if self.ref.lib.ref_gate1(self.h, kind, t, angle) != 0:
            raise ValueError(f"gate {kind} not available")

    def gate2(self, kind, a, b, angle=0.0):
        if self.ref.lib.ref_gate2(self.h, kind, a, b, angle) != 0:
            raise ValueError(f"gate {kind} not available")

    def prob(self, t, o):
        return self.ref.lib.ref_prob(self.h, t, o)

    def postselect(self, t, o):
        return self.ref.lib.ref_postselect(self.h, t, o)

    def measure(self, t):
        return self.ref.lib.ref_measure(self.h, t)

    def measure_all(self):
        return self.ref.lib.ref_measure_all(self.h)

    def sample_all(self, nshots):
        dense = np.zeros(1 << self.nqubits(), np.uint64)
        self.ref.lib.ref_sample_all(self.h, nshots, _ptr(dense))
        return dense

    def sample(self, t, nshots):
        counts = np.zeros(2, np.uint64)
        self.ref.lib.ref_sample(self.h, t, nshots, _ptr(counts))
        return counts

    def sample_all2(self, nshots):
        dense = np.zeros(1 << self.nqubits(), np.uint64)
        if self.ref.lib.ref_sample_all2(self.h, nshots, _ptr(dense)) != 0:
            raise ValueError("sampleAll2 not available")
        return dense

    def nqubits(self):
        return self.ref.lib.ref_nqubits(self.h)

    # ---- Resize backend (src/qubits/resize/) ----
    def set_state_any(self, st):
        st = np.ascontiguousarray(st, self.dtype)
        n = int(st.size).bit_length() - 1
        if self.ref.lib.ref_set_state_n(self.h, _ptr(st), n) != 0:
            raise ValueError("reference setState threw or is not a Resize object")

    def get_state_any(self):
        out = np.zeros(1 << self.nqubits(), self.dtype)
        self.ref.lib.ref_get_state(self.h, _ptr(out))
        return out

    def add_qubit(self, t):
        if self.ref.lib.ref_add_qubit(self.h, t) != 0:
            raise ValueError("not a Resize object")

    def measure_out(self, t):
        return self.ref.lib.ref_measure_out(self.h, t)

    def postselect_out(self, t, o):
        return self.ref.lib.ref_postselect_out(self.h, t, o)

    def reallocate(self):
        if self.ref.lib.ref_reallocate(self.h) != 0:
            raise ValueError("not a Resize object")


def near_cdf_boundary(cdf, u, eps):
    """Which uniforms lie within eps of a value of the CDF `cdf` (sorted):
    the draws whose outcome may differ between two summation orders of the
    same probabilities.  Sampling parity is exact outside them."""
    cdf = np.asarray(cdf, np.float64)
    u = np.asarray(u, np.float64)
    j = np.searchsorted(cdf, u)
    lo = np.abs(u - cdf[np.clip(j - 1, 0, len(cdf) - 1)])
    hi = np.abs(u - cdf[np.clip(j, 0, len(cdf) - 1)])
    return np.minimum(lo, hi) < eps


def random_state(rng, n, dtype=np.complex128):
    """The recipe of makeRandomState (src/utils/random.cpp:48-68): re and im
    uniform on (-1, 1), then normalised; drawn from a seeded numpy rng."""
    rt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    re = rng.uniform(-1, 1, 1 << n).astype(rt)
    im = rng.uniform(-1, 1, 1 << n).astype(rt)
    st = (re + 1j * im).astype(dtype)
    Oracle_singleton().normalise(st)
    return st


def random_np_state(rng, n, k, dtype=np.complex128):
    """makeRandomNPState (src/utils/random.cpp:85-108): random amplitudes on
    the weight-k strings only, then normalised."""
    o = Oracle_singleton()
    rt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    st = np.zeros(1 << n, dtype)
    lst = o.generate_lookup(n, k).astype(np.int64)
    re = rng.uniform(-1, 1, len(lst)).astype(rt)
    im = rng.uniform(-1, 1, len(lst)).astype(rt)
    st[lst] = (re + 1j * im).astype(dtype)
    o.normalise(st)
    return st


_ORACLE = None


def Oracle_singleton():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
