"""Loader for libqslhip.so (the HIP backend's C ABI, include/qslhip.h).

The library is built in-tree by __graft_entry__.build() (make -C
qsl_amd/csrc).  There is no fallback: if the library or a HIP device is
missing, calls fail loudly with QslHipError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqslhip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "qslhip.h")

# status codes
OK, EINVAL, ESIZE, ENOTNP, ENOMEM, ENODEV, EDEVICE, EKIND = range(8)

# gate kinds
PAULI_X, PAULI_Y, PAULI_Z, HADAMARD, PHASE, ROT_X, ROT_Y, ROT_Z = range(8)
CNOT, CY, CZ, CROT_X, CROT_Y, CROT_Z, CPHASE, CHADAMARD = range(16, 24)
SWAP, FSWAP, NP_ROT_X, NP_ROT_Y, NP_HADAMARD = range(24, 29)

MODE_PER_GATE, MODE_FUSED, MODE_DENSE, MODE_REORDER = 0, 1, 2, 3

OP_DTYPE = np.dtype([("kind", "<i4"), ("q0", "<u4"), ("q1", "<u4"),
                     ("pad", "<u4"), ("angle", "<f8")])


class QslHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class SizeError(QslHipError, ValueError):
    """The reference's std::logic_error for size mismatches."""


class NotNPError(QslHipError, ValueError):
    """The reference's std::logic_error for nones / non-NP states."""


_lib = None

_vp = C.c_void_p
_u64 = C.c_uint64
_U = C.c_uint
_I = C.c_int
_D = C.c_double

SIGNATURES = {
    "qslhip_status_string": (C.c_char_p, [_I]),
    "qslhip_last_error": (C.c_char_p, []),
    "qslhip_device_count": (_I, [C.POINTER(_I)]),
    "qslhip_create": (_I, [_U, _I, C.POINTER(_vp)]),
    "qslhip_create_np": (_I, [_U, _U, _I, C.POINTER(_vp)]),
    "qslhip_destroy": (_I, [_vp]),
    "qslhip_clone": (_I, [_vp, C.POINTER(_vp)]),
    "qslhip_assign": (_I, [_vp, _vp]),
    "qslhip_info": (_I, [_vp, C.POINTER(_U), C.POINTER(_U), C.POINTER(_I),
                         C.POINTER(_u64)]),
    "qslhip_set_stream": (_I, [_vp, _vp]),
    "qslhip_get_stream": (_I, [_vp, C.POINTER(_vp)]),
    "qslhip_sync": (_I, [_vp]),
    "qslhip_plan": (_I, [_U, _I, _vp, _u64, C.POINTER(_u64), _vp, _vp, _u64]),
    "qslhip_plan_mode": (_I, [_U, _I, _vp, _u64, _I, C.POINTER(_u64), _vp, _vp, _u64, _vp]),
    "qslhip_jit_source": (_I, [_U, _I, _vp, _u64, _I, _u64, C.c_char_p, C.c_size_t,
                               C.POINTER(_u64)]),
    "qslhip_jit_check": (_I, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    "qslhip_jit_wait": (_I, []),
    "qslhip_reset": (_I, [_vp]),
    "qslhip_set_state": (_I, [_vp, _vp, _u64]),
    "qslhip_get_state": (_I, [_vp, _vp, _u64]),
    "qslhip_set_basis": (_I, [_vp, _u64]),
    "qslhip_set_num_ones": (_I, [_vp, _U]),
    "qslhip_get_amps": (_I, [_vp, _vp, _u64, _u64]),
    "qslhip_set_amps": (_I, [_vp, _vp, _u64, _u64]),
    "qslhip_gate1": (_I, [_vp, _I, _U, _D]),
    "qslhip_gate2": (_I, [_vp, _I, _U, _U, _D]),
    "qslhip_apply_ops": (_I, [_vp, _vp, _u64, _I]),
    "qslhip_prob": (_I, [_vp, _U, _U, C.POINTER(_D)]),
    "qslhip_collapse": (_I, [_vp, _U, _U, _D]),
    "qslhip_measure": (_I, [_vp, _U, _D, C.POINTER(_I)]),
    "qslhip_postselect": (_I, [_vp, _U, _U, C.POINTER(_D)]),
    "qslhip_measure_all": (_I, [_vp, _D, C.POINTER(_u64)]),
    "qslhip_sample_all": (_I, [_vp, _vp, _u64, _vp]),
    "qslhip_sample_all_noalloc": (_I, [_vp, _vp, _u64, _vp]),
    "qslhip_create_resize": (_I, [_U, _I, C.POINTER(_vp)]),
    "qslhip_resize_set_state": (_I, [_vp, _vp, _u64]),
    "qslhip_add_qubit": (_I, [_vp, _U]),
    "qslhip_collapse_out": (_I, [_vp, _U, _U, _D]),
    "qslhip_measure_out": (_I, [_vp, _U, _D, C.POINTER(_I)]),
    "qslhip_postselect_out": (_I, [_vp, _U, _U, C.POINTER(_D)]),
    "qslhip_reallocate": (_I, [_vp]),
    "qslhip_capacity": (_I, [_vp, C.POINTER(_u64)]),
    "qslhip_sample": (_I, [_vp, _U, _vp, _u64, _vp]),
    "qslhip_sample_all2": (_I, [_vp, _vp, _u64, _vp, _u64]),
    "qslhip_norm": (_I, [_vp, C.POINTER(_D)]),
    "qslhip_uniforms": (_I, [_u64, _u64, _vp]),
    "qslhip_expval": (_I, [_vp, _vp, _u64, C.POINTER(_D)]),
    "qslhip_expval_grad": (_I, [_vp, _vp, _u64, _vp, _u64, C.POINTER(_D), _vp]),
    "qslhip_batch_create": (_I, [_I, C.c_uint32, _vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "qslhip_batch_run": (_I, [_vp, _vp, _vp, _vp]),
    "qslhip_batch_stream": (_I, [_vp, C.POINTER(_vp)]),
    "qslhip_batch_destroy": (_I, [_vp]),
    "qslhip_tuning_get": (_I, [C.c_char_p, C.POINTER(_I)]),
    "qslhip_tuning_set": (_I, [C.c_char_p, _I]),
    "qslhip_autotune": (_I, [_I, _U, C.c_char_p, C.c_size_t]),
    "qslhip_plan_dense": (_I, [_U, _vp, _u64, _I, C.POINTER(_u64), C.POINTER(_u64),
                               C.POINTER(_u64), _vp, _vp, _vp, _u64, _u64, _u64]),
}

PAULI_DTYPE = np.dtype([("coef", "<f8"), ("xmask", "<u8"), ("zmask", "<u8")])


def lib():
    """Load libqslhip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QslHipError(EDEVICE, f"{LIB_PATH} is not built: run "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc == OK:
        return
    L = lib()
    msg = L.qslhip_last_error().decode() or L.qslhip_status_string(rc).decode()
    if rc == ESIZE:
        raise SizeError(rc, msg)
    if rc == ENOTNP:
        raise NotNPError(rc, msg)
    raise QslHipError(rc, f"{L.qslhip_status_string(rc).decode()}: {msg}")


def ptr(a):
    return a.ctypes.data_as(_vp)


def uniforms(seed, count):
    """mt19937_64(splitmix64(seed)) uniforms (x >> 11) * 2^-53 in [0, 1)."""
    out = np.empty(count, np.float64)
    check(lib().qslhip_uniforms(seed, count, ptr(out)))
    return out


def device_count():
    n = _I(0)
    rc = lib().qslhip_device_count(C.byref(n))
    return n.value if rc == OK else 0


def plan(nqubits, prec, ops, cap=4096):
    """Tile passes qslhip_apply_ops(FUSED) would run: (masks, op_counts)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    npasses = _u64(0)
    masks = np.zeros(cap, np.uint64)
    counts = np.zeros(cap, np.uint64)
    check(lib().qslhip_plan(nqubits, prec, ptr(ops), len(ops), C.byref(npasses),
                            ptr(masks), ptr(counts), cap))
    m = min(cap, npasses.value)
    return masks[:m].copy(), counts[:m].copy()


def plan_order(nqubits, prec, ops, mode=MODE_REORDER, cap=4096):
    """Tile passes and execution order of qslhip_apply_ops(mode):
    (masks, op_counts, order) with order[i] the index of the i-th gate run."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    npasses = _u64(0)
    masks = np.zeros(cap, np.uint64)
    counts = np.zeros(cap, np.uint64)
    order = np.zeros(max(1, len(ops)), np.uint64)
    check(lib().qslhip_plan_mode(nqubits, prec, ptr(ops), len(ops), mode, C.byref(npasses),
                                 ptr(masks), ptr(counts), cap, ptr(order)))
    m = min(cap, npasses.value)
    return masks[:m].copy(), counts[:m].copy(), order[:len(ops)].copy()


def jit_source(nqubits, prec, ops, mode, pass_index):
    """Device source the run-time compiler generates for one pass."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    n = _u64(0)
    check(lib().qslhip_jit_source(nqubits, prec, ptr(ops), len(ops), mode, pass_index, None, 0,
                                  C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib().qslhip_jit_source(nqubits, prec, ptr(ops), len(ops), mode, pass_index, buf,
                                  n.value + 1, C.byref(n)))
    return buf.value.decode()


def jit_check(src, arch="gfx950"):
    """Compile a generated pass with hipRTC (host only): (ok, log)."""
    log = C.create_string_buffer(1 << 16)
    rc = lib().qslhip_jit_check(src.encode(), arch.encode(), log, len(log))
    return rc == OK, log.value.decode(errors="replace")


def tuning_get(key):
    v = _I(0)
    check(lib().qslhip_tuning_get(key.encode(), C.byref(v)))
    return v.value


def jit_wait():
    """Block until every pass structure queued for compilation is compiled."""
    check(lib().qslhip_jit_wait())


def tuning_set(key, value):
    check(lib().qslhip_tuning_set(key.encode(), int(value)))


def autotune(prec, nqubits):
    """Time tile / fusion widths on the device; returns the JSON report."""
    import json
    buf = C.create_string_buffer(1 << 14)
    check(lib().qslhip_autotune(prec, nqubits, buf, len(buf)))
    return json.loads(buf.value.decode())


def plan_dense(nqubits, ops, F):
    """Dense-mode plan: (pass_info [P,3], op_info [M,9], mats float32)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    a, b, c = _u64(0), _u64(0), _u64(0)
    check(lib().qslhip_plan_dense(nqubits, ptr(ops), len(ops), F, C.byref(a), C.byref(b),
                                  C.byref(c), None, None, None, 0, 0, 0))
    pi = np.zeros((a.value, 3), np.uint64)
    oi = np.zeros((b.value, 9), np.int64)
    m = np.zeros(max(1, c.value), np.float32)
    check(lib().qslhip_plan_dense(nqubits, ptr(ops), len(ops), F, C.byref(a), C.byref(b),
                                  C.byref(c), ptr(pi), ptr(oi), ptr(m), a.value, b.value,
                                  c.value))
    return pi, oi, m[:c.value]
