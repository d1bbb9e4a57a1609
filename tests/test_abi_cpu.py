"""Host-side checks of the C-ABI library that need no GPU: it loads, it
exports every function include/qslhip.h declares, its enums agree with the
oracle's and the Python mirror's, argument validation rejects bad input
before any device work, and without a device every entry point fails loudly
(there is no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L

HDR = open(L.HEADER_PATH).read()


def _declared_functions():
    body = re.sub(r"/\*.*?\*/", "", HDR, flags=re.S)
    return sorted(set(re.findall(r"\b(qslhip_[a-z0-9_]+)\s*\(", body)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds all of them
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def _enum(name):
    m = re.search(r"\b%s\s*=\s*(\d+)" % name, HDR)
    return int(m.group(1))


def test_enums_agree_with_oracle_and_mirror():
    oh = open(os.path.join(os.path.dirname(L.HERE), "oracle", "qsl_oracle.h")).read()
    for n in ["PAULI_X", "PAULI_Y", "PAULI_Z", "HADAMARD", "PHASE", "ROT_X", "ROT_Y",
              "ROT_Z", "CNOT", "CY", "CZ", "CROT_X", "CROT_Y", "CROT_Z", "CPHASE",
              "CHADAMARD", "SWAP", "FSWAP", "NP_ROT_X", "NP_ROT_Y", "NP_HADAMARD"]:
        v = _enum("QSLHIP_" + n)
        assert int(re.search(r"\bORC_%s\s*=\s*(\d+)" % n, oh).group(1)) == v
        assert getattr(L, n) == v
    for i, n in enumerate(["OK", "EINVAL", "ESIZE", "ENOTNP", "ENOMEM", "ENODEV",
                           "EDEVICE", "EKIND"]):
        assert _enum("QSLHIP_" + n) == i == getattr(L, n)
    assert L.OP_DTYPE.itemsize == 24


def test_argument_validation_before_device_work():
    lib = L.lib()
    h = C.c_void_p()
    assert lib.qslhip_create(0, 8, C.byref(h)) == L.EINVAL
    assert lib.qslhip_create(41, 8, C.byref(h)) == L.EINVAL
    assert lib.qslhip_create(10, 2, C.byref(h)) == L.EINVAL
    assert lib.qslhip_create_np(10, 0, 8, C.byref(h)) == L.ENOTNP
    assert lib.qslhip_create_np(4, 4, 8, C.byref(h)) == L.ENOTNP
    assert lib.qslhip_create(10, 8, None) == L.EINVAL
    assert lib.qslhip_gate1(None, 0, 0, 0.0) == L.EINVAL
    assert lib.qslhip_prob(None, 0, 0, None) == L.EINVAL
    assert b"null handle" in lib.qslhip_last_error()
    assert lib.qslhip_status_string(L.ESIZE) == b"state size mismatch"


def test_no_device_fails_loudly():
    if qsl_amd.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(qsl_amd.QslHipError) as e:
        qsl_amd.Qubits(4)
    assert e.value.code == L.ENODEV


def test_uniform_stream_is_seeded_and_in_range():
    a = qsl_amd.uniforms(20261202, 1000)
    b = qsl_amd.uniforms(20261202, 2000)
    assert np.array_equal(a, b[:1000])
    assert a.min() >= 0 and a.max() < 1
    assert not np.array_equal(a, qsl_amd.uniforms(20261203, 1000))
    # (x >> 11) * 2^-53 is a multiple of 2^-53
    assert np.all(np.floor(a * 2.0 ** 53) == a * 2.0 ** 53)
    assert abs(a.mean() - 0.5) < 0.05
