"""qsl_amd -- MI355X (gfx950) state-vector backend for qsl.

The product is libqslhip.so (HIP kernels behind the C ABI of
include/qslhip.h) and the C++20 headers include/qsl/qubits/hip*.hpp.  This
package is the Python mirror of the same API used by the tests and
bench.py; it has no CPU fallback.
"""
from ._lib import (QslHipError, SizeError, NotNPError, uniforms, device_count,  # noqa: F401
                   OP_DTYPE, MODE_PER_GATE, MODE_FUSED, MODE_DENSE, MODE_REORDER, LIB_PATH,
                   HEADER_PATH)
from . import _lib as abi  # noqa: F401
from .qubits import Qubits, QubitsNP, QubitsResize, Batch  # noqa: F401
from ._lib import PAULI_DTYPE  # noqa: F401
