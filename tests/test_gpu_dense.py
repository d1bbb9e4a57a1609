"""Dense fused unitaries on FP32 matrix cores (QSLHIP_MODE_DENSE).

The block matrices are products of the gate matrices in FP64, rounded once
to FP32, and applied by v_mfma_f32_32x32x2f32 on LDS tiles, so results are
compared with the FP64 oracle within a stated FP32 bound: max amplitude
error <= 2e-6 * (number of dense ops a column vector passes through) + 1e-6,
checked on circuits of every gate kind and on the Haar-layer family of the
34-qubit workload at 14-18 qubits.
"""
import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L
from qsl_amd import workloads as W
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2

pytestmark = pytest.mark.gpu
O = Oracle()


def _fp64_reference(n, ops):
    st = np.zeros(1 << n, np.complex128)
    st[0] = 1
    O.apply_ops(st, ops)
    return st


@pytest.fixture(params=[5, 4])
def fusion_width(request):
    old = L.tuning_get("dense_f")
    L.tuning_set("dense_f", request.param)
    yield request.param
    L.tuning_set("dense_f", old)


@pytest.mark.parametrize("n", [12, 14, 17])
def test_haar_layers_vs_fp64(n, fusion_width):
    ops, _ = W.c5_haar_layers(n, 6, W.Stream(50 + n))
    q = qsl_amd.Qubits(n, dtype=np.complex64)
    q.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
    got = q.getState().astype(np.complex128)
    want = _fp64_reference(n, ops)
    err = np.abs(got - want).max()
    assert err <= 2e-6 * 6 * 4 + 1e-6, err
    assert abs(np.linalg.norm(got) - 1) < 1e-5


@pytest.fixture(params=[0, 1], ids=["program_order", "reordered"])
def dense_reorder(request):
    old = L.tuning_get("dense_reorder")
    L.tuning_set("dense_reorder", request.param)
    yield request.param
    L.tuning_set("dense_reorder", old)


@pytest.mark.parametrize("n", [13, 15])
def test_every_kind_vs_fp64(n, fusion_width, dense_reorder):
    # both dense plans: tile passes in program order and commutation-aware
    rng = np.random.default_rng(n)
    ops = np.zeros(300, OP_DTYPE)
    kinds = GATE1 + GATE2
    for i in range(len(ops)):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops[i] = (k, a, b if k >= 16 else 0, 0, float(rng.uniform(0, 2 * np.pi)))
    q = qsl_amd.Qubits(n, dtype=np.complex64)
    q.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
    got = q.getState().astype(np.complex128)
    want = _fp64_reference(n, ops)
    assert np.abs(got - want).max() <= 1e-4
    # the arithmetic-preserving FP32 tile path agrees to FP32 rounding too
    q2 = qsl_amd.Qubits(n, dtype=np.complex64)
    q2.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    assert np.abs(q2.getState().astype(np.complex128) - got).max() <= 1e-4


def test_dense_mode_validation():
    q = qsl_amd.Qubits(12)  # FP64: dense mode is FP32-only
    with pytest.raises(qsl_amd.QslHipError):
        q.apply_ops(np.zeros(1, OP_DTYPE), mode=qsl_amd.MODE_DENSE)
    q = qsl_amd.Qubits(10, dtype=np.complex64)  # needs >= 12 qubits
    with pytest.raises(qsl_amd.QslHipError):
        q.apply_ops(np.zeros(1, OP_DTYPE), mode=qsl_amd.MODE_DENSE)


def test_autotune_sets_valid_widths():
    keys = ("logt_f64", "low_f64", "logt_f32", "low_f32", "lds_max_target_f32", "dense_f",
            "rthreads_f64", "rthreads_f32")
    saved = {k: L.tuning_get(k) for k in keys}
    saved_reg = L.tuning_get("reg_tile")
    try:
        rep = L.autotune(4, 22)
        kern = L.tuning_get("rthreads_f32") if L.tuning_get("reg_tile") else 0
        assert rep["tile_best"] == [kern, L.tuning_get("logt_f32"), L.tuning_get("low_f32")]
        assert L.tuning_get("dense_f") in (4, 5)
        # the tuned configuration still computes the same circuit
        n = 16
        ops, _ = W.c2_random(n, 80, W.Stream(60))
        a = qsl_amd.Qubits(n, dtype=np.complex64)
        a.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        b = qsl_amd.Qubits(n, dtype=np.complex64)
        b.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
        assert np.array_equal(a.getState().view(np.uint8), b.getState().view(np.uint8))
    finally:
        L.tuning_set("reg_tile", 1)  # admits every width the LDS kernel admits
        for k, v in saved.items():
            L.tuning_set(k, v)
        L.tuning_set("reg_tile", saved_reg)
