"""Passes compiled at run time: the generated source of every register-tile
pass compiles with hipRTC for gfx950 (host only, no device).

Random circuits of every gate kind, both precisions, the exact (FUSED) and
tolerance-bound (REORDER) arithmetic and every register width (2^10, 2^11,
2^12 tiles with 256 threads, 2^12 with 512), so each helper instantiation the
generator can emit is compiled at least once.  Parity of the compiled passes
is checked on the GPU (tests/test_gpu_regtile.py)."""
import numpy as np
import pytest

from qsl_amd import _lib as L
from oracle.pyoracle import OP_DTYPE, GATE1, GATE2


def _ops(rng, n, count):
    kinds = GATE1 + GATE2
    ops = np.zeros(count, OP_DTYPE)
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops[i] = (k, a, b if k >= 16 else 0, 0, float(rng.uniform(-6, 6)))
    return ops


def _widths(sfx, logt, low):
    try:
        L.tuning_set("logt_" + sfx, logt)
        L.tuning_set("low_" + sfx, low)
    except L.QslHipError:
        L.tuning_set("low_" + sfx, low)
        L.tuning_set("logt_" + sfx, logt)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("mode", [L.MODE_FUSED, L.MODE_REORDER])
@pytest.mark.parametrize("cfg", [(256, 10, 3), (256, 11, 3), (256, 12, 4), (512, 12, 4)],
                         ids=lambda c: "k%d_t%d_low%d" % c)
def test_generated_passes_compile(prec, mode, cfg):
    kern, logt, low = cfg
    sfx = "f64" if prec == 8 else "f32"
    saved = (L.tuning_get("logt_" + sfx), L.tuning_get("low_" + sfx),
             L.tuning_get("rthreads_" + sfx), L.tuning_get("reg_tile"))
    try:
        L.tuning_set("reg_tile", 1)
        _widths(sfx, 11, 3)
        L.tuning_set("rthreads_" + sfx, kern)
        _widths(sfx, logt, low)
        n = 20
        rng = np.random.default_rng(50 + logt + prec + mode)
        ops = _ops(rng, n, 120)
        npass = len(L.plan_order(n, prec, ops, mode)[0])
        for p in range(min(npass, 3)):
            src = L.jit_source(n, prec, ops, mode, p)
            assert "qslhip_pass" in src
            ok, log = L.jit_check(src, "gfx950")
            assert ok, log[:3000]
    finally:
        _widths(sfx, 11, 3)
        L.tuning_set("rthreads_" + sfx, saved[2])
        _widths(sfx, saved[0], saved[1])
        L.tuning_set("reg_tile", saved[3])


def test_source_is_angle_free():
    # the compiled pass is reused for new angles: they are not in the source
    n = 20
    rng = np.random.default_rng(3)
    ops = _ops(rng, n, 60)
    other = ops.copy()
    other["angle"] = rng.uniform(-6, 6, len(ops))
    for p in range(2):
        assert L.jit_source(n, 8, ops, L.MODE_REORDER, p) == \
            L.jit_source(n, 8, other, L.MODE_REORDER, p)
