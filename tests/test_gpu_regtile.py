"""Register-resident tile passes (the default fused path for 10..12 tile
qubits) against the oracle, bit for bit.

Each pass keeps 2^(K-8) amplitudes per thread in registers and goes through
LDS only to remap which tile qubits index the registers, as planned on the
host for the targets that follow.  Every tile width, contiguous low run and
both precisions run the same random all-kinds circuit; the LDS-resident
kernel (reg_tile = 0) runs it too, as a second implementation that must
agree bit for bit.
"""
import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2, random_state

pytestmark = pytest.mark.gpu

O = Oracle()


def _set_widths(sfx, logt, low):
    # tile widths are validated pairwise; one of the two orders is legal
    try:
        L.tuning_set("logt_" + sfx, logt)
        L.tuning_set("low_" + sfx, low)
    except qsl_amd.QslHipError:
        L.tuning_set("low_" + sfx, low)
        L.tuning_set("logt_" + sfx, logt)


@pytest.fixture
def tuning():
    saved = {s: (L.tuning_get("logt_" + s), L.tuning_get("low_" + s)) for s in ("f64", "f32")}
    saved_reg = L.tuning_get("reg_tile")

    saved_thr = {s: L.tuning_get("rthreads_" + s) for s in ("f64", "f32")}
    saved_min = L.tuning_get("reg_min_qubits")
    L.tuning_set("reg_min_qubits", 0)  # register tiles on small states too

    def setter(kern, prec, logt, low):
        # kern: 0 = LDS-resident tile kernel, 256 / 512 = register tiles
        sfx = "f64" if prec == 8 else "f32"
        L.tuning_set("reg_tile", 1)  # admits every width the LDS kernel admits
        if kern:
            _set_widths(sfx, 11, 3)
            L.tuning_set("rthreads_" + sfx, kern)
        _set_widths(sfx, logt, low)
        L.tuning_set("reg_tile", 1 if kern else 0)

    yield setter
    L.tuning_set("reg_tile", 1)
    for s, (logt, low) in saved.items():
        _set_widths(s, 11, 3)
        L.tuning_set("rthreads_" + s, saved_thr[s])
        _set_widths(s, logt, low)
    L.tuning_set("reg_tile", saved_reg)
    L.tuning_set("reg_min_qubits", saved_min)


def _ops(rng, n, count, dtype):
    kinds = GATE1 + GATE2
    ops = np.zeros(count, OP_DTYPE)
    for i in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ang = float(rng.uniform(-2 * np.pi, 2 * np.pi))
        if dtype == np.complex64:
            ang = float(np.float32(ang))
        ops[i] = (k, a, b if k >= 16 else 0, 0, ang)
    return ops


CONFIGS = [(256, 10, 2), (256, 10, 3), (256, 11, 3), (256, 11, 5), (256, 12, 4),
           (256, 12, 6), (512, 12, 4), (512, 13, 5), (0, 10, 3), (0, 11, 3)]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [13, 19])
@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "k%d_t%d_low%d" % c)
def test_tile_kernels_match_oracle(tuning, dtype, n, cfg):
    kern, logt, low = cfg
    prec = 8 if dtype == np.complex128 else 4
    if kern == 512 and logt == 13 and prec == 8:
        pytest.skip("2^13 FP64 amplitudes exceed the 64 KiB LDS remap buffer")
    tuning(kern, prec, logt, low)
    # states narrower than 2^(logt+6) run narrower tiles (at least 64
    # workgroups), never below 2^10 with register tiles on
    rng = np.random.default_rng(1000 * logt + 10 * low + n + prec)
    st0 = random_state(rng, n, dtype)
    ops = _ops(rng, n, 300, dtype)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    got = q.getState()
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.abs(got - want).max() <= (1e-12 if prec == 8 else 1e-5)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), cfg


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_register_remaps_on_every_qubit(tuning, dtype):
    # targets sweep every tile position in both directions, so each layout
    # is left for every other one; controls and diagonals on thread bits,
    # register bits and group bits
    n = 18
    prec = 8 if dtype == np.complex128 else 4
    tuning(256, prec, 12, 4)
    rng = np.random.default_rng(18)
    seq = []
    for t in list(range(n)) + list(range(n - 1, -1, -1)):
        seq.append((5, t, 0, 0, 0.37 + t))
        seq.append((16, (t + 5) % n, t, 0, 0.0))
        seq.append((21, t, (t + 9) % n, 0, 1.1 * t))
        seq.append((22, (t + 3) % n, t, 0, 0.5))
        seq.append((24, t, (t + 1) % n, 0, 0.0))
        seq.append((26, (t + 2) % n, t, 0, 0.9))
    ops = np.array(seq, OP_DTYPE)
    if dtype == np.complex64:
        ops["angle"] = ops["angle"].astype(np.float32)
    st0 = random_state(rng, n, dtype)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.array_equal(q.getState().view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("kern", [256, 512])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_register_remaps_at_512_threads(tuning, kern, dtype):
    n = 18
    prec = 8 if dtype == np.complex128 else 4
    tuning(kern, prec, 12, 4)
    rng = np.random.default_rng(170 + kern)
    st0 = random_state(rng, n, dtype)
    ops = _ops(rng, n, 400, dtype)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.array_equal(q.getState().view(np.uint8), want.view(np.uint8))


def test_tuning_rejects_lds_widths_without_register_tiles(tuning):
    tuning(256, 8, 12, 4)  # 2^12 FP64 amplitudes: only the register kernel fits
    with pytest.raises(qsl_amd.QslHipError):
        L.tuning_set("reg_tile", 0)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("cfg", [(256, 11, 3), (512, 12, 4), (0, 10, 3)],
                         ids=lambda c: "k%d_t%d_low%d" % c)
def test_reorder_mode_matches_program_order(tuning, dtype, cfg):
    # commutation-aware reordering: the same product, so the state matches
    # the oracle's program-order run to rounding (the stated FP64 bound,
    # 1e-12 on every amplitude; FP32 to its own rounding)
    kern, logt, low = cfg
    prec = 8 if dtype == np.complex128 else 4
    tuning(kern, prec, logt, low)
    n = 18
    rng = np.random.default_rng(4000 + logt + prec)
    st0 = random_state(rng, n, dtype)
    ops = _ops(rng, n, 400, dtype)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.abs(q.getState() - want).max() <= (1e-12 if prec == 8 else 1e-5)


def test_reorder_mode_at_22_qubits_from_zero_state(tuning):
    from qsl_amd import workloads as W
    tuning(256, 8, 11, 3)
    n = 22
    ops, _ = W.c2_random(n, 200, W.Stream(22))
    a = qsl_amd.Qubits(n)
    b = qsl_amd.Qubits(n)
    a.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    b.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    sa, sb = a.getState(), b.getState()
    assert np.abs(sa - sb).max() <= 1e-12
    assert abs(np.vdot(sb, sb).real - 1.0) <= 1e-12


@pytest.fixture
def jit_mode():
    saved = L.tuning_get("jit")

    def setter(v):
        L.tuning_set("jit", v)

    yield setter
    L.tuning_set("jit", saved)


@pytest.fixture
def pass_shape():
    # free_layout: load / store layouts of the first / last gate (else the
    # top register positions); jit_prefetch: 0 one group per workgroup,
    # 1 a group loop with the next group's loads in flight
    # (pass_tune off: the shape asked for is the shape that runs)
    keys = ("free_layout", "jit_prefetch", "pass_tune")
    saved = {k: L.tuning_get(k) for k in keys}

    def setter(free_layout, prefetch):
        L.tuning_set("free_layout", free_layout)
        L.tuning_set("jit_prefetch", prefetch)
        L.tuning_set("pass_tune", 0)

    yield setter
    for k, v in saved.items():
        L.tuning_set(k, v)


@pytest.mark.parametrize("shape", [(1, 0), (1, 1), (0, 1)], ids=["free_single", "free_loop",
                                                               "top_loop"])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("cfg", [(256, 10, 3), (256, 11, 3), (512, 12, 4)],
                         ids=lambda c: "k%d_t%d_low%d" % c)
def test_compiled_passes_match_oracle(tuning, jit_mode, pass_shape, dtype, cfg, shape):
    # every pass compiled at run time (jit = 2): program order bit for bit,
    # the reordered mode to rounding, with each pass shape
    kern, logt, low = cfg
    prec = 8 if dtype == np.complex128 else 4
    tuning(kern, prec, logt, low)
    jit_mode(2)
    pass_shape(*shape)
    before = L.tuning_get("jit_compiled")
    n = 18
    rng = np.random.default_rng(7000 + logt + prec)
    st0 = random_state(rng, n, dtype)
    ops = _ops(rng, n, 150, dtype)
    want = st0.copy()
    O.apply_ops(want, ops)
    q = qsl_amd.Qubits(n, dtype=dtype)
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    got = q.getState()
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    q.setState(st0)
    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    assert np.abs(q.getState() - want).max() <= (1e-12 if prec == 8 else 1e-5)
    assert L.tuning_get("jit_compiled") > before
    assert L.tuning_get("jit_failed") == 0


@pytest.mark.parametrize("mode", [qsl_amd.MODE_FUSED, qsl_amd.MODE_REORDER])
def test_compiled_pass_equals_interpreted_pass(tuning, jit_mode, mode):
    # jit = 1: the first application interprets the decoded ops, the second
    # queues the structure for the compiler thread (and still interprets),
    # the third runs the compiled pass; the same arithmetic in the same order
    # gives the same bits, also with new angles (the compiled pass is
    # angle-free)
    tuning(256, 8, 11, 3)
    jit_mode(1)
    n = 19
    rng = np.random.default_rng(900 + mode)
    st0 = random_state(rng, n, np.complex128)
    ops = _ops(rng, n, 120, np.complex128)
    q = qsl_amd.Qubits(n)
    q.setState(st0)
    q.apply_ops(ops, mode=mode)
    first = q.getState()
    before = L.tuning_get("jit_compiled")
    q.setState(st0)
    q.apply_ops(ops, mode=mode)
    assert np.array_equal(q.getState().view(np.uint8), first.view(np.uint8))
    L.jit_wait()
    assert L.tuning_get("jit_compiled") > before
    q.setState(st0)
    q.apply_ops(ops, mode=mode)  # compiled passes
    assert np.array_equal(q.getState().view(np.uint8), first.view(np.uint8))
    ops2 = ops.copy()
    ops2["angle"] = rng.uniform(-6, 6, len(ops))
    q.setState(st0)
    q.apply_ops(ops2, mode=mode)  # compiled passes, new coefficients
    want = st0.copy()
    O.apply_ops(want, ops2)
    if mode == qsl_amd.MODE_FUSED:
        assert np.array_equal(q.getState().view(np.uint8), want.view(np.uint8))
    else:
        assert np.abs(q.getState() - want).max() <= 1e-12


@pytest.fixture
def group_order():
    saved = {k: L.tuning_get(k) for k in ("group_order", "jit")}
    yield lambda order, jit: (L.tuning_set("group_order", order), L.tuning_set("jit", jit))
    for k, v in saved.items():
        L.tuning_set(k, v)


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_group_orders_give_the_same_bits(group_order, jit, dtype):
    # passes of 2^13 groups (24 qubits, tiles of 2^11) in both group orders
    # and the default mix (the XCD sweep for the passes from a reset): the
    # same groups with the same arithmetic, so the states are bit-identical,
    # program order and reordered, and so are the shots sampled from the
    # last pass's run sums after a reset
    n = 24
    rng = np.random.default_rng(4100 + jit + (dtype == np.complex64))
    st0 = random_state(rng, n, dtype)
    ops = _ops(rng, n, 120, dtype)
    u = np.sort(rng.random(4096))
    got = {}
    for order in (0, 1, 2):
        group_order(order, jit)
        q = qsl_amd.Qubits(n, dtype=dtype)
        q.setState(st0)
        q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        fused = q.getState()
        q.setState(st0)
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        reord = q.getState()
        q.sampleAllOutcomes(np.array([0.5]))  # the handle holds its sampling buffer
        q.reset()
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        shots = q.sampleAllOutcomes(u)
        got[order] = (fused, reord, shots, q.getState())
    for order in (1, 2):
        for a, b in zip(got[0], got[order]):
            assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), order
    want = st0.copy()
    O.apply_ops(want, ops)
    assert np.array_equal(got[0][0].view(np.uint8), want.view(np.uint8))


def test_pass_shape_tuner_gives_the_same_bits(jit_mode):
    # FP64 passes that read whole tiles try their four launch shapes on the
    # first launches (one group per workgroup or a group loop, group order
    # 0 or 1) and keep the fastest: every application along the way, and
    # after the choice, gives the bits of the fixed shape
    saved = {k: L.tuning_get(k) for k in ("pass_tune", "reg_min_qubits")}
    try:
        L.tuning_set("reg_min_qubits", 0)
        jit_mode(2)
        n = 22
        rng = np.random.default_rng(5150)
        st0 = random_state(rng, n, np.complex128)
        ops = _ops(rng, n, 100, np.complex128)
        L.tuning_set("pass_tune", 0)
        q = qsl_amd.Qubits(n)
        q.setState(st0)
        q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        want = q.getState()
        L.tuning_set("pass_tune", 1)
        for _ in range(6):
            q.setState(st0)
            q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
            q.sync()
            assert np.array_equal(q.getState().view(np.uint8), want.view(np.uint8))
    finally:
        for k, v in saved.items():
            L.tuning_set(k, v)
