"""Sampling sums left by the last gate pass.

The last register-tile pass of apply_ops writes the |a|^2 sum of every
8-amplitude run of the state it stores; sampling then skips its own pass
over the state.  The sums use the association the sampler's own pass uses,
so outcomes are the same bit for bit whether or not a gate pass left the
sums, and they follow the oracle (src/qubits/default/measure.cpp:60-78,
164-188) except for uniforms within 1e-12 of a CDF boundary.  Every entry
point that changes the state afterwards must drop the sums.
"""
import numpy as np
import pytest

import qsl_amd
from qsl_amd import _lib as L
from oracle.pyoracle import Oracle, OP_DTYPE, GATE1, GATE2, near_cdf_boundary

pytestmark = pytest.mark.gpu
O = Oracle()


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


_near_boundary = near_cdf_boundary


def _check_oracle(st, u, got, dtype):
    want = O.sample_all(st, u)
    _, cdf = O.gen_dist(st)
    uf = u if dtype == np.complex128 else u.astype(np.float32).astype(np.float64)
    eps = 1e-12 if dtype == np.complex128 else 1e-5
    ok = (got == want) | _near_boundary(cdf, uf, eps)
    assert ok.all(), np.nonzero(~ok)[0][:10]
    # below the total mass, outcomes never land on zero-probability entries
    inside = uf < cdf[-1] - eps
    assert np.all(np.abs(st[got[inside].astype(np.int64)]) > 0)


def _counts2(q, u):
    u = np.ascontiguousarray(u, np.float64)
    counts = np.zeros(1 << q.nqubits, np.uint64)
    L.check(L.lib().qslhip_sample_all2(q._h, L.ptr(u), len(u), L.ptr(counts), counts.size))
    return counts


def _sampled(q):
    # a first sampling gives the handle its sums buffer; gate passes only
    # write run sums into one that exists
    q.sampleAllOutcomes(np.array([0.5]))
    return q


def _used():
    return L.tuning_get("run_sums_used")


def _drop_sums(q):
    # a state change that leaves the amplitudes as they are
    q.setAmps(q.getAmps(0, 1), 0)


def _set_widths(sfx, logt, low):
    try:
        L.tuning_set("logt_" + sfx, logt)
        L.tuning_set("low_" + sfx, low)
    except qsl_amd.QslHipError:
        L.tuning_set("low_" + sfx, low)
        L.tuning_set("logt_" + sfx, logt)


@pytest.fixture
def widths():
    saved = {s: (L.tuning_get("logt_" + s), L.tuning_get("low_" + s),
                 L.tuning_get("rthreads_" + s)) for s in ("f64", "f32")}

    def setter(sfx, threads, logt, low):
        _set_widths(sfx, 11, 3)
        L.tuning_set("rthreads_" + sfx, threads)
        _set_widths(sfx, logt, low)

    yield setter
    for sfx, (logt, low, thr) in saved.items():
        _set_widths(sfx, 11, 3)
        L.tuning_set("rthreads_" + sfx, thr)
        _set_widths(sfx, logt, low)


@pytest.fixture
def jit_mode():
    saved = L.tuning_get("jit")
    yield lambda v: L.tuning_set("jit", v)
    L.tuning_set("jit", saved)


@pytest.mark.parametrize("jit", [0, 2])
@pytest.mark.parametrize("mode", [qsl_amd.MODE_FUSED, qsl_amd.MODE_REORDER])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_outcomes_equal_with_and_without_pass_sums(jit_mode, jit, mode, dtype):
    n = 20
    jit_mode(jit)
    rng = np.random.default_rng(60 + jit + 4 * mode)
    ops = _ops(rng, n, 150, dtype)
    q = _sampled(qsl_amd.Qubits(n, dtype=dtype))
    q.apply_ops(ops, mode=mode)
    u = qsl_amd.uniforms(7 + mode, 30000)
    u[:3] = [0.0, 0.5, 1.0 - 2 ** -53]
    used = _used()
    a = q.sampleAllOutcomes(u)
    assert _used() == used + 1  # the pass's sums were taken
    _drop_sums(q)
    b = q.sampleAllOutcomes(u)
    assert np.array_equal(a, b)
    # sampleAll2's strict merge rule (`rand < sum`), counts per outcome
    q.apply_ops(ops[:40], mode=mode)
    ca = _counts2(q, u)
    _drop_sums(q)
    assert np.array_equal(ca, _counts2(q, u))
    if dtype == np.complex128:
        # against the oracle's merge (an FP64 running sum there too): a
        # uniform may change outcome only within 1e-12 of a running sum
        st = q.getState()
        want2 = O.sample_all2(st, u)
        _, cdf = O.gen_dist(st)
        near = int(near_cdf_boundary(cdf, u, 1e-12).sum())
        assert np.abs(ca.astype(np.int64) - want2.astype(np.int64)).sum() <= 2 * near
    _check_oracle(q.getState(), u, q.sampleAllOutcomes(u), dtype)
    q.apply_ops(ops[:0], mode=mode)
    _check_oracle(q.getState(), u, q.sampleAllOutcomes(u), dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("cfg", [(256, 10, 3), (256, 12, 4), (512, 12, 4), (512, 13, 5),
                                 (256, 11, 6)], ids=lambda c: "k%d_t%d_low%d" % c)
def test_run_layout_of_every_tile_width(widths, cfg, dtype):
    # the sums a pass leaves are regrouped from its own group order: every
    # tile width, and last tiles with qubits 3..5 inside and outside
    threads, logt, low = cfg
    sfx = "f64" if dtype == np.complex128 else "f32"
    if threads == 512 and logt == 13 and sfx == "f64":
        pytest.skip("2^13 FP64 amplitudes exceed the 64 KiB LDS remap buffer")
    widths(sfx, threads, logt, low)
    n = 21
    u = qsl_amd.uniforms(31 + logt, 20000)
    for seed in range(3):
        rng = np.random.default_rng(100 * logt + 10 * low + seed)
        ops = _ops(rng, n, 60 + 40 * seed, dtype)
        q = _sampled(qsl_amd.Qubits(n, dtype=dtype))
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        used = _used()
        a = q.sampleAllOutcomes(u)
        assert _used() == used + 1
        _drop_sums(q)
        assert np.array_equal(a, q.sampleAllOutcomes(u)), seed
    _check_oracle(q.getState(), u, a, dtype)


@pytest.mark.parametrize("change", ["gate1", "gate2", "apply_per_gate", "collapse", "set_state",
                                    "reset", "set_basis", "expval_grad"])
def test_state_changes_drop_the_sums(change):
    # (a handle that never sampled gets no sums: see the first test)
    n = 20
    rng = np.random.default_rng(80)
    q = _sampled(qsl_amd.Qubits(n))
    q.apply_ops(_ops(rng, n, 120, np.complex128), mode=qsl_amd.MODE_FUSED)
    if change == "gate1":
        q.hadamard(0)
        q.rotateY(17, 0.3)
    elif change == "gate2":
        q.controlNot(19, 0)
    elif change == "apply_per_gate":
        q.apply_ops(_ops(rng, n, 5, np.complex128), mode=qsl_amd.MODE_PER_GATE)
    elif change == "collapse":
        q.postselect(3, 1)
    elif change == "set_state":
        st = q.getState()
        q.setState(np.roll(st, 5))
    elif change == "reset":
        q.reset()
        q.hadamard(4)
    elif change == "set_basis":
        q.setBasisState(77)
    elif change == "expval_grad":
        terms = np.zeros(1, qsl_amd.PAULI_DTYPE)
        terms[0]["coef"] = 1.0
        terms[0]["zmask"] = 1
        q.expvalGrad(np.array([(5, 2, 0, 0, 0.4)], OP_DTYPE), terms)
    u = qsl_amd.uniforms(81, 20000)
    used = _used()
    got = q.sampleAllOutcomes(u)
    assert _used() == used  # the change dropped them
    _check_oracle(q.getState(), u, got, np.complex128)


def test_measure_all_after_gates():
    n = 20
    rng = np.random.default_rng(90)
    ops = _ops(rng, n, 100, np.complex128)
    q = _sampled(qsl_amd.Qubits(n))
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    st = q.getState()
    used = _used()
    o = q.measureAll(0.4321)
    assert _used() == used + 1
    assert o == O.sample_all(st, np.array([0.4321]))[0]
    s = q.getState()
    assert s[o] == 1 and np.count_nonzero(s) == 1


def test_no_sums_before_the_first_sampling():
    # a handle that never sampled holds no sums buffer, so its gate passes
    # write none and its first sampling makes its own pass over the state
    n = 20
    rng = np.random.default_rng(91)
    q = qsl_amd.Qubits(n)
    q.apply_ops(_ops(rng, n, 80, np.complex128), mode=qsl_amd.MODE_REORDER)
    used = _used()
    u = qsl_amd.uniforms(92, 5000)
    got = q.sampleAllOutcomes(u)
    assert _used() == used
    _check_oracle(q.getState(), u, got, np.complex128)
    q.apply_ops(_ops(rng, n, 80, np.complex128), mode=qsl_amd.MODE_REORDER)
    got = q.sampleAllOutcomes(u)
    assert _used() == used + 1
    _check_oracle(q.getState(), u, got, np.complex128)


def test_sampling_at_26_qubits():
    # 2^14 chunk prefixes: the multi-workgroup scan (k_scan_local, k_scan_add)
    n = 26
    rng = np.random.default_rng(26)
    ops = _ops(rng, n, 120, np.complex128)
    q = _sampled(qsl_amd.Qubits(n))
    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    u = qsl_amd.uniforms(2626, 20000)
    u[:2] = [0.0, 1.0 - 2 ** -53]
    used = _used()
    a = q.sampleAllOutcomes(u)
    assert _used() == used + 1
    _drop_sums(q)
    assert np.array_equal(a, q.sampleAllOutcomes(u))
    _check_oracle(q.getState(), u, a, np.complex128)
