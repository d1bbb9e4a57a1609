"""Headline benchmark: 30-qubit FP64 state vector on one MI355X.

One step = reset to |0..0>, the 200-gate seeded random circuit of config 2
(40% RX/RY/RZ, 20% H, 20% CNOT, 20% CPhase, qubits uniform over 0..29),
then 10^6 shots sampled from the final state.  `value` is gates per second
over the whole timed region (reset + gates + sampling, inputs resident in
HBM).  The gates run in QSLHIP_MODE_REORDER: commutation-aware packing into
register-tile passes (the same product; every run checks the full 2^30
state against the bit-exact program-order result, max |diff| <= 1e-12).
Printed as one JSON line with the roofline of the dominant kernel (the
register tile pass, timed with HIP events on the library's stream) and the
reference's Omp backend timed on the host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--mode reorder|fused|per-gate]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("30q FP64 gates/s + HBM GB/s; batched adjoint grads/s, "
          "FP32 vs FP64 plain")
HBM_PEAK_GBPS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
N = 30
NGATES = 200
NSHOTS = 10 ** 6


class HipEvents:
    """hipEvent timing on an explicit stream, via the HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p,
                                                 C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def stream_of(q):
    from qsl_amd import _lib as L
    s = C.c_void_p()
    L.check(L.lib().qslhip_get_stream(q._h, C.byref(s)))
    return s


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(ops, gpu_state_fn, budget_s=20.0, leg_budget_s=8.0):
    """The reference's Omp backend (oracle/_ref, compiled from qsl's own
    sources) -- or the oracle's restatement of it when that library is
    absent -- applying the first gates of the same circuit to a 30-qubit
    FP64 state on the host cores, until ~budget_s of gate time is spent."""
    from oracle import pyoracle as P
    cores = host_cores()
    kind = "reference"
    try:
        ref = P.RefLib()
        sim = P.RefSim(ref, P.RefLib.OMP, np.complex128, N, cores)

        def apply(op):
            k = int(op["kind"])
            if k < 16:
                sim.gate1(k, int(op["q0"]), float(op["angle"]))
            else:
                sim.gate2(k, int(op["q0"]), int(op["q1"]), float(op["angle"]))
        get = sim.get_state
    except (FileNotFoundError, OSError):
        kind = "port"
        o = P.Oracle()
        st = np.zeros(1 << N, np.complex128)
        st[0] = 1

        def apply(op):
            k = int(op["kind"])
            if k < 16:
                o.omp_gate1(st, k, int(op["q0"]), float(op["angle"]), cores)
            else:
                o.omp_gate2(st, k, int(op["q0"]), int(op["q1"]), float(op["angle"]),
                            cores)

        def get():
            return st
    done, spent = 0, 0.0
    for op in ops:
        t0 = time.perf_counter()
        apply(op)
        spent += time.perf_counter() - t0
        done += 1
        if spent > budget_s:
            break
    # parity of the sample at full size: the GPU replays the same prefix
    cpu = get()
    gpu = gpu_state_fn(ops[:done])
    err = float(np.max(np.abs(cpu - gpu)))
    bit_exact = bool(np.array_equal(cpu.view(np.uint8), gpu.view(np.uint8)))
    del cpu, gpu
    legs = []
    if kind == "reference":
        # the reference's own defaults beside it: Qubits<Omp> at its default
        # 4 threads, and the single-threaded Qubits<Default>
        for typ, threads, name in ((P.RefLib.OMP, 4, "Qubits<Omp,double>"),
                                   (P.RefLib.DEFAULT, 1, "Qubits<Default,double>")):
            sim = None
            sim = P.RefSim(ref, typ, np.complex128, N, threads if typ == P.RefLib.OMP else 0)

            def apply_leg(op, sim=sim):
                k = int(op["kind"])
                if k < 16:
                    sim.gate1(k, int(op["q0"]), float(op["angle"]))
                else:
                    sim.gate2(k, int(op["q0"]), int(op["q1"]), float(op["angle"]))
            n_leg, t_leg = 0, 0.0
            for op in ops:
                t0 = time.perf_counter()
                apply_leg(op)
                t_leg += time.perf_counter() - t0
                n_leg += 1
                if t_leg > leg_budget_s:
                    break
            del sim
            legs.append({"backend": name, "threads": threads,
                         "gates_per_s": round(n_leg / t_leg, 3), "gates": n_leg,
                         "seconds": round(t_leg, 1)})
    return {"value": done / spent, "unit": "gates/s", "cores": cores, "kind": kind,
            "cpu_model": cpu_model(), "legs": legs,
            "sample": f"first {done} gates of the same 200-gate circuit on a "
                      f"30-qubit FP64 state ({spent:.1f} s of gate time); "
                      f"{'qsl Qubits<Omp,double> built from the reference sources' if kind == 'reference' else 'oracle restatement of the Omp loops'}",
            "max_abs_err_vs_gpu": err, "bit_exact_vs_gpu": bit_exact}


def step_pass_bytes(masks, n, amp_bytes=16):
    """HBM bytes the reordered tile passes move when the circuit starts
    from a basis state (the library's rule, regtile.hip pass_extras): pass p
    runs only the groups that agree with the basis outside the qubits
    earlier tiles held; with every qubit in some tile, the first pass only
    writes the basis index's group and a later pass reads only what an
    earlier one wrote; otherwise the first pass writes every group.  The
    last pass also writes 2^n bytes of run sums when it covers every group."""
    full = (1 << n) - 1
    allq = 0
    for m in masks:
        allq |= int(m)
    fill = (allq & full) != full
    support, total = 0, 0
    for i, m in enumerate(masks):
        m = int(m)
        t = bin(m).count("1")
        outside = full & ~m
        if i == 0 and fill:
            groups, reads = 1 << (n - t), 0  # every group written, none read
        elif i == 0:
            groups, reads = 1, 0  # the basis index's group, generated in registers
        else:
            groups = 1 << bin(outside & support).count("1")
            reads = (1 << bin(m & support).count("1")) if not fill else (1 << t)
        total += groups * ((1 << t) + reads) * amp_bytes
        support |= m
    if masks is not None and len(masks):
        last = int(masks[-1])
        if (full & ~last & ~support) == 0:
            total += 1 << n
    return total


def read_traffic():
    """Measured HBM bytes per tile-pass launch (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel's reading), if profiled."""
    path = os.path.join(ROOT, "profiles", "qslhip_pass_f64_n30_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def bench_c1():
    """Config 1: 16-qubit FP64, 20 layers (480 gates), prob of every qubit,
    10^4 shots -- against the reference's serial Default backend."""
    import qsl_amd
    from qsl_amd import workloads as W
    ops, s = W.c1_layers(16, 20)
    shots = s.take(10 ** 4)
    q = qsl_amd.Qubits(16)
    ev = HipEvents()
    st = stream_of(q)
    e0, e1 = ev.event(), ev.event()
    for _ in range(2):
        q.reset()
        ev.record(e0, st)
        q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        ev.record(e1, st)
        probs = [q.prob(t, o) for t in range(16) for o in (0, 1)]
        out = q.sampleAllOutcomes(shots)
    gates_ms = ev.elapsed_ms(e0, e1)
    t0 = time.perf_counter()
    q.reset()
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    probs = [q.prob(t, o) for t in range(16) for o in (0, 1)]
    out = q.sampleAllOutcomes(shots)
    wall = time.perf_counter() - t0
    res = {"gates": len(ops), "gpu_gate_ms": round(gates_ms, 4),
           "gpu_wall_ms_gates_probs_shots": round(1000 * wall, 3)}
    try:  # this config's CPU baseline: the reference's Default backend, serial, and parity
        from oracle import pyoracle as P
        ref = P.RefLib()
        sim = P.RefSim(ref, P.RefLib.DEFAULT, np.complex128, 16)
        t0 = time.perf_counter()
        for op in ops:
            k = int(op["kind"])
            if k < 16:
                sim.gate1(k, int(op["q0"]), float(op["angle"]))
            else:
                sim.gate2(k, int(op["q0"]), int(op["q1"]), float(op["angle"]))
        res["cpu_reference_default_ms"] = round(1000 * (time.perf_counter() - t0), 2)
        cpu = sim.get_state()
        q.reset()
        q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        gpu = q.getState()
        res["bit_exact_vs_reference"] = bool(np.array_equal(cpu.view(np.uint8),
                                                            gpu.view(np.uint8)))
        o = P.Oracle()
        want = o.sample_all(cpu, shots)
        res["shots_equal_to_oracle"] = int(np.sum(want == out))
    except (FileNotFoundError, OSError) as e:
        res["cpu_reference"] = f"unavailable: {e}"
    return res


def np_pass_bytes(ops, n, k, low_bits=12, hi_max=8):
    """HBM bytes of the passes the number-preserving backend launches for a
    list of phases and npRotateX/Y bonds (its grouping rules for these two
    kinds): a run of phases is one pass over the compact vector; two or more
    disjoint adjacent bonds below qubit `low_bits` one pass over it; two to
    `hi_max` disjoint bonds with both qubits in the high part (qubits >= L)
    one pass over the segments in which some pair's bits differ; any other
    bond reads and writes both amplitudes of its C(n-2, k-1) pairs."""
    from qsl_amd import _lib as L
    Lq = min(n, max(16, n - 16))
    namps = math.comb(n, k)
    his = np.arange(1 << (n - Lq), dtype=np.int64)
    pc = np.array([bin(int(h)).count("1") for h in his])
    ok = (pc <= k) & (k - pc <= Lq)
    seg = np.array([math.comb(Lq, k - int(p)) if o else 0 for p, o in zip(pc, ok)], np.int64)
    hi_cache = {}

    def hi_bytes(pairs):
        key = tuple(sorted(pairs))
        if key not in hi_cache:
            act = np.zeros(his.size, bool)
            for a, b in key:
                act |= ((his >> (a - Lq)) & 1) != ((his >> (b - Lq)) & 1)
            hi_cache[key] = int(seg[act].sum()) * 2 * 16
        return hi_cache[key]

    kinds = [int(o["kind"]) for o in ops]
    qs = [(int(o["q0"]), int(o["q1"])) for o in ops]
    bond = lambda i: kinds[i] in (L.NP_ROT_X, L.NP_ROT_Y)
    tot, i = 0, 0
    while i < len(ops):
        if kinds[i] == L.PHASE:
            while i < len(ops) and kinds[i] == L.PHASE:
                i += 1
            tot += 2 * namps * 16
            continue
        assert bond(i), "np_pass_bytes counts phases and npRotateX/Y bonds only"
        j, used = i, set()
        while j < len(ops) and bond(j) and abs(qs[j][0] - qs[j][1]) == 1 \
                and max(qs[j]) < low_bits and not used & set(qs[j]):
            used |= set(qs[j])
            j += 1
        if j - i >= 2:
            tot += 2 * namps * 16
            i = j
            continue
        j, used = i, set()
        while j < len(ops) and j - i < hi_max and bond(j) and min(qs[j]) >= Lq \
                and not used & set(qs[j]):
            used |= set(qs[j])
            j += 1
        if j - i >= 2:
            tot += hi_bytes(qs[i:j])
            i = j
            continue
        tot += 2 * 2 * math.comb(n - 2, k - 1) * 16
        i += 1
    return tot


def bench_c3(layers=100):
    """Config 3: number-preserving 32 qubits, 16 ones, FP64 compact
    (9.62 GB): npRotateX/Y on alternating bonds + phase on every qubit."""
    import qsl_amd
    from qsl_amd import workloads as W
    from qsl_amd import _lib as L
    n, k = 32, 16
    ops, r0, _ = W.c3_np_layers(n, layers)
    q = qsl_amd.QubitsNP(n, k)
    x0 = W.colex_unrank(r0, k)
    q.setBasisState(x0)
    ev = HipEvents()
    st = stream_of(q)
    e0, e1 = ev.event(), ev.event()
    q.apply_ops(ops[: 47])  # warm
    q.setBasisState(x0)
    q.sync()
    ev.record(e0, st)
    q.apply_ops(ops)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    namps = math.comb(n, k)
    nb = sum(1 for o in ops if int(o["kind"]) in (L.NP_ROT_X, L.NP_ROT_Y))
    nph = len(ops) - nb
    # bytes one launch per gate would move (bonds read and write both
    # amplitudes of C(n-2, k-1) pairs, a phase the C(n-1, k-1) strings with
    # its bit set)
    alg = (nb * 2 * 2 * math.comb(n - 2, k - 1) + nph * 2 * math.comb(n - 1, k - 1)) * 16
    kern = np_pass_bytes(ops, n, k)
    norm = q.norm()
    return {"layers": layers, "gates": len(ops), "ms": round(ms, 2),
            "gates_per_s": round(len(ops) / (ms * 1e-3), 1),
            "per_gate_algorithmic_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
            "kernel_hbm_GBps": round(kern / (ms * 1e-3) / 1e9, 1),
            "kernel_hbm_note": "bytes of the passes launched for this circuit (a run of "
                               "phases, the low bonds of a layer and its high bonds one pass "
                               "each, the bonds between qubits 12 and 16 one each) over the "
                               "time; per_gate_algorithmic counts every gate as its own pass, "
                               "so it can exceed the HBM peak",
            "amplitudes": namps, "state_GB": round(namps * 16 / 1e9, 2),
            "norm_err": abs(norm - 1)}


def bench_c4(steps=3, plain_sample=64):
    """Config 4: 4096 variational circuits (10-20 qubits), energies and
    adjoint gradients: batched FP32 (one hipGraph-captured persistent
    kernel) against the plain per-circuit FP64 GPU path on a stratified
    sample of the same circuits."""
    import qsl_amd
    from qsl_amd import workloads as W
    circuits = W.c4_batch(4096)
    t0 = time.perf_counter()
    b = qsl_amd.Batch(circuits)
    create_s = time.perf_counter() - t0
    e, g = b.run()
    b_stream = b.stream()
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, b_stream)
    for _ in range(steps):
        e, g = b.run()
    ev.record(e1, b_stream)
    wall = (time.perf_counter() - t0) / steps
    gpu_ms = ev.elapsed_ms(e0, e1) / steps
    nparams = int(sum(16 * c[0] for c in circuits))
    # plain FP64 per-circuit path on every (4096 / plain_sample)-th circuit
    idx = list(range(0, len(circuits), len(circuits) // plain_sample))
    off = np.concatenate([[0], np.cumsum([len(c[1]) for c in circuits])])
    maxerr, maxe = 0.0, 0.0
    t0 = time.perf_counter()
    for i in idx:
        n, ops, terms = circuits[i]
        e64, g64 = qsl_amd.Qubits(n).expvalGrad(ops, terms)
        maxerr = max(maxerr, float(np.abs(g[off[i]:off[i + 1]] - g64).max()))
        maxe = max(maxe, abs(float(e[i]) - e64))
    plain_s = (time.perf_counter() - t0) / len(idx)
    batch_gps = len(circuits) / wall
    plain_gps = 1.0 / plain_s
    return {"circuits": len(circuits), "params": nparams, "batch_create_s": round(create_s, 2),
            "batch_ms_per_run": round(1000 * wall, 2), "batch_gpu_ms": round(gpu_ms, 2),
            "gradients_per_s_fp32_batched": round(batch_gps, 1),
            "params_per_s_fp32_batched": round(nparams / wall, 1),
            "gradients_per_s_fp64_plain": round(plain_gps, 2),
            "speedup_vs_fp64_plain": round(batch_gps / plain_gps, 1),
            "plain_sample": f"{len(idx)} circuits, every {len(circuits) // plain_sample}th, "
                            "one qslhip_expval_grad call each (FP64, one kernel per gate)",
            "max_grad_err_vs_fp64": maxerr, "max_energy_err_vs_fp64": maxe}


def bench_c5(layers34=30, layers30=30):
    """Config 5: 34-qubit FP32 (128 GiB) Haar layers fused into dense
    4- or 5-qubit unitaries (tuning dense_f) on matrix cores, the other
    width timed beside it; error of the same family at 30 qubits against
    the FP64 path (bit-identical to the oracle)."""
    import qsl_amd
    from qsl_amd import workloads as W
    from qsl_amd import _lib as L
    F = L.tuning_get("dense_f")
    res = {"fusion_width": F}
    n = 34
    ops, _ = W.c5_haar_layers(n, layers34)
    q = qsl_amd.Qubits(n, dtype=np.complex64)
    ev = HipEvents()
    st = stream_of(q)
    e0, e1 = ev.event(), ev.event()
    q.apply_ops(ops[: 3 * n], mode=qsl_amd.MODE_DENSE)  # warm
    q.reset()
    q.sync()
    ev.record(e0, st)
    q.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    res.update({"nqubits": n, "layers": layers34, "gates": len(ops), "ms": round(ms, 1),
                "gates_per_s": round(len(ops) / (ms * 1e-3), 1),
                "norm_err": abs(q.norm() - 1)})
    # matrix-core work of the plan: each dense unitary is a real GEMM with
    # the interleaved embedding, 2^F x 2^F complex = 2^(F+1) x 2^(F+1) real,
    # i.e. 2 * 2^(F+1) * 2 = 8 << F flops per amplitude (v_mfma_f32_32x32x2f32)
    _, oi, _ = L.plan_dense(n, ops, F)
    ndense = int((oi[:, 0] == 1).sum())
    flops = ndense * (8 << F) * float(1 << n)
    tflops = flops / (ms * 1e-3) / 1e12
    res["mfma"] = {"dense_unitaries": ndense, "single_gate_ops": int((oi[:, 0] == 0).sum()),
                   "flops": flops, "achieved_TFLOPs": round(tflops, 1),
                   "peak_TFLOPs_fp32": MFMA_FP32_PEAK_TFLOPS,
                   "utilisation": round(tflops / MFMA_FP32_PEAK_TFLOPS, 3),
                   "note": "plan flops over the whole dense application (loads, stores and "
                           "single-gate tile ops included in the time)"}
    # the other fusion width on the same circuit, for comparison
    F2 = 9 - F
    L.tuning_set("dense_f", F2)
    try:
        q.reset()
        q.sync()
        ev.record(e0, st)
        q.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
        ev.record(e1, st)
        ms2 = ev.elapsed_ms(e0, e1)
        _, oi2, _ = L.plan_dense(n, ops, F2)
        nd2 = int((oi2[:, 0] == 1).sum())
        tf2 = nd2 * (8 << F2) * float(1 << n) / (ms2 * 1e-3) / 1e12
        res["other_fusion_width"] = {
            "fusion_width": F2, "ms": round(ms2, 1),
            "gates_per_s": round(len(ops) / (ms2 * 1e-3), 1), "dense_unitaries": nd2,
            "achieved_TFLOPs": round(tf2, 1),
            "utilisation": round(tf2 / MFMA_FP32_PEAK_TFLOPS, 3)}
    finally:
        L.tuning_set("dense_f", F)
    # the first 6 layers on FP32 register-tile passes, reordered (every pass
    # is compiled at run time, so the comparison stays at 6 layers): the
    # first two applications interpret (the second queues the structures
    # for the compiler thread), the third is timed compiled
    ops = ops[: 6 * len(ops) // layers34]
    times = []
    for i in range(3):
        if i == 2:
            L.jit_wait()
        q.reset()
        q.sync()
        ev.record(e0, st)
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
        ev.record(e1, st)
        times.append(ev.elapsed_ms(e0, e1))
    npass = len(L.plan_order(n, 4, ops, L.MODE_REORDER, cap=1 << 16)[0])
    norm_err = abs(q.norm() - 1)
    for _ in range(2):  # on the dense result: every group; interpret, then queue
        q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    L.jit_wait()
    q.sync()
    ev.record(e0, st)
    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    ev.record(e1, st)
    dense_ms = ev.elapsed_ms(e0, e1)
    res["register_tile_fp32"] = {
        "mode": "reorder", "layers": 6, "gates": len(ops), "passes": npass,
        "ms": round(times[2], 1),
        "gates_per_s": round(len(ops) / (times[2] * 1e-3), 1),
        "interpreted_ms": round(times[0], 1), "norm_err": norm_err,
        "from_reset": "the passes skip what is still zero",
        "dense_ms": round(dense_ms, 1),
        "dense_gates_per_s": round(len(ops) / (dense_ms * 1e-3), 1),
        "dense_hbm_GBps": round(npass * 2 * (8 << n) / (dense_ms * 1e-3) / 1e9, 1)}
    del q
    n = 30
    ops, _ = W.c5_haar_layers(n, layers30)
    # expectation value of H = sum_i Z_i Z_{i+1} + 0.5 sum_i X_i (FP64
    # accumulation on the device for every precision)
    terms = np.array([(1.0, 0, 3 << i) for i in range(n - 1)] +
                     [(0.5, 1 << i, 0) for i in range(n)], L.PAULI_DTYPE)
    q32 = qsl_amd.Qubits(n, dtype=np.complex64)
    q32.apply_ops(ops, mode=qsl_amd.MODE_DENSE)
    e32 = q32.expval(terms)
    s32 = q32.getState()
    q32.reset()
    q32.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    e32r = q32.expval(terms)
    s32r = q32.getState()
    del q32
    q64 = qsl_amd.Qubits(n)
    q64.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    e64 = q64.expval(terms)
    s64 = q64.getState()
    del q64
    res["max_amp_err_30q_vs_fp64"] = float(np.abs(s32.astype(np.complex128) - s64).max())
    res["register_tile_fp32"]["max_amp_err_30q_vs_fp64"] = float(
        np.abs(s32r.astype(np.complex128) - s64).max())
    res["max_amp_30q"] = float(np.abs(s64).max())
    res["expval_30q"] = {
        "observable": "sum_i Z_i Z_(i+1) + 0.5 sum_i X_i", "fp64": e64,
        "dense_mfma_fp32_err": abs(e32 - e64), "register_tile_fp32_err": abs(e32r - e64)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["reorder", "fused", "per-gate"], default="reorder")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short fixed run for rocprofv3 counter passes")
    ap.add_argument("--profile-c4", action="store_true",
                    help="short fixed run of the 4096-circuit batched adjoint for rocprofv3")
    ap.add_argument("--profile-c3", action="store_true",
                    help="short fixed run of the number-preserving path (32 qubits, weight 16, "
                         "FP64, 10 layers) for rocprofv3")
    ap.add_argument("--profile-c5", action="store_true",
                    help="short fixed run of the dense MFMA path (34 qubits FP32, 3 Haar layers) "
                         "for rocprofv3")
    ap.add_argument("--dense-f", type=int, choices=(4, 5), default=None,
                    help="qubits per fused dense unitary for --profile-c5 (default: the "
                         "library's tuning, 4)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline 30-qubit workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus "
              f"{args.gpus}); multi-GPU sharding is a design note, not built.",
              file=sys.stderr)
        sys.exit(2)

    import qsl_amd
    from qsl_amd import workloads as W
    from qsl_amd import _lib as L

    if args.profile_c3:
        # the compact number-preserving state: 10 of the 100 layers, from the
        # same start basis string
        n, k = 32, 16
        ops3, r0, _ = W.c3_np_layers(n, 10)
        q = qsl_amd.QubitsNP(n, k)
        q.setBasisState(W.colex_unrank(r0, k))
        q.sync()
        print(f"profile: NP {n}q weight {k}, {len(ops3)} gates", flush=True)
        q.apply_ops(ops3)
        q.sync()
        print("profile: done", flush=True)
        return
    if args.profile_c5:
        # the dense fused-unitary path on matrix cores: the first 3 of the
        # 30 Haar layers at 34 qubits, applied once from |0...0>
        if args.dense_f is not None:
            L.tuning_set("dense_f", args.dense_f)
        F = L.tuning_get("dense_f")
        ops5, _ = W.c5_haar_layers(34, 3)
        _, oi, _ = L.plan_dense(34, ops5, F)
        nd = int((oi[:, 0] == 1).sum())
        q = qsl_amd.Qubits(34, dtype=np.complex64)
        print(f"profile: dense application, {len(ops5)} gates, {nd} dense unitaries, "
              f"{nd * (8 << F) * float(1 << 34):.4g} plan flops", flush=True)
        q.apply_ops(ops5, mode=qsl_amd.MODE_DENSE)
        q.sync()
        print("profile: done", flush=True)
        return
    if args.profile_c4:
        # the batched variational workload: build once, then three runs of
        # the captured graph (one persistent k_batch launch each)
        b = qsl_amd.Batch(W.c4_batch(4096))
        for rep in (1, 2, 3):
            print(f"profile: batch run ({rep})", flush=True)
            b.run()
        print("profile: done", flush=True)
        return
    mode = {"reorder": qsl_amd.MODE_REORDER, "fused": qsl_amd.MODE_FUSED,
            "per-gate": qsl_amd.MODE_PER_GATE}[args.mode]
    ops, stream = W.c2_random(N, NGATES)
    if args.profile:
        # a short fixed run for rocprofv3 counter passes: one sampling, the
        # circuit three times reordered and three times in program order (on
        # the dense state the sampling wrote), rotateX on a low and a high
        # qubit, and one
        # device-to-device copy of the 16 GiB state (2^30*16 B read and
        # written: the calibration for FETCH_SIZE / WRITE_SIZE)
        q = qsl_amd.Qubits(N, dtype=np.complex128)
        # one sampling first: the handle then holds the sampling buffer, so
        # the last pass of each application writes its run sums, as in the
        # timed steps
        q.sampleAllOutcomes(np.array([0.5]))
        # the first two applications interpret (the second queues each pass
        # structure for the compiler thread), the third runs the compiled
        # passes back to back
        for rep in (1, 2, 3):
            if rep == 3:
                L.jit_wait()
            print(f"profile: reordered circuit ({rep})", flush=True)
            q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
            q.sync()
        for rep in (1, 2, 3):
            if rep == 3:
                L.jit_wait()
            print(f"profile: program-order circuit ({rep})", flush=True)
            q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
            q.sync()
        print("profile: rotateX t=5 x3, t=29 x3", flush=True)
        for t in (5, 29):
            for _ in range(3):
                q.rotateX(t, 0.3)
        q.sync()
        print("profile: state copy", flush=True)
        c = q.copy()
        c.sync()
        del c
        print("profile: done", flush=True)
        return
    shots_u = stream.take(NSHOTS)
    reorder_masks = L.plan_order(N, 8, ops, L.MODE_REORDER)[0]
    npasses = len(reorder_masks)
    npasses_fused = len(L.plan_order(N, 8, ops, L.MODE_FUSED)[0])
    hl_mode = mode if mode != qsl_amd.MODE_PER_GATE else qsl_amd.MODE_REORDER
    hl_passes = npasses if hl_mode == qsl_amd.MODE_REORDER else npasses_fused

    q = qsl_amd.Qubits(N, dtype=np.complex128)
    ev = HipEvents()
    s = stream_of(q)

    def step():
        q.reset()
        q.apply_ops(ops, mode=mode)
        return q.sampleAllOutcomes(shots_u)

    warm_ms = []
    jit0 = L.tuning_get("jit_compiled")
    for _ in range(args.warmup):
        tw = time.perf_counter()
        out = step()
        q.sync()
        warm_ms.append(round(1000 * (time.perf_counter() - tw), 1))
    # the second warmup queued the repeating pass structures for the
    # library's compiler thread; the timed steps run them compiled
    L.jit_wait()
    jit_passes = L.tuning_get("jit_compiled") - jit0
    q.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    q.sync()
    t1 = time.perf_counter()
    ms_per_step = 1000 * (t1 - t0) / args.steps
    value = NGATES * args.steps / (t1 - t0)

    # --- phase breakdown with HIP events on the library stream -------------
    e0, e1, e2 = ev.event(), ev.event(), ev.event()
    q.reset()
    q.sync()
    ev.record(e0, s)
    q.apply_ops(ops, mode=hl_mode)
    ev.record(e1, s)
    q.sync()
    gate_ms = ev.elapsed_ms(e0, e1)
    # program order: the first application interprets the decoded ops, the
    # second interprets and queues the passes for compilation, the third runs
    # them compiled
    fused_runs = []
    for i in range(3):
        if i == 2:
            L.jit_wait()
        q.reset()
        q.sync()
        ev.record(e0, s)
        q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
        ev.record(e1, s)
        q.sync()
        fused_runs.append(ev.elapsed_ms(e0, e1))
    fused_interp_ms, fused_ms = fused_runs[0], fused_runs[2]
    q.reset()
    q.apply_ops(ops, mode=hl_mode)
    q.sync()
    state_bytes = (1 << N) * 16
    pass_bytes = 2 * state_bytes
    # roofline of the pass kernel: applications on the current state, so
    # that every pass reads and writes it (from a reset the passes skip what
    # is still zero); the last pass also writes the 2^30 B of run sums.  The
    # plan for a state that is not a basis state spreads the diagonal gates
    # over the passes, so it is its own structure: the first application
    # interprets, the second compiles.  The next ones let each pass try its
    # four launch shapes and keep the fastest (tuning pass_tune).  The timed
    # region is three compiled applications queued behind an untimed one,
    # so the host's planning of each runs while the device works and the
    # passes follow each other without gaps: steady_ms is one application's
    # share
    for _ in range(2):
        q.apply_ops(ops, mode=hl_mode)
    L.jit_wait()
    q.sync()
    for _ in range(5):
        q.apply_ops(ops, mode=hl_mode)
        q.sync()
    ev.record(e0, s)
    for _ in range(3):
        q.apply_ops(ops, mode=hl_mode)
    ev.record(e1, s)
    q.sync()
    steady_ms = ev.elapsed_ms(e0, e1) / 3
    tile_avg_ms = steady_ms / hl_passes
    steady_bytes = hl_passes * pass_bytes + (1 << N)
    t_s0 = time.perf_counter()
    q.sampleAllOutcomes(shots_u)
    sample_ms = 1000 * (time.perf_counter() - t_s0)

    # norm is preserved to FP64 rounding (size-independent check)
    norm_err = abs(q.norm() - 1.0)

    # --- the per-gate streaming kernel: dense 1-qubit FP64 gate -------------
    per_target = {}
    reps = 10
    for t in (0, 1, 4, 8, 12, 20, 29):
        q.sync()
        ev.record(e0, s)
        for _ in range(reps):
            q.rotateX(t, 0.3)
        ev.record(e2, s)
        ms = ev.elapsed_ms(e0, e2) / reps
        per_target[str(t)] = round(pass_bytes / (ms * 1e-3) / 1e9, 1)
    ms_pg = []
    q.sync()
    ev.record(e0, s)
    q.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    ev.record(e2, s)
    per_gate_ms = ev.elapsed_ms(e0, e2)
    alg_bytes = W.algorithmic_bytes(ops, N, 16)

    # --- full-size checks: fused bit-identical to per-gate (program order);
    # the reordered state within 1e-12 of it on every amplitude ------------
    q.reset()
    q.apply_ops(ops, mode=qsl_amd.MODE_PER_GATE)
    ref_state = q.getState()
    q.reset()
    q.apply_ops(ops, mode=qsl_amd.MODE_FUSED)
    fused_state = q.getState()
    fused_bit_identical = bool(np.array_equal(ref_state.view(np.uint8),
                                              fused_state.view(np.uint8)))
    del fused_state
    q.reset()
    q.apply_ops(ops, mode=qsl_amd.MODE_REORDER)
    reorder_state = q.getState()
    reorder_diff = 0.0
    for i in range(0, ref_state.size, 1 << 26):  # chunked: no 16 GiB temporaries
        j = i + (1 << 26)
        reorder_diff = max(reorder_diff, float(np.abs(reorder_state[i:j] - ref_state[i:j]).max()))
    del ref_state, reorder_state

    achieved = steady_bytes / (steady_ms * 1e-3) / 1e9  # per launch: both / passes
    step_bytes = step_pass_bytes(reorder_masks if hl_mode == qsl_amd.MODE_REORDER else
                                 L.plan_order(N, 8, ops, L.MODE_FUSED)[0], N)
    traffic = read_traffic()
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "30-qubit FP64 state (16 GiB) from |0..0>: 200 seeded random "
                        "gates (40% RX/RY/RZ, 20% H, 20% CNOT, 20% CPhase, qubits "
                        "uniform over 0-29), then 1e6 shots; step = reset + gates + "
                        "sampling",
            "nqubits": N, "gates": NGATES, "shots": NSHOTS, "mode": args.mode,
            "seed": W.MASTER_SEED + 2, "tile_passes": hl_passes,
            "tile_qubits": L.tuning_get("logt_f64"), "low_run": L.tuning_get("low_f64"),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "kernel": ("qslhip_pass (register tile pass compiled at run time; "
                       if jit_passes else "k_rtile<double,256,8> (register tile pass; ")
                      + "unit = one pass, 2*2^30*16 B per launch, the last one also "
                      "2^30 B of sampling run sums)",
            "avg_launch_ms": round(tile_avg_ms, 4),
            "gates_per_launch": round(NGATES / hl_passes, 1),
            "note": "measured HBM bytes per pass equal the algorithmic bytes (traffic); "
                    "each pass runs within ~5% of a gate-free copy with the same "
                    "addressing, whose rate (5.9-8.7 ms per 16 GiB in place) depends on "
                    "the tile's qubits",
        },
        "jit": {
            "mode": L.tuning_get("jit"),
            "passes_compiled_in_warmup": jit_passes,
            "warmup_step_ms": warm_ms,
            "failed": L.tuning_get("jit_failed"),
        },
        "gate_phase": {
            "mode": args.mode,
            "gate_ms": round(gate_ms, 3),
            "gates_per_s": round(NGATES / (gate_ms * 1e-3), 1),
            "from_reset": "from |0..0> each pass runs only the groups that can be nonzero "
                          "and reads only what an earlier pass wrote",
            "steady_state_ms": round(steady_ms, 3),
            "dense_gates_per_s": round(NGATES / (steady_ms * 1e-3), 1),
            "algorithmic_GBps": round(W.algorithmic_bytes(ops, N, 16) / (gate_ms * 1e-3) / 1e9, 1),
            "reorder_max_abs_diff_vs_program_order": reorder_diff,
            "reorder_within_1e-12": reorder_diff <= 1e-12,
            "fused_passes": npasses_fused,
            "fused_ms": round(fused_ms, 3),
            "fused_gates_per_s": round(NGATES / (fused_ms * 1e-3), 1),
            "fused_interpreted_ms": round(fused_interp_ms, 3),
            "per_gate_ms": round(per_gate_ms, 3),
            "per_gate_gates_per_s": round(NGATES / (per_gate_ms * 1e-3), 1),
            "algorithmic_GB": round(alg_bytes / 1e9, 2),
            "per_gate_hbm_GBps": round(alg_bytes / (per_gate_ms * 1e-3) / 1e9, 1),
            "fused_hbm_GBps_actual": round(npasses_fused * pass_bytes / (fused_ms * 1e-3) / 1e9, 1),
            "fused_bit_identical_to_per_gate": fused_bit_identical,
            "norm_err": norm_err,
        },
        # the same 200 gates on a dense state (every pass reads and writes
        # all of it), and the HBM rate of the timed step's gate phase counted
        # from the bytes its passes actually move from |0..0>
        "dense_gates_per_s": round(NGATES / (steady_ms * 1e-3), 1),
        "step_gate_phase_hbm": {
            "bytes": step_bytes, "ms": round(gate_ms, 3),
            "GBps": round(step_bytes / (gate_ms * 1e-3) / 1e9, 1),
            "frac": round(step_bytes / (gate_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "rule": "reordered passes from a basis state run only the groups that "
                    "can be nonzero and read only what an earlier pass wrote",
        },
        "rotateX_f64_GBps_by_target": per_target,
        "sampling": {"shots": NSHOTS, "ms": round(sample_ms, 3),
                     "shots_per_s": round(NSHOTS / (sample_ms * 1e-3), 1)},
    }
    if not args.no_cpu:
        def gpu_state(prefix):
            q.reset()
            q.apply_ops(prefix, mode=qsl_amd.MODE_PER_GATE)
            return q.getState()
        result["cpu_baseline"] = cpu_baseline(ops, gpu_state)
    else:
        result["cpu_baseline"] = None
    del q
    if not args.no_extras:
        for name, fn in (("c1_16q_fp64", bench_c1), ("c4_batched_adjoint", bench_c4),
                         ("c3_np_32q_fp64", bench_c3), ("c5_dense_mfma", bench_c5)):
            t0 = time.perf_counter()
            try:
                result[name] = fn()
            except Exception as e:  # report, keep the headline line intact
                result[name] = {"error": f"{type(e).__name__}: {e}"}
            result[name]["wall_s"] = round(time.perf_counter() - t0, 2)
            print(f"bench: {name} done in {result[name]['wall_s']} s", file=sys.stderr,
                  flush=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
