// autotune.cpp -- picks tile and fusion widths by timing them on the device.
//
// Candidates are timed with hipEvents on a scratch state (at most 2^26
// amplitudes) running a seeded random circuit of the gate mix the tuned
// path serves:
//   * LDS tile passes (fused mode): tile width logt and contiguous low run
//     per precision, subject to logt - low <= 8 and a 64 KiB tile;
//   * per-gate mode: the highest target that still goes through a one-op
//     LDS tile instead of the streaming kernel;
//   * dense MFMA mode (FP32): fusion width F in {4, 5}.
// The winners are stored in the library's tuning table (qslhip_tuning_get
// reads it back); the report lists every timing.
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "internal.h"

using namespace qslhip;

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        QH_CHECK(hipEventCreate(&a));
        QH_CHECK(hipEventCreate(&b));
    }
    ~Timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

std::vector<qslhip_op> mix_circuit(unsigned n, int count, uint64_t seed, bool haar)
{
    std::mt19937_64 g(seed);
    auto u = [&] { return (double)(g() >> 11) * 0x1.0p-53; };
    std::vector<qslhip_op> ops;
    for (int i = 0; i < count; i++) {
        qslhip_op o{};
        double r = u();
        unsigned a = (unsigned)(u() * n) % n, b = (unsigned)(u() * (n - 1)) % (n - 1);
        b += b >= a;
        if (haar) {  // rotations on every qubit plus controlPhase pairs
            o.kind = r < 0.85 ? (r < 0.55 ? QSLHIP_ROT_Z : QSLHIP_ROT_Y) : QSLHIP_CPHASE;
        } else {
            o.kind = r < 0.4 ? QSLHIP_ROT_X + (int)(u() * 3) % 3
                   : r < 0.6 ? QSLHIP_HADAMARD
                   : r < 0.8 ? QSLHIP_CNOT : QSLHIP_CPHASE;
        }
        o.q0 = a;
        o.q1 = is_gate2(o.kind) ? b : 0;
        o.angle = 6.283185307179586 * u();
        ops.push_back(o);
    }
    return ops;
}

float time_apply(qslhip_t h, const std::vector<qslhip_op> &ops, int mode, Timer &t)
{
    if (qslhip_apply_ops(h, ops.data(), ops.size(), mode) != QSLHIP_OK)
        throw Error(QSLHIP_EDEVICE, qslhip_last_error());
    void *s = nullptr;
    qslhip_get_stream(h, &s);
    QH_CHECK(hipEventRecord(t.a, (hipStream_t)s));
    if (qslhip_apply_ops(h, ops.data(), ops.size(), mode) != QSLHIP_OK)
        throw Error(QSLHIP_EDEVICE, qslhip_last_error());
    QH_CHECK(hipEventRecord(t.b, (hipStream_t)s));
    QH_CHECK(hipEventSynchronize(t.b));
    float ms = 0;
    QH_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
    return ms;
}

// A tile width the fused path can run: the LDS kernel keeps the tile and
// an offset table within 64 KiB; register tiles (10..12 tile qubits) use
// LDS only as the 2^logt-amplitude remap buffer.
bool tile_fits(const Tuning &t, int logt, int low, int ab)
{
    if (logt < 6 || logt > 13 || low < 1 || low > logt - 2 || logt - low > 8) return false;
    const int prec = ab == 16 ? 8 : 4;
    if (t.reg_tile && reg_tile_ok(logt, prec == 8 ? t.rthreads_f64 : t.rthreads_f32, prec))
        return true;
    return ((size_t)ab << logt) <= 64 * 1024 - 2048;
}

}  // namespace

extern "C" {

int qslhip_tuning_get(const char *key, int *value)
{
    return guarded_call([&] {
        QH_REQUIRE(key && value, QSLHIP_EINVAL, "null argument");
        Tuning &t = tuning();
        std::string k = key;
        if (k == "logt_f64") *value = t.logt_f64;
        else if (k == "low_f64") *value = t.low_f64;
        else if (k == "logt_f32") *value = t.logt_f32;
        else if (k == "low_f32") *value = t.low_f32;
        else if (k == "lds_max_target_f64") *value = t.lds_max_target_f64;
        else if (k == "lds_max_target_f32") *value = t.lds_max_target_f32;
        else if (k == "dense_f") *value = t.dense_f;
        else if (k == "reg_tile") *value = t.reg_tile;
        else if (k == "rthreads_f64") *value = t.rthreads_f64;
        else if (k == "rthreads_f32") *value = t.rthreads_f32;
        else if (k == "rgroups") *value = t.rgroups;
        else if (k == "reg_min_qubits") *value = t.reg_min_qubits;
        else if (k == "jit") *value = t.jit;
        else if (k == "jit_waves") *value = t.jit_waves;
        else if (k == "stream_min_qubits") *value = t.stream_min_qubits;
        else if (k == "lane_gates") *value = t.lane_gates;
        else if (k == "np_adj_max") *value = t.np_adj_max;
        else if (k == "np_adj_policy") *value = t.np_adj_policy;
        else if (k == "np_stream_min_q") *value = t.np_stream_min_q;
        else if (k == "np_bonds") *value = t.np_bonds;
        else if (k == "np_lowbonds") *value = t.np_lowbonds;
        else if (k == "pass_stream") *value = t.pass_stream;
        else if (k == "np_hibonds") *value = t.np_hibonds;
        else if (k == "np_phase_lut") *value = t.np_phase_lut;
        else if (k == "free_layout") *value = t.free_layout;
        else if (k == "dense_reorder") *value = t.dense_reorder;
        else if (k == "jit_prefetch") *value = t.jit_prefetch;
        else if (k == "pass_tune") *value = t.pass_tune;
        else if (k == "pass_tune_pending") *value = pass_tune_pending();
        else if (k == "group_order") *value = t.group_order;
        else if (k == "jit_compiled") *value = g_jit_compiled;
        else if (k == "jit_failed") *value = g_jit_failed;
        else if (k == "run_sums_used") *value = g_run_sums_used;
        else throw Error(QSLHIP_EINVAL, "unknown tuning key " + k);
    });
}

int qslhip_tuning_set(const char *key, int value)
{
    return guarded_call([&] {
        QH_REQUIRE(key, QSLHIP_EINVAL, "null key");
        Tuning &t = tuning();
        std::string k = key;
        auto tile_ok = [&](int logt, int low, int ab) { return tile_fits(t, logt, low, ab); };
        if (k == "logt_f64") {
            QH_REQUIRE(tile_ok(value, t.low_f64, 16), QSLHIP_EINVAL, "bad logt_f64");
            t.logt_f64 = value;
        } else if (k == "low_f64") {
            QH_REQUIRE(tile_ok(t.logt_f64, value, 16), QSLHIP_EINVAL, "bad low_f64");
            t.low_f64 = value;
        } else if (k == "logt_f32") {
            QH_REQUIRE(tile_ok(value, t.low_f32, 8), QSLHIP_EINVAL, "bad logt_f32");
            t.logt_f32 = value;
        } else if (k == "low_f32") {
            QH_REQUIRE(tile_ok(t.logt_f32, value, 8), QSLHIP_EINVAL, "bad low_f32");
            t.low_f32 = value;
        } else if (k == "lds_max_target_f64") {
            QH_REQUIRE(value >= -1 && value < t.logt_f64, QSLHIP_EINVAL, "bad threshold");
            t.lds_max_target_f64 = value;
        } else if (k == "lds_max_target_f32") {
            QH_REQUIRE(value >= -1 && value < t.logt_f32, QSLHIP_EINVAL, "bad threshold");
            t.lds_max_target_f32 = value;
        } else if (k == "dense_f") {
            QH_REQUIRE(value == 4 || value == 5, QSLHIP_EINVAL, "dense_f must be 4 or 5");
            t.dense_f = value;
        } else if (k == "reg_tile") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "reg_tile must be 0 or 1");
            Tuning u = t;
            u.reg_tile = value;
            QH_REQUIRE(tile_fits(u, u.logt_f64, u.low_f64, 16) &&
                           tile_fits(u, u.logt_f32, u.low_f32, 8),
                       QSLHIP_EINVAL, "current tile widths need register tiles");
            t.reg_tile = value;
        } else if (k == "jit") {
            QH_REQUIRE(value >= 0 && value <= 2, QSLHIP_EINVAL, "jit must be 0, 1 or 2");
            t.jit = value;
        } else if (k == "pass_stream") {
            QH_REQUIRE(value >= 0 && value <= 3, QSLHIP_EINVAL, "pass_stream must be in [0, 3]");
            t.pass_stream = value;
        } else if (k == "np_hibonds") {
            QH_REQUIRE(value >= 0 && value <= 2, QSLHIP_EINVAL, "np_hibonds must be 0, 1 or 2");
            t.np_hibonds = value;
        } else if (k == "group_order") {
            QH_REQUIRE(value >= 0 && value <= 2, QSLHIP_EINVAL, "group_order must be 0, 1 or 2");
            t.group_order = value;
        } else if (k == "pass_tune") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "pass_tune must be 0 or 1");
            t.pass_tune = value;
        } else if (k == "jit_prefetch") {
            QH_REQUIRE(value >= 0 && value <= 2, QSLHIP_EINVAL, "jit_prefetch must be 0, 1 or 2");
            t.jit_prefetch = value;
        } else if (k == "dense_reorder") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "dense_reorder must be 0 or 1");
            t.dense_reorder = value;
        } else if (k == "free_layout") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "free_layout must be 0 or 1");
            t.free_layout = value;
        } else if (k == "np_phase_lut") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "np_phase_lut must be 0 or 1");
            t.np_phase_lut = value;
        } else if (k == "np_lowbonds") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "np_lowbonds must be 0 or 1");
            t.np_lowbonds = value;
        } else if (k == "np_bonds") {
            QH_REQUIRE(value >= 1 && value <= 4, QSLHIP_EINVAL, "np_bonds must be in [1, 4]");
            t.np_bonds = value;
        } else if (k == "np_stream_min_q") {
            QH_REQUIRE(value >= 0 && value <= 64, QSLHIP_EINVAL,
                       "np_stream_min_q must be in [0, 64]");
            t.np_stream_min_q = value;
        } else if (k == "np_adj_policy") {
            QH_REQUIRE(value >= 0 && value <= 3, QSLHIP_EINVAL, "np_adj_policy must be in [0, 3]");
            t.np_adj_policy = value;
        } else if (k == "np_adj_max") {
            QH_REQUIRE(value >= -1 && value <= 62, QSLHIP_EINVAL, "np_adj_max must be in [-1, 62]");
            t.np_adj_max = value;
        } else if (k == "lane_gates") {
            QH_REQUIRE(value == 0 || value == 1, QSLHIP_EINVAL, "lane_gates must be 0 or 1");
            t.lane_gates = value;
        } else if (k == "stream_min_qubits") {
            QH_REQUIRE(value >= 0 && value <= 64, QSLHIP_EINVAL,
                       "stream_min_qubits must be in [0, 64]");
            t.stream_min_qubits = value;
        } else if (k == "jit_waves") {
            QH_REQUIRE(value >= 0 && value <= 8, QSLHIP_EINVAL, "jit_waves must be in [0, 8]");
            t.jit_waves = value;
        } else if (k == "reg_min_qubits") {
            QH_REQUIRE(value >= 0 && value <= (int)kMaxQubits, QSLHIP_EINVAL,
                       "reg_min_qubits must be in [0, 40]");
            t.reg_min_qubits = value;
        } else if (k == "rgroups") {
            QH_REQUIRE(value >= 0 && value <= 1024, QSLHIP_EINVAL, "rgroups must be in [0, 1024]");
            t.rgroups = value;
        } else if (k == "rthreads_f64" || k == "rthreads_f32") {
            QH_REQUIRE(value == 256 || value == 512, QSLHIP_EINVAL, "threads must be 256 or 512");
            Tuning u = t;
            (k == "rthreads_f64" ? u.rthreads_f64 : u.rthreads_f32) = value;
            QH_REQUIRE(tile_fits(u, u.logt_f64, u.low_f64, 16) &&
                           tile_fits(u, u.logt_f32, u.low_f32, 8),
                       QSLHIP_EINVAL, "current tile widths need the other workgroup size");
            t = u;
        } else {
            throw Error(QSLHIP_EINVAL, "unknown tuning key " + k);
        }
    });
}

int qslhip_autotune(int prec, unsigned nqubits, char *report, size_t cap)
{
    return guarded_call([&] {
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(nqubits >= 14 && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "autotune needs 14..40 qubits");
        const unsigned n = nqubits < 26 ? nqubits : 26;
        qslhip_t h = nullptr;
        if (qslhip_create(n, prec, &h) != QSLHIP_OK)
            throw Error(QSLHIP_ENOMEM, qslhip_last_error());
        Tuning &t = tuning();
        const Tuning saved = t;
        std::string rep = "{\"nqubits\": " + std::to_string(n) + ", \"prec\": " +
                          std::to_string(prec) + ", \"tile\": [";
        try {
            Timer tm;
            auto ops = mix_circuit(n, 64, 20261202, false);
            // tile passes
            int best_l = 0, best_c = 0;
            float best = 1e30f;
            const int ab = prec == 8 ? 16 : 8;
            bool first = true;
            // kernel: 0 = LDS-resident tile, 256 / 512 = register tile with
            // that many threads per workgroup
            int best_k = 0;
            int &thr_ref = prec == 8 ? t.rthreads_f64 : t.rthreads_f32;
            for (int kern : {0, 256, 512}) {
                for (int logt = 9; logt <= 13; logt++) {
                    for (int low = 2; low <= 6; low++) {
                        t.reg_tile = kern != 0;
                        if (kern) thr_ref = kern;
                        if (kern && !reg_tile_ok(logt, kern, prec)) continue;
                        if (!tile_fits(t, logt, low, ab)) continue;
                        if (prec == 8) {
                            t.logt_f64 = logt;
                            t.low_f64 = low;
                        } else {
                            t.logt_f32 = logt;
                            t.low_f32 = low;
                        }
                        float ms = time_apply(h, ops, QSLHIP_MODE_FUSED, tm);
                        char buf[96];
                        std::snprintf(buf, sizeof buf, "%s[%d, %d, %d, %.4f]", first ? "" : ", ",
                                      kern, logt, low, ms);
                        rep += buf;
                        first = false;
                        if (ms < best) {
                            best = ms;
                            best_l = logt;
                            best_c = low;
                            best_k = kern;
                        }
                    }
                }
            }
            t.reg_tile = best_k != 0;
            thr_ref = best_k ? best_k : 256;
            if (prec == 8) {
                t.logt_f64 = best_l;
                t.low_f64 = best_c;
            } else {
                t.logt_f32 = best_l;
                t.low_f32 = best_c;
            }
            // the other precision's widths must stay runnable
            if (!tile_fits(t, t.logt_f64, t.low_f64, 16)) {
                t.logt_f64 = 10;
                t.low_f64 = 3;
            }
            if (!tile_fits(t, t.logt_f32, t.low_f32, 8)) {
                t.logt_f32 = 11;
                t.low_f32 = 3;
            }
            rep += "], \"tile_best\": [" + std::to_string(best_k) + ", " +
                   std::to_string(best_l) + ", " + std::to_string(best_c) + "]";
            // per-gate threshold: one-op tile vs streaming, target by target
            int thr = -1;
            rep += ", \"per_gate\": [";
            for (int tq = 0; tq < 8 && tq < best_l; tq++) {
                std::vector<qslhip_op> one(8);
                for (auto &o : one) o = qslhip_op{QSLHIP_ROT_X, (uint32_t)tq, 0, 0, 0.3};
                int &lim = prec == 8 ? t.lds_max_target_f64 : t.lds_max_target_f32;
                lim = tq;
                float ms_tile = time_apply(h, one, QSLHIP_MODE_PER_GATE, tm);
                lim = -1;
                float ms_str = time_apply(h, one, QSLHIP_MODE_PER_GATE, tm);
                char buf[96];
                std::snprintf(buf, sizeof buf, "%s[%d, %.4f, %.4f]", tq ? ", " : "", tq,
                              ms_tile, ms_str);
                rep += buf;
                if (ms_tile < ms_str && thr == tq - 1) thr = tq;
            }
            (prec == 8 ? t.lds_max_target_f64 : t.lds_max_target_f32) = thr;
            rep += "], \"lds_max_target\": " + std::to_string(thr);
            // dense fusion width
            if (prec == 4) {
                auto hops = mix_circuit(n, 96, 20261205, true);
                float ms4, ms5;
                t.dense_f = 4;
                ms4 = time_apply(h, hops, QSLHIP_MODE_DENSE, tm);
                t.dense_f = 5;
                ms5 = time_apply(h, hops, QSLHIP_MODE_DENSE, tm);
                t.dense_f = ms4 < ms5 ? 4 : 5;
                char buf[128];
                std::snprintf(buf, sizeof buf, ", \"dense\": [[4, %.4f], [5, %.4f]], \"dense_f\": %d",
                              ms4, ms5, t.dense_f);
                rep += buf;
            }
            rep += "}";
        } catch (...) {
            t = saved;
            qslhip_destroy(h);
            throw;
        }
        qslhip_destroy(h);
        if (report && cap) {
            std::strncpy(report, rep.c_str(), cap - 1);
            report[cap - 1] = 0;
        }
    });
}

}  // extern "C"
