// qsl/benchmark/hip_sweep.hpp -- gate-timing sweeps over the qubit count,
// for any simulators with the qsl member API, written in qsl's results
// format (qsl/benchmark/hip_results.hpp).
//
// nqubitSweep<Gate, Sims...>(file) follows the reference's nqubitSweep
// (include/qsl/benchmark/time.hpp:126-163) and its single-simulator timing
// (time_single.tpp:54-180): for n = nmin .. nmax-1 qubits, each simulator
// gets a random state; the gate is applied test_len = max(1, 2^len_factor /
// 2^n) times to every qubit (one-qubit gates) or every pair n < m
// (two-qubit gates); the row holds, per simulator, the mean over
// qubits/pairs of the time per application, in seconds.
//
// What differs: a simulator whose member calls return before the device has
// finished (the MI355X backend: gates are queued on its stream) is
// synchronised before the clock starts and before it stops, so the time is
// the device's, not the launch rate's.  A simulator is treated so when it
// has a `sync()` member.  Random states and angles come from a seeded
// mt19937_64 (the reference's generators are unseedable).
#ifndef QSL_BENCHMARK_HIP_SWEEP_HPP
#define QSL_BENCHMARK_HIP_SWEEP_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "qsl/benchmark/hip_results.hpp"
#include "qsl/utils/complex.hpp"

namespace qsl::hipbench {

enum class Gate { pauliX, hadamard, rotateX, rotateY, rotateZ, phase, controlNot, controlPhase };

inline const char *gateName(Gate g)
{
    switch (g) {
    case Gate::pauliX: return "pauliX";
    case Gate::hadamard: return "hadamard";
    case Gate::rotateX: return "rotateX";
    case Gate::rotateY: return "rotateY";
    case Gate::rotateZ: return "rotateZ";
    case Gate::phase: return "phase";
    case Gate::controlNot: return "controlNot";
    case Gate::controlPhase: return "controlPhase";
    }
    return "?";
}

inline bool twoQubit(Gate g) { return g == Gate::controlNot || g == Gate::controlPhase; }

template <typename Sim>
void syncDevice(Sim &s)
{
    if constexpr (requires { s.sync(); }) s.sync();
}

template <typename Sim>
void applyGate(Sim &s, Gate g, unsigned a, unsigned b, typename Sim::Fp_type angle)
{
    switch (g) {
    case Gate::pauliX: s.pauliX(a); break;
    case Gate::hadamard: s.hadamard(a); break;
    case Gate::rotateX: s.rotateX(a, angle); break;
    case Gate::rotateY: s.rotateY(a, angle); break;
    case Gate::rotateZ: s.rotateZ(a, angle); break;
    case Gate::phase: s.phase(a, angle); break;
    case Gate::controlNot: s.controlNot(a, b); break;
    case Gate::controlPhase: s.controlPhase(a, b, angle); break;
    }
}

// Uniform (-1, 1) real and imaginary parts, normalised: the recipe of
// makeRandomState (src/utils/random.cpp:48-68), from a seeded generator.
template <std::floating_point Fp>
std::vector<complex<Fp>> randomState(unsigned n, std::uint64_t seed)
{
    std::mt19937_64 g(seed);
    auto u = [&] { return 2 * ((g() >> 11) * 0x1.0p-53) - 1; };
    std::vector<complex<Fp>> v(std::size_t(1) << n);
    double s = 0;
    for (auto &a : v) {
        const double re = u(), im = u();
        a.real = (Fp)re;
        a.imag = (Fp)im;
        s += re * re + im * im;
    }
    const double f = 1 / std::sqrt(s);
    for (auto &a : v) {
        a.real = (Fp)(a.real * f);
        a.imag = (Fp)(a.imag * f);
    }
    return v;
}

// Per-qubit (or per-pair) times of `test_len` applications on one random
// n-qubit state: the table of Time<SingleSim>::impl (time_single.tpp:54-180),
// headed "qubit <gate>" or "Control Target <gate>".
template <typename Sim>
Table<unsigned, double> timeGate(Gate g, unsigned n, std::size_t test_len, std::uint64_t seed = 1)
{
    using Fp = typename Sim::Fp_type;
    Sim sim{n};
    sim.setState(randomState<Fp>(n, seed));
    std::mt19937_64 rg(seed + 1);
    auto angle = [&] { return (Fp)(2 * M_PI * ((rg() >> 11) * 0x1.0p-53)); };
    const std::string name = gateName(g);
    Table<unsigned, double> r = twoQubit(g) ? Table<unsigned, double>({"Control", "Target", name})
                                            : Table<unsigned, double>({"qubit", name});
    r.addMeta("Testing '" + name + "' member function");
    r.addMeta("Using simulator '" + Sim::name + "'");
    r.addMeta("Number of qubits: " + std::to_string(n));
    r.addMeta("Test length (function repeats per row): " + std::to_string(test_len));
    r.addMeta("Using a single random simulator for every function evaluation");
    r.addMeta("Format: space separated, one header row");
    using clock = std::chrono::steady_clock;
    auto time_loop = [&](unsigned a, unsigned b) {
        const Fp th = angle();
        syncDevice(sim);
        const auto t0 = clock::now();
        for (std::size_t i = 0; i < test_len; i++) applyGate(sim, g, a, b, th);
        syncDevice(sim);
        return std::chrono::duration<double>(clock::now() - t0).count();
    };
    if (twoQubit(g)) {
        for (unsigned a = 0; a < n; a++)
            for (unsigned b = a + 1; b < n; b++) r.addRow({a, b}, {time_loop(a, b)});
    } else {
        for (unsigned a = 0; a < n; a++) r.addRow({a}, {time_loop(a, 0)});
    }
    return r;
}

// The sweep: one row per qubit count, one column per simulator (headed by
// its name), each the mean time of one application in seconds.
template <typename... Sims>
Table<unsigned, double> nqubitSweep(Gate g, unsigned len_factor = 20, unsigned nmin = 2,
                                    unsigned nmax = 19, std::ostream *log = nullptr)
{
    Table<unsigned, double> res({std::string("nqubits"), Sims::name...});
    const std::size_t base = std::size_t(1) << len_factor;
    const unsigned col = twoQubit(g) ? 2 : 1;
    for (unsigned n = std::max(nmin, twoQubit(g) ? 2u : 1u); n < nmax; n++) {
        const std::size_t test_len = std::max<std::size_t>(1, base >> n);
        std::vector<double> avg{(timeGate<Sims>(g, n, test_len).mean(col) / (double)test_len)...};
        if (log) {
            *log << n;
            for (double a : avg) *log << ' ' << a;
            *log << std::endl;
        }
        res.addRow({n}, avg);
    }
    res.addMeta(std::string("Averaged gate times for ") + gateName(g));
    res.addMeta("Results are in seconds");
    return res;
}

template <typename... Sims>
void nqubitSweep(Gate g, const std::string &results_file, unsigned len_factor = 20,
                 unsigned nmin = 2, unsigned nmax = 19)
{
    nqubitSweep<Sims...>(g, len_factor, nmin, nmax, &std::cout).writeToFile(results_file);
    std::cout << std::endl << "Written results to " << results_file << std::endl;
}

}  // namespace qsl::hipbench

#endif
