// test_qubits_hip.cpp -- typed tests of qsl::Qubits<Type::Hip, Fp> and
// qsl::Qubits<Type::HipNP, Fp> through the C++20 API, in the shape of the
// reference's gtest suite (src/gtest/sim-constructors.cpp, gates.cpp,
// measurements.cpp).  The checker is the CPU oracle (oracle/liboracle.so),
// whose gate arithmetic is pinned bit for bit to the reference's own build.
//
//   ./test_qubits_hip            run everything (needs an MI355X)
//   ./test_qubits_hip --compile  only the compile-time concept checks
#include <qsl/qubits.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <sstream>
#include <string>
#include <vector>

extern "C" {
#include "qsl_oracle.h"
}

// ---- compile-time contract (include/qsl/concepts.hpp) ----------------------
static_assert(qsl::Simulator<qsl::Qubits<qsl::Type::Hip, double>>);
static_assert(qsl::Simulator<qsl::Qubits<qsl::Type::Hip, float>>);
static_assert(qsl::HasAllGates<qsl::Qubits<qsl::Type::Hip, double>>);
static_assert(qsl::NonNPSimulator<qsl::Qubits<qsl::Type::Hip, float>>);
static_assert(qsl::NPSimulator<qsl::Qubits<qsl::Type::Hip, double>>);
static_assert(qsl::Simulator<qsl::Qubits<qsl::Type::HipNP, double>>);
static_assert(qsl::NPSimulator<qsl::Qubits<qsl::Type::HipNP, double>>);
static_assert(qsl::NPSimulator<qsl::Qubits<qsl::Type::HipNP, float>>);
static_assert(!qsl::HasNonNPGates<qsl::Qubits<qsl::Type::HipNP, double>>);
static_assert(qsl::Simulator<qsl::Qubits<qsl::Type::HipResize, double>>);
static_assert(qsl::HasAllGates<qsl::Qubits<qsl::Type::HipResize, float>>);
static_assert(qsl::HasMeasurement<qsl::Qubits<qsl::Type::HipResize, double>>);

// ---- a minimal test runner ---------------------------------------------------
static int g_fail = 0, g_checks = 0;
static std::vector<std::pair<std::string, std::function<void()>>> &registry()
{
    static std::vector<std::pair<std::string, std::function<void()>>> r;
    return r;
}
struct Reg {
    Reg(const char *n, std::function<void()> f) { registry().emplace_back(n, f); }
};
#define TEST(name)                        \
    static void name();                   \
    static Reg reg_##name(#name, name);   \
    static void name()
#define EXPECT(cond)                                                              \
    do {                                                                          \
        g_checks++;                                                               \
        if (!(cond)) {                                                            \
            g_fail++;                                                             \
            std::fprintf(stderr, "  FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
        }                                                                         \
    } while (0)
#define EXPECT_THROW(stmt, exc)              \
    do {                                     \
        bool thrown_ = false;                \
        try {                                \
            stmt;                            \
        } catch (const exc &) {              \
            thrown_ = true;                  \
        }                                    \
        EXPECT(thrown_ && #stmt " throws " #exc); \
    } while (0)

template <typename Fp>
using Hip = qsl::Qubits<qsl::Type::Hip, Fp>;
template <typename Fp>
using HipNP = qsl::Qubits<qsl::Type::HipNP, Fp>;

// oracle dispatch by precision
template <typename Fp>
struct Orc;
template <>
struct Orc<double> {
    static void g1(double *s, unsigned n, int k, unsigned t, double a) { orc_gate1_d(s, n, k, t, a); }
    static void g2(double *s, unsigned n, int k, unsigned a, unsigned b, double x)
    {
        orc_gate2_d(s, n, k, a, b, x);
    }
    static double prob(const double *s, unsigned n, unsigned t, unsigned o) { return orc_prob_d(s, n, t, o); }
    static int measure(double *s, unsigned n, unsigned t, double u) { return orc_measure_d(s, n, t, u); }
    static void sample_all(const double *s, unsigned n, const double *u, size_t m, uint64_t *o)
    {
        orc_sample_all_d(s, n, u, m, o);
    }
    static void sample_all2(const double *s, unsigned n, const double *u, size_t m, uint64_t *c)
    {
        orc_sample_all2_d(s, n, u, m, c);
    }
};
template <>
struct Orc<float> {
    static void g1(float *s, unsigned n, int k, unsigned t, float a) { orc_gate1_f(s, n, k, t, a); }
    static void g2(float *s, unsigned n, int k, unsigned a, unsigned b, float x)
    {
        orc_gate2_f(s, n, k, a, b, x);
    }
    static float prob(const float *s, unsigned n, unsigned t, unsigned o) { return orc_prob_f(s, n, t, o); }
    static int measure(float *s, unsigned n, unsigned t, float u) { return orc_measure_f(s, n, t, u); }
    static void sample_all(const float *s, unsigned n, const float *u, size_t m, uint64_t *o)
    {
        orc_sample_all_f(s, n, u, m, o);
    }
    static void sample_all2(const float *s, unsigned n, const float *u, size_t m, uint64_t *c)
    {
        orc_sample_all2_f(s, n, u, m, c);
    }
};

template <typename Fp>
static Fp *raw(std::vector<qsl::complex<Fp>> &v)
{
    return reinterpret_cast<Fp *>(v.data());
}

template <typename Fp>
static bool same_bits(const std::vector<qsl::complex<Fp>> &a, const std::vector<qsl::complex<Fp>> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(a[0])) == 0;
}

template <typename Fp>
static double max_diff(const std::vector<qsl::complex<Fp>> &a, const std::vector<qsl::complex<Fp>> &b)
{
    double m = 0;
    for (size_t i = 0; i < a.size(); i++)
        m = std::fmax(m, std::fabs((double)a[i].real - b[i].real) +
                             std::fabs((double)a[i].imag - b[i].imag));
    return m;
}

// ---- constructors and utilities (sim-constructors.cpp:107-260) --------------
template <typename Fp>
static void constructor_test()
{
    using Sim = Hip<Fp>;
    const unsigned n = 5;
    Sim q{n};
    const auto state = qsl::makeRandomState<Fp>(n, 11);
    q.setState(state);
    Sim q_copy{q};
    EXPECT(q.getNumQubits() == q_copy.getNumQubits());
    EXPECT(same_bits(q.getState(), q_copy.getState()));
    Sim q_from_state{state};
    EXPECT(same_bits(q.getState(), q_from_state.getState()));
    Sim q_new{n};
    q_new = q;
    EXPECT(same_bits(q.getState(), q_new.getState()));
    Sim q_wrong{n + 1};
    EXPECT_THROW(q_wrong = q, std::logic_error);
    // utility functions
    Sim q3{3};
    EXPECT_THROW(q3.setState(qsl::makeRandomState<Fp>(4, 1)), std::logic_error);
    q3.setBasisState(4);
    auto b = q3.getState();
    EXPECT(b[4].real == 1 && std::fabs(qsl::norm(b) - 1) < 1e-6);
    EXPECT_THROW(q3.hadamard(3), std::out_of_range);
    EXPECT(Sim::name == (sizeof(Fp) == 8 ? "Qub<hip,double>" : "Qub<hip,float>"));
}
TEST(ConstructorTestDouble) { constructor_test<double>(); }
TEST(ConstructorTestFloat) { constructor_test<float>(); }

template <typename Fp>
static void np_constructor_test()
{
    using Sim = HipNP<Fp>;
    const unsigned n = 5;
    Sim q{n};
    const auto state = qsl::makeRandomNPState<Fp>(n, 3, 5);
    q.setState(state);
    EXPECT(q.getNumOnes() == 3);
    Sim q_copy{q};
    EXPECT(same_bits(q.getState(), q_copy.getState()));
    Sim q_from_state{state};
    EXPECT(same_bits(q.getState(), q_from_state.getState()));
    Sim q_new{n};
    q_new = q;
    EXPECT(same_bits(q.getState(), q_new.getState()));
    Sim q_wrong{n + 1};
    EXPECT_THROW(q_wrong = q, std::logic_error);
    // sim-constructors.cpp:202-233
    EXPECT_THROW(Sim(4, 5), std::logic_error);
    Sim q4{4};
    EXPECT_THROW(q4.setNumOnes(5), std::logic_error);
    q4.setNumOnes(2);
    EXPECT(q4.getNumOnes() == 2);
    EXPECT_THROW(q4.setState(qsl::makeRandomNPState<Fp>(5, 2, 1)), std::logic_error);
    q4.setBasisState(4);
    auto b = q4.getState();
    EXPECT(b[4].real == 1 && q4.getNumOnes() == 1);
}
TEST(NPConstructorTestDouble) { np_constructor_test<double>(); }
TEST(NPConstructorTestFloat) { np_constructor_test<float>(); }

// print format (sim-constructors.cpp:254-293)
TEST(PrintTest)
{
    Hip<double> q{3};
    std::stringstream s1, s2;
    q.print(s1);
    s2 << "Number of qubits = 3" << std::endl << "(1, 0)" << std::endl;
    for (int i = 0; i < 7; i++) s2 << "(0, 0)" << std::endl;
    s2 << std::endl;
    EXPECT(s1.str() == s2.str());

    HipNP<float> p{3, 2};
    std::stringstream t1, t2;
    p.print(t1);
    t2 << "Number of qubits = 3" << std::endl << "Number of ones = 2" << std::endl;
    for (int i = 0; i < 3; i++) t2 << "(0, 0)" << std::endl;
    t2 << "(1, 0)" << std::endl;
    for (int i = 0; i < 4; i++) t2 << "(0, 0)" << std::endl;
    t2 << std::endl;
    EXPECT(t1.str() == t2.str());
}

// ---- gates against the oracle (gates.cpp:313-878 shape: n = 8, targ 4,
// ctrl 2, random state, random angle), here over every target pair ---------
template <typename Fp>
static void gate_test()
{
    const unsigned n = 8;
    Hip<Fp> q{n};
    const auto st0 = qsl::makeRandomState<Fp>(n, 42);
    using G1 = void (Hip<Fp>::*)(unsigned);
    using G1a = void (Hip<Fp>::*)(unsigned, Fp);
    using G2 = void (Hip<Fp>::*)(unsigned, unsigned);
    using G2a = void (Hip<Fp>::*)(unsigned, unsigned, Fp);
    const std::vector<std::pair<int, G1>> g1 = {
        {QSLHIP_PAULI_X, &Hip<Fp>::pauliX}, {QSLHIP_PAULI_Y, &Hip<Fp>::pauliY},
        {QSLHIP_PAULI_Z, &Hip<Fp>::pauliZ}, {QSLHIP_HADAMARD, &Hip<Fp>::hadamard}};
    const std::vector<std::pair<int, G1a>> g1a = {
        {QSLHIP_PHASE, &Hip<Fp>::phase}, {QSLHIP_ROT_X, &Hip<Fp>::rotateX},
        {QSLHIP_ROT_Y, &Hip<Fp>::rotateY}, {QSLHIP_ROT_Z, &Hip<Fp>::rotateZ}};
    const std::vector<std::pair<int, G2>> g2 = {
        {QSLHIP_CNOT, &Hip<Fp>::controlNot}, {QSLHIP_CY, &Hip<Fp>::controlY},
        {QSLHIP_CZ, &Hip<Fp>::controlZ}, {QSLHIP_CHADAMARD, &Hip<Fp>::controlHadamard},
        {QSLHIP_SWAP, &Hip<Fp>::swap}, {QSLHIP_FSWAP, &Hip<Fp>::fswap},
        {QSLHIP_NP_HADAMARD, &Hip<Fp>::npHadamard}};
    const std::vector<std::pair<int, G2a>> g2a = {
        {QSLHIP_CROT_X, &Hip<Fp>::controlRotateX}, {QSLHIP_CROT_Y, &Hip<Fp>::controlRotateY},
        {QSLHIP_CROT_Z, &Hip<Fp>::controlRotateZ}, {QSLHIP_CPHASE, &Hip<Fp>::controlPhase},
        {QSLHIP_NP_ROT_X, &Hip<Fp>::npRotateX}, {QSLHIP_NP_ROT_Y, &Hip<Fp>::npRotateY}};
    const Fp angle = static_cast<Fp>(M_PI * 0.7316);
    for (unsigned t = 0; t < n; t++) {
        for (auto &[k, f] : g1) {
            q.setState(st0);
            (q.*f)(t);
            auto want = st0;
            Orc<Fp>::g1(raw(want), n, k, t, 0);
            EXPECT(same_bits(q.getState(), want));
        }
        for (auto &[k, f] : g1a) {
            q.setState(st0);
            (q.*f)(t, angle);
            auto want = st0;
            Orc<Fp>::g1(raw(want), n, k, t, angle);
            EXPECT(same_bits(q.getState(), want));
        }
        for (unsigned c = 0; c < n; c++) {
            if (c == t) continue;
            for (auto &[k, f] : g2) {
                q.setState(st0);
                (q.*f)(c, t);
                auto want = st0;
                Orc<Fp>::g2(raw(want), n, k, c, t, 0);
                EXPECT(same_bits(q.getState(), want));
            }
            for (auto &[k, f] : g2a) {
                q.setState(st0);
                (q.*f)(c, t, angle);
                auto want = st0;
                Orc<Fp>::g2(raw(want), n, k, c, t, angle);
                EXPECT(same_bits(q.getState(), want));
            }
        }
    }
}
TEST(GatesDouble) { gate_test<double>(); }
TEST(GatesFloat) { gate_test<float>(); }

// ---- measurement (measurements.cpp:80-856) -------------------------------------
template <typename Fp>
static void measurement_test()
{
    const unsigned n = 8;
    const double tol = sizeof(Fp) == 8 ? 1e-12 : 1e-5;
    auto st = qsl::makeRandomState<Fp>(n, 7);
    Hip<Fp> q{st};
    for (unsigned t = 0; t < n; t++)
        for (unsigned o = 0; o < 2; o++)
            EXPECT(std::fabs(q.prob(t, o) - Orc<Fp>::prob(raw(st), n, t, o)) < tol);
    // postselect: prob(targ, outcome) becomes 1 (measurements.cpp:160-283)
    {
        Hip<Fp> p{st};
        Fp pr = p.postselect(3, 1);
        EXPECT(std::fabs(pr - Orc<Fp>::prob(raw(st), n, 3, 1)) < tol);
        EXPECT(std::fabs(p.prob(3, 1) - 1) < 10 * tol);
    }
    // measure with injected uniforms: identical outcome and collapse
    for (unsigned t = 0; t < n; t++) {
        Hip<Fp> m{st};
        auto want = st;
        Fp u = static_cast<Fp>(0.37);
        EXPECT(m.measure(t, u) == Orc<Fp>::measure(raw(want), n, t, u));
        EXPECT(max_diff(m.getState(), want) < 10 * tol);
    }
    // sampleAll / sampleAll2 with injected uniforms
    std::vector<Fp> u(4000);
    std::mt19937_64 g(3);
    for (auto &x : u) x = static_cast<Fp>((g() >> 11) * 0x1.0p-53);
    auto m1 = q.sampleAll(u);
    std::vector<uint64_t> o(u.size());
    Orc<Fp>::sample_all(raw(st), n, u.data(), u.size(), o.data());
    std::map<std::size_t, std::size_t> m2;
    for (auto x : o) m2[x]++;
    EXPECT(m1 == m2);
    auto a1 = q.sampleAll2(u);
    std::vector<uint64_t> c(1u << n);
    Orc<Fp>::sample_all2(raw(st), n, u.data(), u.size(), c.data());
    bool same = a1.size() == c.size();
    for (size_t i = 0; same && i < c.size(); i++) same = a1[i] == c[i];
    EXPECT(same);
    // statistical check of the self-drawn streams, as the reference does
    // (measurements.cpp:39), on a qubit with prob(0) = cos^2(theta/2) = 0.9,
    // far from 1/2, so that a reversed `u < prob0` rule (which gives 0.1)
    // fails: 2000 samples, |p_est - 0.9| < 0.05 (about 7 sigma)
    {
        Hip<Fp> b{3};
        b.seed(1234);
        b.rotateY(1, static_cast<Fp>(2 * std::acos(std::sqrt(0.9))));
        EXPECT(std::fabs(b.prob(1, 0) - 0.9) < 10 * tol);
        auto cnt = b.sample(1, 2000);
        EXPECT(cnt[0] + cnt[1] == 2000);
        EXPECT(std::fabs(cnt[0] / 2000.0 - 0.9) < 0.05);
        // measure draws its own uniform: outcome 0 about 9 times in 10 (a
        // copy carries its source's generator state, so each is reseeded)
        std::size_t zeros = 0;
        for (int i = 0; i < 400; i++) {
            Hip<Fp> c{b};
            c.seed(1000 + (std::uint64_t)i);
            zeros += c.measure(1) == 0;
        }
        EXPECT(std::fabs(zeros / 400.0 - 0.9) < 0.08);
    }
    auto all = q.sampleAll(1000);
    size_t tot = 0;
    for (auto &[k, v] : all) tot += v;
    EXPECT(tot == 1000);
    // measureAll collapses onto the drawn basis state
    std::size_t x = q.measureAll();
    auto s2 = q.getState();
    EXPECT(s2[x].real == 1 && qsl::norm(s2) == 1);
}
TEST(MeasurementDouble) { measurement_test<double>(); }
TEST(MeasurementFloat) { measurement_test<float>(); }

template <typename Fp>
static void np_gate_test()
{
    const unsigned n = 8, k = 4;
    auto st0 = qsl::makeRandomNPState<Fp>(n, k, 9);
    HipNP<Fp> q{st0};
    const Fp angle = static_cast<Fp>(1.2345);
    for (unsigned a = 0; a < n; a++) {
        q.setState(st0);
        q.rotateZ(a, angle);
        auto want = st0;
        if constexpr (sizeof(Fp) == 8) orc_np_gate1_d(raw(want), n, k, QSLHIP_ROT_Z, a, angle);
        else orc_np_gate1_f(raw(want), n, k, QSLHIP_ROT_Z, a, angle);
        EXPECT(same_bits(q.getState(), want));
        for (unsigned b = 0; b < n; b++) {
            if (a == b) continue;
            q.setState(st0);
            q.npRotateX(a, b, angle);
            q.fswap(a, b);
            q.controlPhase(a, b, angle);
            auto w = st0;
            if constexpr (sizeof(Fp) == 8) {
                orc_np_gate2_d(raw(w), n, k, QSLHIP_NP_ROT_X, a, b, angle);
                orc_np_gate2_d(raw(w), n, k, QSLHIP_FSWAP, a, b, 0);
                orc_np_gate2_d(raw(w), n, k, QSLHIP_CPHASE, a, b, angle);
            } else {
                orc_np_gate2_f(raw(w), n, k, QSLHIP_NP_ROT_X, a, b, angle);
                orc_np_gate2_f(raw(w), n, k, QSLHIP_FSWAP, a, b, 0);
                orc_np_gate2_f(raw(w), n, k, QSLHIP_CPHASE, a, b, angle);
            }
            EXPECT(same_bits(q.getState(), w));
        }
    }
    // the compact vector is the weight-k amplitudes in generateLookup order
    auto compact = q.getCompactState();
    auto full = q.getState();
    auto lst = qsl::generateLookup(n, k);
    bool ok = compact.size() == lst.size();
    for (size_t i = 0; ok && i < lst.size(); i++)
        ok = std::memcmp(&compact[i], &full[lst[i]], sizeof(compact[i])) == 0;
    EXPECT(ok);
}
TEST(NPGatesDouble) { np_gate_test<double>(); }
TEST(NPGatesFloat) { np_gate_test<float>(); }

// ---- resizeable backend (src/gtest/resize-test.cpp) -------------------------------
template <typename Fp>
using HipRz = qsl::Qubits<qsl::Type::HipResize, Fp>;

template <typename Fp>
struct OrcRz;
template <>
struct OrcRz<double> {
    static void add(double *s, unsigned n, unsigned t) { orc_add_qubit_d(s, n, t); }
    static double post(double *s, unsigned n, unsigned t, unsigned o)
    {
        return orc_postselect_out_d(s, n, t, o);
    }
    static void out(double *s, unsigned n, unsigned t, unsigned o, double f)
    {
        orc_collapse_out_d(s, n, t, o, f);
    }
    static double prob(const double *s, unsigned n, unsigned t, unsigned o) { return orc_prob_d(s, n, t, o); }
};
template <>
struct OrcRz<float> {
    static void add(float *s, unsigned n, unsigned t) { orc_add_qubit_f(s, n, t); }
    static float post(float *s, unsigned n, unsigned t, unsigned o)
    {
        return orc_postselect_out_f(s, n, t, o);
    }
    static void out(float *s, unsigned n, unsigned t, unsigned o, float f)
    {
        orc_collapse_out_f(s, n, t, o, f);
    }
    static float prob(const float *s, unsigned n, unsigned t, unsigned o) { return orc_prob_f(s, n, t, o); }
};

template <typename Fp>
static void resize_test()
{
    using Sim = HipRz<Fp>;
    const double tol = sizeof(Fp) == 8 ? 1e-12 : 1e-5;
    // constructors and assignment (resize-test.cpp:32-76)
    const unsigned n = 5;
    const auto state = qsl::makeRandomState<Fp>(n, 21);
    Sim q{n};
    q.setState(state);
    Sim q_copy{q};
    EXPECT(q_copy.getNumQubits() == n && same_bits(q_copy.getState(), state));
    Sim q_new{n + 2};
    q_new = q;  // any size
    EXPECT(q_new.getNumQubits() == n && same_bits(q_new.getState(), state));
    EXPECT(Sim::name == (sizeof(Fp) == 8 ? "Qub<hipresize,double>" : "Qub<hipresize,float>"));
    // setState of another size (resize-test.cpp:79-100)
    Sim u{3};
    u.setState(qsl::makeRandomState<Fp>(4, 2));
    EXPECT(u.getNumQubits() == 4);
    u.setBasisState(4);
    EXPECT(u.getState()[4].real == 1);
    // appendQubit (resize-test.cpp:260-282): the old vector then zeros
    {
        Sim a{state};
        a.appendQubit();
        auto v = a.getState();
        EXPECT(a.getNumQubits() == n + 1 && v.size() == state.size() * 2);
        bool ok = true;
        for (std::size_t i = 0; i < v.size(); i++) {
            const auto w = i < state.size() ? state[i] : qsl::complex<Fp>{0, 0};
            ok = ok && v[i].real == w.real && v[i].imag == w.imag;
        }
        EXPECT(ok);
    }
    // addQubit at every index, against the oracle (bit for bit)
    for (unsigned t = 0; t <= n; t++) {
        Sim a{state};
        a.addQubit(t);
        std::vector<qsl::complex<Fp>> want(state.size() * 2);
        std::copy(state.begin(), state.end(), want.begin());
        OrcRz<Fp>::add(raw(want), n, t);
        EXPECT(same_bits(a.getState(), want));
    }
    // postselectOut (resize-test.cpp:193-258)
    for (unsigned t = 0; t < n; t++)
        for (unsigned o = 0; o < 2; o++) {
            Sim a{state};
            Fp p = a.postselectOut(t, o);
            auto want = state;
            Fp wp = OrcRz<Fp>::post(raw(want), n, t, o);
            want.resize(state.size() / 2);
            EXPECT(std::fabs(p - wp) < tol && a.getNumQubits() == n - 1);
            EXPECT(max_diff(a.getState(), want) < 10 * tol);
        }
    // measureOut (resize-test.cpp:103-190): outcome frequencies and the
    // collapsed state for every outcome drawn
    for (unsigned t = 0; t < n; t++) {
        std::size_t ones = 0;
        const int samples = 300;
        for (int k = 0; k < samples; k++) {
            Sim a{state};
            a.seed(77 + 1000 * t + (std::uint64_t)k);
            int o = a.measureOut(t);
            ones += o;
            auto want = state;
            const double p0 = OrcRz<Fp>::prob(raw(want), n, t, 0);
            const double f = o == 0 ? 1 / std::sqrt(p0) : 1 / std::sqrt(1 - p0);
            OrcRz<Fp>::out(raw(want), n, t, (unsigned)o, (Fp)f);
            want.resize(state.size() / 2);
            EXPECT(max_diff(a.getState(), want) < 100 * tol);
        }
        const double p1 = OrcRz<Fp>::prob(raw(const_cast<std::vector<qsl::complex<Fp>> &>(state)), n, t, 1);
        EXPECT(std::fabs((double)ones / samples - p1) < 0.12);  // 4 sigma at 300
    }
    // remove then add again inside the allocation, then trim it
    {
        Sim a{state};
        a.postselectOut(2, 0);
        a.addQubit(2);
        a.reallocateState();
        EXPECT(a.getNumQubits() == n && a.prob(2, 1) == 0);
    }
    EXPECT_THROW(Sim{1}.postselectOut(0, 0), std::out_of_range);
}
TEST(ResizeDouble) { resize_test<double>(); }
TEST(ResizeFloat) { resize_test<float>(); }

// ---- expectation values, adjoint gradients, batches (no reference
// counterpart; pinned by the mathematics: parameter shift is exact for
// rotations, central differences for the rest, all on the FP64 oracle) ----
static double energy_of(const std::vector<double> &st, unsigned n,
                        const std::vector<qsl::hip::Pauli> &terms)
{
    // P|j> = i^nY (-1)^popcount(j & z) |j ^ x>; E = sum coef Re <psi|P|psi>
    static const double ir[4] = {1, 0, -1, 0}, ii[4] = {0, 1, 0, -1};
    double e = 0;
    for (const auto &t : terms) {
        const int ny = __builtin_popcountll(t.xmask & t.zmask) & 3;
        double acc = 0;
        for (std::uint64_t j = 0; j < (std::uint64_t(1) << n); j++) {
            const double sg = (__builtin_popcountll(j & t.zmask) & 1) ? -1.0 : 1.0;
            const double ar = st[2 * j], ai = st[2 * j + 1];
            const double cr = sg * (ir[ny] * ar - ii[ny] * ai);
            const double ci = sg * (ir[ny] * ai + ii[ny] * ar);
            const std::uint64_t k = j ^ t.xmask;
            acc += st[2 * k] * cr + st[2 * k + 1] * ci;
        }
        e += t.coef * acc;
    }
    return e;
}

static double oracle_energy(unsigned n, const std::vector<qsl::hip::Op> &ops,
                            const std::vector<qsl::hip::Pauli> &terms)
{
    std::vector<double> st(std::size_t(2) << n, 0.0);
    st[0] = 1;
    static_assert(sizeof(orc_op) == sizeof(qsl::hip::Op));
    orc_apply_ops_d(st.data(), n, reinterpret_cast<const orc_op *>(ops.data()), ops.size());
    return energy_of(st, n, terms);
}

// the config-4 ansatz: per layer RY, RZ on every qubit, then a CNOT ladder
static std::vector<qsl::hip::Op> ansatz(unsigned n, unsigned layers, std::mt19937_64 &g)
{
    using qsl::hip::gate;
    auto ang = [&] { return M_PI * (2 * ((g() >> 11) * 0x1.0p-53) - 1); };
    std::vector<qsl::hip::Op> ops;
    for (unsigned l = 0; l < layers; l++) {
        for (unsigned q = 0; q < n; q++) ops.push_back(gate(QSLHIP_ROT_Y, q, 0, ang()));
        for (unsigned q = 0; q < n; q++) ops.push_back(gate(QSLHIP_ROT_Z, q, 0, ang()));
        for (unsigned q = 0; q + 1 < n; q++) ops.push_back(gate(QSLHIP_CNOT, q, q + 1));
    }
    return ops;
}

static std::vector<qsl::hip::Pauli> c4_terms(unsigned n)
{
    std::vector<qsl::hip::Pauli> t{qsl::hip::pauli(1.0, {{0, 'Z'}, {1, 'Z'}})};
    for (unsigned q = 0; q < n; q++) t.push_back(qsl::hip::pauli(0.5, {{q, 'X'}}));
    return t;
}

TEST(ExpvalGradDouble)
{
    using qsl::hip::gate;
    const unsigned n = 6;
    std::mt19937_64 g(41);
    auto ops = ansatz(n, 3, g);
    // every differentiable kind, fixed gates in between
    ops.push_back(gate(QSLHIP_CROT_X, 2, 4, 0.7));
    ops.push_back(gate(QSLHIP_HADAMARD, 1));
    ops.push_back(gate(QSLHIP_CROT_Y, 5, 0, -1.1));
    ops.push_back(gate(QSLHIP_CROT_Z, 0, 3, 2.3));
    ops.push_back(gate(QSLHIP_PHASE, 2, 0, 0.4));
    ops.push_back(gate(QSLHIP_CPHASE, 1, 5, -0.9));
    ops.push_back(gate(QSLHIP_ROT_X, 4, 0, 1.9));
    ops.push_back(gate(QSLHIP_NP_ROT_X, 1, 2, 0.3));
    ops.push_back(gate(QSLHIP_NP_ROT_Y, 3, 4, -0.6));
    auto terms = c4_terms(n);
    terms.push_back(qsl::hip::pauli(-0.3, {{2, 'Y'}, {4, 'X'}}));
    terms.push_back(qsl::hip::pauli(0.8, {{1, 'Y'}, {3, 'Z'}, {5, 'Y'}}));

    Hip<double> q{n};
    auto [e, grad] = q.expvalGrad(ops, terms);
    EXPECT(grad.size() == ops.size());
    EXPECT(std::fabs(e - oracle_energy(n, ops, terms)) < 1e-12);
    // the object keeps U|0>: expval of the same observable is E again
    EXPECT(std::fabs(q.expval(terms) - e) < 1e-12);
    for (std::size_t i = 0; i < ops.size(); i++) {
        const int k = ops[i].kind;
        const bool rot = k == QSLHIP_ROT_X || k == QSLHIP_ROT_Y || k == QSLHIP_ROT_Z;
        const bool param = rot || k == QSLHIP_PHASE || k == QSLHIP_CROT_X ||
                           k == QSLHIP_CROT_Y || k == QSLHIP_CROT_Z || k == QSLHIP_CPHASE ||
                           k == QSLHIP_NP_ROT_X || k == QSLHIP_NP_ROT_Y;
        if (!param) {
            EXPECT(grad[i] == 0.0);
            continue;
        }
        auto shifted = [&](double d) {
            auto o = ops;
            o[i].angle += d;
            return oracle_energy(n, o, terms);
        };
        if (rot) {
            // parameter shift: exact for exp(-i t P / 2)
            const double ps = (shifted(M_PI / 2) - shifted(-M_PI / 2)) / 2;
            EXPECT(std::fabs(grad[i] - ps) < 1e-10);
        } else {
            const double h = 1e-3;
            const double fd = (shifted(h) - shifted(-h)) / (2 * h);
            EXPECT(std::fabs(grad[i] - fd) < 1e-6);
        }
    }
    // a refused observable (qubit outside the register) throws
    std::vector<qsl::hip::Pauli> bad{qsl::hip::pauli(1.0, {{n, 'Z'}})};
    EXPECT_THROW(q.expval(bad), std::out_of_range);
}

TEST(BatchAgainstPlainFP64)
{
    std::mt19937_64 g(7);
    std::vector<qsl::HipBatch::Circuit> cs;
    for (unsigned n : {14u, 6u, 13u, 9u, 12u, 10u}) cs.push_back({n, ansatz(n, 4, g), c4_terms(n)});
    qsl::HipBatch b{cs};
    EXPECT(b.numCircuits() == cs.size());
    std::size_t total = 0;
    for (auto &c : cs) total += c.ops.size();
    EXPECT(b.numOps() == total);
    auto check = [&](const qsl::HipBatch::Result &r) {
        for (std::size_t c = 0; c < cs.size(); c++) {
            Hip<double> q{cs[c].nqubits};
            auto [e64, g64] = q.expvalGrad(cs[c].ops, cs[c].terms);
            EXPECT(std::fabs(r.energies[c] - e64) < 1e-5);
            auto gc = r.gradient(c);
            EXPECT(gc.size() == g64.size());
            double m = 0;
            for (std::size_t i = 0; i < gc.size(); i++) m = std::fmax(m, std::fabs(gc[i] - g64[i]));
            EXPECT(m <= 1e-4);
        }
    };
    check(b.run());
    // replays with new angles (the hipGraph of the batch), each checked
    for (int rep = 0; rep < 2; rep++) {
        std::vector<double> angles;
        for (auto &c : cs)
            for (auto &o : c.ops) {
                if (o.kind != QSLHIP_CNOT) o.angle = M_PI * (2 * ((g() >> 11) * 0x1.0p-53) - 1);
                angles.push_back(o.angle);
            }
        check(b.run(angles));
    }
    // one circuit against the oracle's parameter shift
    {
        auto r = b.run();
        const auto &c = cs[1];
        auto gc = r.gradient(1);
        for (std::size_t i = 0; i < c.ops.size(); i += 5) {
            if (c.ops[i].kind == QSLHIP_CNOT) continue;
            auto o = c.ops;
            o[i].angle += M_PI / 2;
            double ep = oracle_energy(c.nqubits, o, c.terms);
            o[i].angle -= M_PI;
            double em = oracle_energy(c.nqubits, o, c.terms);
            EXPECT(std::fabs(gc[i] - (ep - em) / 2) <= 1e-4);
        }
    }
    EXPECT_THROW(b.run(std::vector<double>(3)), std::invalid_argument);
    qsl::HipBatch moved{std::move(b)};
    EXPECT(moved.numCircuits() == cs.size());
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "--compile") == 0) {
        std::printf("compile-time concept checks passed\n");
        return 0;
    }
    for (auto &[name, fn] : registry()) {
        int before = g_fail;
        try {
            fn();
        } catch (const std::exception &e) {
            g_fail++;
            std::fprintf(stderr, "  EXCEPTION in %s: %s\n", name.c_str(), e.what());
        }
        std::printf("[%s] %s\n", g_fail == before ? " OK " : "FAIL", name.c_str());
        std::fflush(stdout);
    }
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
