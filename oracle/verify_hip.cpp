// verify_hip.cpp -- the reference's own cross-simulator harness, Verify
// (include/qsl/verify.hpp:69-134) with its checkers
// (include/qsl/verify/checkers.hpp: ProbChecker :224, MeasureChecker :305,
// SampleChecker :462, SampleAllChecker :589, DefaultGateChecker :945,
// NPGateChecker :1082, PostselectChecker :1178), instantiated with the MI355X
// backends against the reference's Default, NP and Resize backends.
//
// TEST INFRASTRUCTURE.  oracle/Makefile (target `verify`) compiles this file
// against the reference's headers read in place under /root/reference, with
// one of them, qsl/qubits.hpp, patched in the build folder the way the
// integration guide describes (the Hip, HipNP and HipResize tags added to the Type enum,
// and the backend's headers included after the reference's own); it links
// the reference library built from qsl's sources (oracle/_ref/libqslref.so)
// and libqslhip.so.  The binary goes to oracle/_ref/verify_hip.
//
//   verify_hip --compile   the templates instantiated (no device work)
//   verify_hip             every checker on both pairs; exit status 0 when
//                          every bound of the reference's own cross-simulator
//                          test cases (src/tests/basic-tests.cpp:406-521)
//                          holds
#include "qsl/qubits.hpp"
#include "qsl/verify.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

int g_checks = 0, g_fail = 0;

void expect(bool ok, const std::string &what)
{
    g_checks++;
    if (!ok) {
        g_fail++;
        std::printf("  FAIL %s\n", what.c_str());
    }
}

// The bounds the reference's own Default-vs-Omp / Default-vs-NP cases put
// on each checker's results.
// The measure and sample checkers draw 100 samples at 95% Wald intervals by
// default: two such intervals for one probability miss each other about
// once in 200, so an 8-qubit run of either fails about 4% of the time for
// any two simulators.  Both are configured with more samples at 99.9%.
template <typename V>
void configure_statistics(V &verify)
{
    verify.template configureChecker<qsl::MeasureChecker>(2000, 0.999);
    verify.template configureChecker<qsl::SampleChecker>(200000, 0.999);
    verify.template configureChecker<qsl::SampleAllChecker>(1000000, 0.99);
}

template <typename V>
void check_suite(V &verify, const char *label, double state_tol)
{
    std::printf("== %s\n", label);
    auto prob = verify.template check<qsl::ProbChecker>();
    expect(std::abs(prob.template mean<double>(1)) < 1e-10, "prob diff0 mean");
    expect(prob.template variance<double>(1) < 1e-10, "prob diff0 variance");
    expect(std::abs(prob.template mean<double>(2)) < 1e-10, "prob diff1 mean");
    expect(prob.template variance<double>(2) < 1e-10, "prob diff1 variance");

    auto meas = verify.template check<qsl::MeasureChecker>();
    expect(std::abs(meas.template mean<double>(1) - 1) < 1e-10, "measure intervals overlap");
    expect(meas.template variance<double>(1) < 1e-10, "measure overlap variance");

    auto post = verify.template check<qsl::PostselectChecker>();
    for (std::size_t c = 1; c <= 2; c++) {
        expect(std::abs(post.template mean<double>(c)) < 1e-10, "postselect prob diff mean");
        expect(post.template variance<double>(c) < 1e-10, "postselect prob diff variance");
    }
    for (std::size_t c = 3; c <= 4; c++) {
        expect(std::abs(post.template mean<double>(c)) < state_tol, "postselect state distance mean");
        expect(post.template variance<double>(c) < state_tol, "postselect state distance variance");
    }

    auto samp = verify.template check<qsl::SampleChecker>();
    expect(std::abs(samp.template mean<double>(1) - 1) < 1e-10, "sample intervals overlap");

    auto all = verify.template check<qsl::SampleAllChecker>();
    expect(std::abs(all.overlap_rate - 1) < 1e-10, "sampleAll intervals overlap");
    expect(std::abs(all.mean_relative_error) < 0.15, "sampleAll mean relative error < 15%");
    std::printf("   sampleAll: mean relative error %.4f, overlap %.4f, common %.4f\n",
                all.mean_relative_error, all.overlap_rate, all.common_rate);
}

template <typename V, template <typename, typename> class GateChecker>
void check_gates(V &verify)
{
    auto gates = verify.template check<GateChecker>();
    expect(!gates.empty(), "gate checker ran");
    for (const auto &[gate, table] : gates) {
        const double m = table.template mean<double>("distance");
        const double v = table.template variance<double>("distance");
        expect(std::abs(m) < 1e-9, "gate " + gate + " distance mean");
        expect(std::abs(v) < 1e-9, "gate " + gate + " distance variance");
        std::printf("   %-16s mean distance %.3e\n", gate.c_str(), m);
    }
}

using Def = qsl::Qubits<qsl::Type::Default, double>;
using NP = qsl::Qubits<qsl::Type::NP, double>;
using Hip = qsl::Qubits<qsl::Type::Hip, double>;
using HipNP = qsl::Qubits<qsl::Type::HipNP, double>;

using VerifyFull = qsl::Verify<Def, Hip, qsl::DefaultStateGen<double>, qsl::MeasureChecker,
                               qsl::PostselectChecker, qsl::SampleChecker,
                               qsl::SampleAllChecker, qsl::ProbChecker,
                               qsl::DefaultGateChecker>;
using VerifyNP = qsl::Verify<NP, HipNP, qsl::NPStateGen<double>, qsl::MeasureChecker,
                             qsl::PostselectChecker, qsl::SampleChecker, qsl::SampleAllChecker,
                             qsl::ProbChecker, qsl::NPGateChecker>;
// the number-preserving gate set of the full-vector backend against NP
using VerifyFullNP = qsl::Verify<NP, Hip, qsl::NPStateGen<double>, qsl::ProbChecker,
                                 qsl::NPGateChecker>;
// the resizeable pair, as the reference's Default-vs-Resize case
using Resize = qsl::Qubits<qsl::Type::Resize, double>;
using HipResize = qsl::Qubits<qsl::Type::HipResize, double>;
using VerifyResize = qsl::Verify<Resize, HipResize, qsl::DefaultStateGen<double>,
                                 qsl::MeasureChecker, qsl::PostselectChecker, qsl::SampleChecker,
                                 qsl::SampleAllChecker, qsl::ProbChecker, qsl::DefaultGateChecker>;

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "--compile") == 0) {
        static_assert(qsl::Simulator<Hip> && qsl::HasAllGates<Hip>);
        static_assert(qsl::Simulator<HipNP> && qsl::NPSimulator<HipNP>);
        static_assert(qsl::Simulator<HipResize> && qsl::HasAllGates<HipResize>);
        std::printf("verify harness instantiated with Qubits<Hip>, Qubits<HipNP> and "
                    "Qubits<HipResize>\n");
        return 0;
    }
    {
        VerifyFull verify;
        verify.configureState(8);
        configure_statistics(verify);
        // prob sums in another order (FP64 tree vs serial), as the Omp
        // backend's reduction does: the reference bounds the postselected
        // states' Fubini-Study distance by 1e-7 for Default vs Omp
        // (basic-tests.cpp:437-440), since acos near 1 turns a 1e-16 change
        // of the overlap into about 1e-8
        check_suite(verify, "Qubits<Default,double> vs Qubits<Hip,double>, 8 qubits", 1e-7);
        check_gates<VerifyFull, qsl::DefaultGateChecker>(verify);
    }
    {
        VerifyNP verify;
        verify.configureState(8, 5);
        configure_statistics(verify);
        // the same summation-order bound as above
        check_suite(verify, "Qubits<NP,double> vs Qubits<HipNP,double>, 8 qubits, 5 ones", 1e-7);
        check_gates<VerifyNP, qsl::NPGateChecker>(verify);
    }
    {
        VerifyResize verify;
        verify.configureState(8);
        configure_statistics(verify);
        check_suite(verify, "Qubits<Resize,double> vs Qubits<HipResize,double>, 8 qubits", 1e-7);
        check_gates<VerifyResize, qsl::DefaultGateChecker>(verify);
    }
    {
        VerifyFullNP verify;
        verify.configureState(10, 4);
        std::printf("== Qubits<NP,double> vs Qubits<Hip,double>, 10 qubits, 4 ones\n");
        auto prob = verify.check<qsl::ProbChecker>();
        expect(std::abs(prob.mean<double>(1)) < 1e-10, "prob diff0 mean");
        check_gates<VerifyFullNP, qsl::NPGateChecker>(verify);
    }
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
