// qsl/concepts.hpp -- the compile-time contract a qsl simulator satisfies.
//
// Same requirements as the reference's concepts (include/qsl/concepts.hpp:
// 50-128), which its Verify and benchmark harnesses are written against;
// Qubits<Type::Hip, Fp> and Qubits<Type::HipNP, Fp> are checked against them
// by static_asserts in tests/cpp.  (The reference's Simulator concept tests
// std::is_constructible<std::size_t>, which checks nothing; here it is the
// intended T(std::size_t).)
#ifndef QSL_CONCEPTS_HPP
#define QSL_CONCEPTS_HPP

#include <concepts>
#include <cstddef>
#include <type_traits>
#include <vector>

#include "qsl/utils/complex.hpp"

namespace qsl {

template <typename T>
concept Simulator =
    std::is_floating_point_v<typename T::Fp_type> &&
    std::is_constructible_v<T, std::size_t> &&
    std::is_constructible_v<T, std::vector<qsl::complex<typename T::Fp_type>>> &&
    std::is_copy_constructible_v<T> && std::is_copy_assignable_v<T> &&
    requires(T t, std::vector<qsl::complex<typename T::Fp_type>> state) {
        t.reset();
        t.getNumQubits();
        t.print();
        t.getState();
        t.setState(state);
    };

template <typename Sim>
concept HasNPGates = requires(Sim sim, unsigned ctrl, unsigned targ,
                              typename Sim::Fp_type param) {
    sim.pauliZ(targ);
    sim.rotateZ(targ, param);
    sim.phase(targ, param);
    sim.controlZ(ctrl, targ);
    sim.controlRotateZ(ctrl, targ, param);
    sim.controlPhase(ctrl, targ, param);
    sim.swap(ctrl, targ);
    sim.fswap(ctrl, targ);
    sim.npRotateX(ctrl, targ, param);
    sim.npRotateY(ctrl, targ, param);
    sim.npHadamard(ctrl, targ);
};

template <typename Sim>
concept HasNonNPGates = requires(Sim sim, unsigned ctrl, unsigned targ,
                                 typename Sim::Fp_type param) {
    sim.pauliX(targ);
    sim.pauliY(targ);
    sim.rotateX(targ, param);
    sim.rotateY(targ, param);
    sim.hadamard(targ);
    sim.controlNot(ctrl, targ);
    sim.controlY(ctrl, targ);
    sim.controlRotateX(ctrl, targ, param);
    sim.controlRotateY(ctrl, targ, param);
    sim.controlHadamard(ctrl, targ);
};

template <typename Sim>
concept HasMeasurement = requires(Sim sim, unsigned targ, unsigned outcome,
                                  std::size_t nsamples) {
    { sim.measure(targ) } -> std::same_as<int>;
    { sim.measureAll() } -> std::same_as<std::size_t>;
    { sim.prob(targ, outcome) } -> std::convertible_to<double>;
    { sim.postselect(targ, outcome) } -> std::convertible_to<double>;
    { sim.sample(targ, nsamples) };
    { sim.sampleAll(nsamples) };
};

template <typename Sim>
concept HasAllGates = HasNPGates<Sim> && HasNonNPGates<Sim>;

template <typename Sim>
concept NPSimulator = HasMeasurement<Sim> && HasNPGates<Sim>;

template <typename Sim>
concept NonNPSimulator = HasMeasurement<Sim> && HasNonNPGates<Sim>;

}  // namespace qsl

#endif
