// qsl/qubits.hpp -- standalone entry header of the MI355X backend.
//
// Mirrors the reference's plugin point (include/qsl/qubits.hpp:53-71): the
// backend tag enum, with the two new tags Hip and HipNP appended, and the
// primary template Qubits<Type, Fp> that each backend specialises.  Inside
// the reference tree the same two tags are added to its own enum and the
// two headers below are included from there (INTEGRATION.md shows the
// three-line change); this file lets the backend build and be tested
// without the reference.
#ifndef QSL_QUBITS_HPP
#define QSL_QUBITS_HPP

#include <concepts>

#include "qsl/utils/complex.hpp"
#include "qsl/utils/quantum.hpp"

namespace qsl {

enum class Type {
    Default,  // reference: serial CPU
    Omp,      // reference: OpenMP CPU
    NP,       // reference: number preserving CPU
    OmpNP,    // reference: number preserving OpenMP CPU
    Resize,   // reference: resizeable CPU
    Hip,      // this backend: full state vector on one MI355X
    HipNP,    // this backend: compact number-preserving vector on one MI355X
    HipResize,  // this backend: resizeable full vector on one MI355X
};

template <Type T = Type::Default, std::floating_point Fp = double>
class Qubits;

}  // namespace qsl

#include "qsl/qubits/hip.hpp"
#include "qsl/qubits/hip_np.hpp"
#include "qsl/qubits/hip_resize.hpp"
#include "qsl/concepts.hpp"

namespace qsl {

// Member-pointer typedefs used by the reference's harnesses
// (include/qsl/qubits.hpp:87-94).
template <typename Sim, typename... Args>
using Gate = void (Sim::*)(Args...);
template <typename Sim, typename... Args>
using OneQubitGate = Gate<Sim, unsigned, Args...>;
template <typename Sim, typename... Args>
using TwoQubitGate = Gate<Sim, unsigned, unsigned, Args...>;

// fubiniStudy between two simulators (include/qsl/qubits.hpp:97-102)
template <Simulator S1, Simulator S2>
    requires std::is_same_v<typename S1::Fp_type, typename S2::Fp_type>
typename S1::Fp_type fubiniStudy(const S1 &s1, const S2 &s2)
{
    return qsl::fubiniStudy(s1.getState(), s2.getState());
}

}  // namespace qsl

#endif
