// qsl/qubits/hip_grad.hpp -- observables, adjoint gradients and batched
// variational circuits for the MI355X backend (no reference counterpart:
// qsl has no observables or gradients).
//
//   qsl::hip::Op / qsl::hip::Pauli   one gate of a gate list / one weighted
//                                    Pauli string (the C ABI's qslhip_op and
//                                    qslhip_pauli)
//   qsl::hip::gate(...), pauli(...)  builders
//   Qubits<Type::Hip, Fp>::expval / expvalGrad   (declared in hip.hpp)
//   qsl::HipBatch                    many circuits, one persistent FP32
//                                    kernel replayed from a hipGraph, FP64
//                                    energies and gradients (RAII owner of a
//                                    qslhip_batch_t)
#ifndef QSL_QUBITS_HIP_GRAD_HPP
#define QSL_QUBITS_HIP_GRAD_HPP

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qslhip.h"

namespace qsl::hip {

using Op = qslhip_op;
using Pauli = qslhip_pauli;

// One gate: kind is a QSLHIP_* gate kind; for two-qubit kinds q0 is the
// control (or the first qubit of swap-like gates) and q1 the target.
inline Op gate(int kind, unsigned q0, unsigned q1 = 0, double angle = 0)
{
    Op o{};
    o.kind = kind;
    o.q0 = q0;
    o.q1 = q1;
    o.pad = 0;
    o.angle = angle;
    return o;
}

// coef times the tensor product of single-qubit Paulis, given as
// {qubit, 'X' | 'Y' | 'Z'} pairs, e.g. pauli(0.5, {{0, 'Z'}, {1, 'Z'}}).
inline Pauli pauli(double coef, std::initializer_list<std::pair<unsigned, char>> factors)
{
    Pauli p{coef, 0, 0};
    for (auto [q, c] : factors) {
        if (q >= 64) throw std::out_of_range("Pauli factor on qubit >= 64");
        const std::uint64_t b = std::uint64_t(1) << q;
        if ((p.xmask | p.zmask) & b) throw std::logic_error("two Pauli factors on one qubit");
        switch (c) {
        case 'X': p.xmask |= b; break;
        case 'Y': p.xmask |= b; p.zmask |= b; break;
        case 'Z': p.zmask |= b; break;
        default: throw std::logic_error(std::string("unknown Pauli factor '") + c + "'");
        }
    }
    return p;
}

inline void check_status(int rc)
{
    if (rc == QSLHIP_OK) return;
    std::string msg = qslhip_last_error();
    if (msg.empty()) msg = qslhip_status_string(rc);
    if (rc == QSLHIP_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

}  // namespace qsl::hip

namespace qsl {

// Batched variational circuits.  Circuit c starts from |0...0> on
// nqubits qubits, applies its ops and measures its observable; run()
// returns every energy and, for every op, dE/d(angle) (0 for fixed gates),
// from one persistent kernel launch captured in a hipGraph at construction.
// States are FP32; energies and gradients accumulate in FP64.  Circuits of
// more than 12 qubits may apply commuting gates in another order (the same
// unitary; results move by FP32 rounding only).
class HipBatch {
public:
    struct Circuit {
        unsigned nqubits = 0;
        std::vector<hip::Op> ops;
        std::vector<hip::Pauli> terms;
    };

    struct Result {
        std::vector<double> energies;  // one per circuit
        std::vector<double> grads;     // one per op, circuits in order
        std::vector<std::uint64_t> offsets;  // circuit c owns grads[offsets[c] .. offsets[c+1])
        // the gradient of circuit c, one entry per op of that circuit
        std::vector<double> gradient(std::size_t c) const
        {
            return std::vector<double>(grads.begin() + (std::ptrdiff_t)offsets.at(c),
                                       grads.begin() + (std::ptrdiff_t)offsets.at(c + 1));
        }
    };

    explicit HipBatch(const std::vector<Circuit> &circuits)
    {
        if (circuits.empty()) throw std::invalid_argument("a batch needs at least one circuit");
        std::vector<std::uint32_t> nq;
        std::vector<hip::Op> ops;
        std::vector<hip::Pauli> terms;
        std::vector<std::uint64_t> toff{0};
        off_.assign(1, 0);
        for (const Circuit &c : circuits) {
            nq.push_back(c.nqubits);
            ops.insert(ops.end(), c.ops.begin(), c.ops.end());
            terms.insert(terms.end(), c.terms.begin(), c.terms.end());
            off_.push_back(ops.size());
            toff.push_back(terms.size());
        }
        hip::check_status(qslhip_batch_create(4, (std::uint32_t)circuits.size(), nq.data(),
                                              off_.data(), ops.data(), toff.data(),
                                              terms.data(), &b_));
    }
    HipBatch(const HipBatch &) = delete;
    HipBatch &operator=(const HipBatch &) = delete;
    HipBatch(HipBatch &&o) noexcept : b_(std::exchange(o.b_, nullptr)), off_(std::move(o.off_)) {}
    HipBatch &operator=(HipBatch &&o) noexcept
    {
        if (this != &o) {
            if (b_) qslhip_batch_destroy(b_);
            b_ = std::exchange(o.b_, nullptr);
            off_ = std::move(o.off_);
        }
        return *this;
    }
    ~HipBatch()
    {
        if (b_) qslhip_batch_destroy(b_);
    }

    std::size_t numCircuits() const { return off_.size() - 1; }
    std::size_t numOps() const { return (std::size_t)off_.back(); }
    // index of circuit c's first op in the flat angle and gradient arrays
    std::size_t opOffset(std::size_t c) const { return (std::size_t)off_.at(c); }

    // With the angles of the op lists (first call) or of the previous run.
    Result run() { return launch(nullptr); }
    // With new angles, one per op in circuit order (numOps() values).
    Result run(const std::vector<double> &angles)
    {
        if (angles.size() != numOps())
            throw std::invalid_argument("HipBatch::run: " + std::to_string(angles.size()) +
                                        " angles for " + std::to_string(numOps()) + " ops");
        return launch(angles.data());
    }

    qslhip_batch_t handle() const { return b_; }

private:
    Result launch(const double *angles)
    {
        Result r;
        r.energies.resize(numCircuits());
        r.grads.resize(numOps());
        r.offsets = off_;
        hip::check_status(qslhip_batch_run(b_, angles, r.energies.data(), r.grads.data()));
        return r;
    }

    qslhip_batch_t b_ = nullptr;
    std::vector<std::uint64_t> off_;
};

}  // namespace qsl

#endif
