// qsl/qubits/hip.hpp -- Qubits<Type::Hip, Fp>: the full state-vector
// simulator of the reference's Default backend (include/qsl/qubits/
// default.hpp:65-713), with the state in MI355X HBM and every gate,
// reduction and sample run by HIP kernels of libqslhip.so.
//
// Member signatures are the reference's; additions are marked (the
// gradient call expvalGrad and the batch class qsl::HipBatch are in
// qsl/qubits/hip_grad.hpp's types).  Gate
// arithmetic is the reference's, operation for operation, so FP64 results
// agree with the Default backend bit for bit (and within 1e-12 by
// contract); prob is accumulated in FP64.
#ifndef QSL_QUBITS_HIP_HPP
#define QSL_QUBITS_HIP_HPP

#include "qsl/qubits/hip_common.hpp"
#include "qsl/qubits/hip_grad.hpp"

namespace qsl {

namespace hipdetail {

// The gate set, observables and sampleAll2 of the full-vector backends
// (Qubits<Hip> and Qubits<HipResize>).
template <std::floating_point Fp>
class FullVector : public Base<Fp> {
protected:
    using Base<Fp>::Base;

public:
    // ---- one-qubit gates (default.hpp:208-333) --------------------------------
    void rotateX(unsigned targ, Fp angle) { this->g1(QSLHIP_ROT_X, targ, angle); }
    void rotateY(unsigned targ, Fp angle) { this->g1(QSLHIP_ROT_Y, targ, angle); }
    void rotateZ(unsigned targ, Fp angle) { this->g1(QSLHIP_ROT_Z, targ, angle); }
    void pauliX(unsigned targ) { this->g1(QSLHIP_PAULI_X, targ); }
    void pauliY(unsigned targ) { this->g1(QSLHIP_PAULI_Y, targ); }
    void pauliZ(unsigned targ) { this->g1(QSLHIP_PAULI_Z, targ); }
    void hadamard(unsigned targ) { this->g1(QSLHIP_HADAMARD, targ); }
    void phase(unsigned targ, Fp angle) { this->g1(QSLHIP_PHASE, targ, angle); }

    // ---- two-qubit gates (default.hpp:354-608) --------------------------------
    void controlNot(unsigned ctrl, unsigned targ) { this->g2(QSLHIP_CNOT, ctrl, targ); }
    void controlY(unsigned ctrl, unsigned targ) { this->g2(QSLHIP_CY, ctrl, targ); }
    void controlZ(unsigned ctrl, unsigned targ) { this->g2(QSLHIP_CZ, ctrl, targ); }
    void controlRotateX(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CROT_X, ctrl, targ, angle);
    }
    void controlRotateY(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CROT_Y, ctrl, targ, angle);
    }
    void controlRotateZ(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CROT_Z, ctrl, targ, angle);
    }
    void controlPhase(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CPHASE, ctrl, targ, angle);
    }
    void controlHadamard(unsigned ctrl, unsigned targ)
    {
        this->g2(QSLHIP_CHADAMARD, ctrl, targ);
    }
    void swap(unsigned q1, unsigned q2) { this->g2(QSLHIP_SWAP, q1, q2); }
    void fswap(unsigned q1, unsigned q2) { this->g2(QSLHIP_FSWAP, q1, q2); }
    void npRotateX(unsigned q1, unsigned q2, Fp angle)
    {
        this->g2(QSLHIP_NP_ROT_X, q1, q2, angle);
    }
    void npRotateY(unsigned q1, unsigned q2, Fp angle)
    {
        this->g2(QSLHIP_NP_ROT_Y, q1, q2, angle);
    }
    void npHadamard(unsigned q1, unsigned q2) { this->g2(QSLHIP_NP_HADAMARD, q1, q2); }

    // ---- observables and adjoint gradients (no reference counterpart) --------
    // <psi|H|psi> of the current state, H = sum of weighted Pauli strings
    // (qsl::hip::pauli), accumulated in FP64.
    double expval(const std::vector<hip::Pauli> &terms) const
    {
        double v = 0;
        hipdetail::check(qslhip_expval(this->h_, terms.data(), terms.size(), &v));
        return v;
    }
    // Applies `ops` to the current state psi0 (the object keeps U psi0) and
    // returns E = <psi|H|psi> and dE/d(angle) of every op (0 for gates
    // without an angle), by the adjoint reverse sweep in Fp with FP64
    // accumulation.  Differentiable kinds: rotateX/Y/Z, phase,
    // controlRotateX/Y/Z, controlPhase, npRotateX/Y.
    std::pair<double, std::vector<double>> expvalGrad(const std::vector<hip::Op> &ops,
                                                      const std::vector<hip::Pauli> &terms)
    {
        double e = 0;
        std::vector<double> g(ops.size());
        hipdetail::check(qslhip_expval_grad(this->h_, ops.data(), ops.size(), terms.data(),
                                            terms.size(), &e, g.data()));
        return {e, std::move(g)};
    }

    // ---- sampleAll2 (measure.cpp:237-267): every index, sorted-merge rule ----
    std::map<std::size_t, std::size_t> sampleAll2(std::size_t nsamples)
    {
        std::vector<Fp> u(nsamples);
        for (auto &x : u) x = this->draw();
        return sampleAll2(u);
    }
    std::map<std::size_t, std::size_t> sampleAll2(const std::vector<Fp> &u)
    {
        std::vector<double> d(u.begin(), u.end());
        std::vector<std::uint64_t> c(std::size_t(1) << this->nqubits_);
        hipdetail::check(qslhip_sample_all2(this->h_, d.data(), d.size(), c.data(), c.size()));
        std::map<std::size_t, std::size_t> m;
        for (std::size_t i = 0; i < c.size(); i++) m[i] = static_cast<std::size_t>(c[i]);
        return m;
    }
};

}  // namespace hipdetail

template <std::floating_point Fp>
class Qubits<Type::Hip, Fp> : public hipdetail::FullVector<Fp> {
    using B = hipdetail::FullVector<Fp>;

    static qslhip_t create(unsigned n)
    {
        qslhip_t h = nullptr;
        hipdetail::check(qslhip_create(n, hipdetail::prec_of<Fp>(), &h));
        return h;
    }

public:
    using Fp_type = Fp;
    static inline const std::string name =
        sizeof(Fp) == 8 ? "Qub<hip,double>" : "Qub<hip,float>";

    // default.hpp:114: |0...0> on nqubits
    Qubits(unsigned nqubits) : B(create(nqubits), nqubits) {}
    // default.hpp:128: from a state vector (size must be a power of two)
    Qubits(const std::vector<complex<Fp>> &state)
        : B(create(checkStateSize(state)), checkStateSize(state))
    {
        this->setState(state);
    }
    Qubits(const Qubits &o) = default;
    Qubits(Qubits &&o) noexcept = default;

    // default.cpp:94-105: throws std::logic_error on a qubit-count mismatch
    void operator=(const Qubits &old) { this->assign(old); }

    // default.cpp:117-122
    void print(std::ostream &os = std::cout) const
    {
        os << "Number of qubits = " << this->nqubits_ << std::endl;
        for (const auto &a : this->getState()) os << a << std::endl;
        os << std::endl;
    }

};

}  // namespace qsl

#endif
