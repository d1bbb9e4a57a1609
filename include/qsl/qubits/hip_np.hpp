// qsl/qubits/hip_np.hpp -- Qubits<Type::HipNP, Fp>: the number-preserving
// simulator of the reference's NP backend (include/qsl/qubits/np.hpp:50-253)
// on one MI355X.  The device holds only the C(n, nones) amplitudes of
// weight nones, in the reference's own enumeration (colex) order, so a
// 32-qubit half-filled state takes 9.6 GB instead of the reference's 64 GiB
// vector plus lookup tables.  getState/setState still exchange the full
// 2^n vector, as the reference's do.
#ifndef QSL_QUBITS_HIP_NP_HPP
#define QSL_QUBITS_HIP_NP_HPP

#include "qsl/qubits/hip_common.hpp"

namespace qsl {

template <std::floating_point Fp>
class Qubits<Type::HipNP, Fp> : public hipdetail::Base<Fp> {
    using B = hipdetail::Base<Fp>;

    static qslhip_t create(unsigned n, unsigned nones)
    {
        qslhip_t h = nullptr;
        hipdetail::check(qslhip_create_np(n, nones, hipdetail::prec_of<Fp>(), &h));
        return h;
    }

public:
    using Fp_type = Fp;
    static inline const std::string name =
        sizeof(Fp) == 8 ? "Qub<hipnp,double>" : "Qub<hipnp,float>";

    // np.hpp:107 / np.cpp:42-55: throws unless 1 <= nones < nqubits
    Qubits(unsigned nqubits, unsigned nones = 1)
        : B(create(nqubits, nones), nqubits) {}
    // np.hpp:115: nqubits from the size, nones from the state
    Qubits(const std::vector<complex<Fp>> &state)
        : B(create(checkStateSize(state), checkStateNP(state)), checkStateSize(state))
    {
        this->setState(state);
    }
    Qubits(const Qubits &o) = default;
    Qubits(Qubits &&o) noexcept = default;

    void operator=(const Qubits &old) { this->assign(old); }

    unsigned getNumOnes() const
    {
        unsigned k = 0;
        hipdetail::check(qslhip_info(this->h_, nullptr, &k, nullptr, nullptr));
        return k;
    }
    // np.cpp:155-165: validates, then resets to the lowest weight-nones string
    void setNumOnes(unsigned nones) { hipdetail::check(qslhip_set_num_ones(this->h_, nones)); }

    // np.cpp:184-189
    void print(std::ostream &os = std::cout) const
    {
        os << "Number of qubits = " << this->nqubits_ << std::endl;
        os << "Number of ones = " << getNumOnes() << std::endl;
        for (const auto &a : this->getState()) os << a << std::endl;
        os << std::endl;
    }

    // Compact device amplitudes (rank order), no reference counterpart.
    std::vector<complex<Fp>> getCompactState() const
    {
        std::uint64_t namps = 0;
        hipdetail::check(qslhip_info(this->h_, nullptr, nullptr, nullptr, &namps));
        std::vector<complex<Fp>> v(namps);
        hipdetail::check(qslhip_get_amps(this->h_, v.data(), 0, namps));
        return v;
    }

    // ---- number-preserving gates (np.hpp:158-208) ---------------------------
    void phase(unsigned targ, Fp angle) { this->g1(QSLHIP_PHASE, targ, angle); }
    void pauliZ(unsigned targ) { this->g1(QSLHIP_PAULI_Z, targ); }
    void rotateZ(unsigned targ, Fp angle) { this->g1(QSLHIP_ROT_Z, targ, angle); }
    void controlZ(unsigned ctrl, unsigned targ) { this->g2(QSLHIP_CZ, ctrl, targ); }
    void controlRotateZ(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CROT_Z, ctrl, targ, angle);
    }
    void controlPhase(unsigned ctrl, unsigned targ, Fp angle)
    {
        this->g2(QSLHIP_CPHASE, ctrl, targ, angle);
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

    // sampleAll2 (np/measure.cpp:231-256): every weight-nones index
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
        for (std::size_t x : generateLookup(this->nqubits_, getNumOnes()))
            m[x] = static_cast<std::size_t>(c[x]);
        return m;
    }
};

}  // namespace qsl

#endif
