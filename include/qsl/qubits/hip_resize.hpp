// qsl/qubits/hip_resize.hpp -- Qubits<Type::HipResize, Fp>: the resizeable
// simulator of the reference's Resize backend (include/qsl/qubits/resize.hpp)
// on one MI355X.  Qubits are added (addQubit, appendQubit) and measured or
// postselected out (measureOut, postselectOut) on the device; the allocation
// doubles when a qubit no longer fits, as the reference's vector does, and
// shrinks only in reallocateState.  The full gate set and measurement API
// are the Qubits<Type::Hip> ones.
#ifndef QSL_QUBITS_HIP_RESIZE_HPP
#define QSL_QUBITS_HIP_RESIZE_HPP

#include "qsl/qubits/hip.hpp"

namespace qsl {

template <std::floating_point Fp>
class Qubits<Type::HipResize, Fp> : public hipdetail::FullVector<Fp> {
    using B = hipdetail::FullVector<Fp>;

    static qslhip_t create(unsigned n)
    {
        qslhip_t h = nullptr;
        hipdetail::check(qslhip_create_resize(n, hipdetail::prec_of<Fp>(), &h));
        return h;
    }

    void refresh()
    {
        unsigned n = 0;
        hipdetail::check(qslhip_info(this->h_, &n, nullptr, nullptr, nullptr));
        this->nqubits_ = n;
    }

public:
    using Fp_type = Fp;
    static inline const std::string name =
        sizeof(Fp) == 8 ? "Qub<hipresize,double>" : "Qub<hipresize,float>";

    // resize.cpp:39-46
    Qubits(unsigned nqubits) : B(create(nqubits), nqubits) {}
    // resize.cpp:48-56
    Qubits(const std::vector<complex<Fp>> &state)
        : B(create(checkStateSize(state)), checkStateSize(state))
    {
        setState(state);
    }
    Qubits(const Qubits &o) = default;
    Qubits(Qubits &&o) noexcept = default;

    // The reference leaves assignment to the compiler (resize.hpp's
    // operator= is commented out), which copies the vector whatever its
    // size: the same here, through a device copy.
    Qubits &operator=(const Qubits &old)
    {
        if (this != &old) {
            Qubits tmp{old};
            std::swap(this->h_, tmp.h_);
            this->nqubits_ = tmp.nqubits_;
        }
        return *this;
    }

    // resize.cpp:70-81: a state of any (power-of-two) size re-initialises
    // the object, its qubit count included
    void setState(const std::vector<complex<Fp>> &state)
    {
        checkStateSize(state);
        hipdetail::check(qslhip_resize_set_state(this->h_, state.data(), state.size()));
        refresh();
    }

    // resize.cpp:107-147: a new qubit in |0> at index targ (qubits k >= targ
    // move to k + 1)
    void addQubit(unsigned targ)
    {
        hipdetail::check(qslhip_add_qubit(this->h_, targ));
        refresh();
    }
    // resize.cpp:101-105
    void appendQubit() { addQubit(this->nqubits_); }
    // resize.cpp:93-98: the allocation trimmed to the current state
    void reallocateState() { hipdetail::check(qslhip_reallocate(this->h_)); }

    // measure.cpp:157-190: measure, then remove the qubit
    int measureOut(unsigned targ) { return measureOut(targ, this->draw()); }
    int measureOut(unsigned targ, Fp u)
    {
        int o = 0;
        hipdetail::check(qslhip_measure_out(this->h_, targ, static_cast<double>(u), &o));
        refresh();
        return o;
    }
    // measure.cpp:235-250: postselect, then remove the qubit; returns the
    // probability of the outcome
    Fp postselectOut(unsigned targ, unsigned outcome)
    {
        double p = 0;
        hipdetail::check(qslhip_postselect_out(this->h_, targ, outcome, &p));
        refresh();
        return static_cast<Fp>(p);
    }

    // resize.cpp:181-193
    void print(std::ostream &os = std::cout) const
    {
        os << "Number of qubits = " << this->nqubits_ << std::endl;
        for (const auto &a : this->getState()) os << a << std::endl;
        os << std::endl;
    }
};

}  // namespace qsl

#endif
