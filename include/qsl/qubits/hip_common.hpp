// qsl/qubits/hip_common.hpp -- the part of Qubits<Type::Hip, Fp> and
// Qubits<Type::HipNP, Fp> they share: ownership of one device state through
// the C ABI of libqslhip.so, the state accessors and the measurement API
// (include/qsl/qubits/default.hpp:114-186, 627-712).
//
// Errors follow the reference: std::logic_error for its logic_error cases
// (size mismatch, bad nones, non-NP state); std::out_of_range for qubit
// indices (the reference leaves them undefined); std::runtime_error for
// device failures.
#ifndef QSL_QUBITS_HIP_COMMON_HPP
#define QSL_QUBITS_HIP_COMMON_HPP

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <iostream>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qslhip.h"
#include "qsl/utils/complex.hpp"
#include "qsl/utils/quantum.hpp"

namespace qsl::hipdetail {

inline void check(int rc)
{
    if (rc == QSLHIP_OK) return;
    std::string msg = qslhip_last_error();
    if (msg.empty()) msg = qslhip_status_string(rc);
    switch (rc) {
    case QSLHIP_ESIZE:
    case QSLHIP_ENOTNP:
    case QSLHIP_EKIND:
        throw std::logic_error(msg);
    case QSLHIP_EINVAL:
        throw std::out_of_range(msg);
    default:
        throw std::runtime_error(msg);
    }
}

template <std::floating_point Fp>
constexpr int prec_of()
{
    return sizeof(Fp) == 8 ? 8 : 4;
}

// Shared implementation.  Derived classes add their gate sets.
template <std::floating_point Fp>
class Base {
protected:
    qslhip_t h_ = nullptr;
    unsigned nqubits_ = 0;
    std::mt19937_64 gen_;

    explicit Base(qslhip_t h, unsigned n) : h_(h), nqubits_(n), gen_(std::random_device{}()) {}

    Base(const Base &o) : nqubits_(o.nqubits_), gen_(o.gen_)
    {
        check(qslhip_clone(o.h_, &h_));
    }
    Base(Base &&o) noexcept
        : h_(std::exchange(o.h_, nullptr)), nqubits_(o.nqubits_), gen_(o.gen_) {}

    ~Base()
    {
        if (h_) qslhip_destroy(h_);
    }

    void assign(const Base &o)
    {
        check(qslhip_assign(h_, o.h_));
    }

    // One uniform on [0, 1) in Fp, as the reference's Random<Fp>(0, 1).
    Fp draw() { return static_cast<Fp>((gen_() >> 11) * 0x1.0p-53); }

    void g1(int kind, unsigned t, double angle = 0)
    {
        check(qslhip_gate1(h_, kind, t, angle));
    }
    void g2(int kind, unsigned a, unsigned b, double angle = 0)
    {
        check(qslhip_gate2(h_, kind, a, b, angle));
    }

public:
    using Fp_type = Fp;

    unsigned getNumQubits() const { return nqubits_; }

    void reset() { check(qslhip_reset(h_)); }

    std::vector<complex<Fp>> getState() const
    {
        std::vector<complex<Fp>> v(std::size_t(1) << nqubits_);
        check(qslhip_get_state(h_, v.data(), v.size()));
        return v;
    }

    void setState(const std::vector<complex<Fp>> &state)
    {
        check(qslhip_set_state(h_, state.data(), state.size()));
    }

    void setBasisState(std::size_t index) { check(qslhip_set_basis(h_, index)); }

    // Reseed the uniform stream behind measure/measureAll/sample*.
    void seed(std::uint64_t s) { gen_.seed(s); }

    // Gate lists in one call (no reference counterpart).
    void apply(const std::vector<qslhip_op> &ops, int mode = QSLHIP_MODE_FUSED)
    {
        check(qslhip_apply_ops(h_, ops.data(), ops.size(), mode));
    }

    // ---- measurement (src/qubits/default/measure.cpp) -----------------------
    Fp prob(unsigned targ, unsigned outcome) const
    {
        double p = 0;
        check(qslhip_prob(h_, targ, outcome, &p));
        return static_cast<Fp>(p);
    }

    Fp postselect(unsigned targ, unsigned outcome)
    {
        double p = 0;
        check(qslhip_postselect(h_, targ, outcome, &p));
        return static_cast<Fp>(p);
    }

    int measure(unsigned targ) { return measure(targ, draw()); }
    int measure(unsigned targ, Fp u)
    {
        int o = 0;
        check(qslhip_measure(h_, targ, static_cast<double>(u), &o));
        return o;
    }

    std::size_t measureAll() { return measureAll(draw()); }
    std::size_t measureAll(Fp u)
    {
        std::uint64_t o = 0;
        check(qslhip_measure_all(h_, static_cast<double>(u), &o));
        return static_cast<std::size_t>(o);
    }

    std::vector<std::size_t> sample(unsigned targ, std::size_t nsamples)
    {
        std::vector<Fp> u(nsamples);
        for (auto &x : u) x = draw();
        return sample(targ, u);
    }
    std::vector<std::size_t> sample(unsigned targ, const std::vector<Fp> &u)
    {
        std::vector<double> d(u.begin(), u.end());
        std::uint64_t c[2] = {0, 0};
        check(qslhip_sample(h_, targ, d.data(), d.size(), c));
        return {static_cast<std::size_t>(c[0]), static_cast<std::size_t>(c[1])};
    }

    std::map<std::size_t, std::size_t> sampleAll(std::size_t nsamples)
    {
        std::vector<Fp> u(nsamples);
        for (auto &x : u) x = draw();
        return sampleAll(u);
    }
    std::map<std::size_t, std::size_t> sampleAll(const std::vector<Fp> &u)
    {
        std::vector<double> d(u.begin(), u.end());
        std::vector<std::uint64_t> out(d.size());
        check(qslhip_sample_all(h_, d.data(), d.size(), out.data()));
        std::map<std::size_t, std::size_t> m;
        for (auto x : out) m[static_cast<std::size_t>(x)]++;
        return m;
    }

    // Waits until every queued operation on this object has finished on the
    // device (no reference counterpart: the reference's members return when
    // done, these return once queued).  Timing harnesses call it around
    // their clocks.
    void sync() const { check(qslhip_sync(h_)); }

    // The opaque handle, for callers mixing the C ABI in.
    qslhip_t handle() const { return h_; }
};

}  // namespace qsl::hipdetail

#endif
