// ref_shim.cpp -- C entry points onto the REAL reference library.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
// with the reference's own sources (read in place under /root/reference;
// nothing is copied into this repository) into oracle/_ref/libqslref.so.
// Tests use it to pin the C oracle bit for bit against qsl itself, and
// bench.py may time its Omp backend as the "reference" CPU baseline.
//
// Handles wrap qsl::Qubits<Type, Fp> for Type in {Default, Omp, NP, OmpNP,
// Resize} (include/qsl/qubits.hpp:53-59) and Fp in {float, double}.  Gate kinds use
// the numbering of oracle/qsl_oracle.h.
#include <qsl/qubits.hpp>
#include <qsl/utils/misc.hpp>

#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct Base {
    virtual ~Base() = default;
    virtual void set_state(const void *p) = 0;
    virtual void get_state(void *p) const = 0;
    virtual int gate1(int kind, unsigned t, double angle) = 0;
    virtual int gate2(int kind, unsigned a, unsigned b, double angle) = 0;
    virtual double prob(unsigned t, unsigned o) const = 0;
    virtual double postselect(unsigned t, unsigned o) = 0;
    virtual int measure(unsigned t) = 0;
    virtual std::uint64_t measure_all() = 0;
    virtual void sample_all(std::uint64_t n, std::uint64_t *dense) = 0;
    virtual void sample(unsigned t, std::uint64_t n, std::uint64_t *counts) = 0;
    virtual int sample_all2(std::uint64_t n, std::uint64_t *dense) = 0;
    virtual unsigned nqubits() const = 0;
    // Resize backend only (return -1 elsewhere)
    virtual int set_state_n(const void *p, unsigned n) = 0;
    virtual int add_qubit(unsigned t) = 0;
    virtual int measure_out(unsigned t) = 0;
    virtual double postselect_out(unsigned t, unsigned o) = 0;
    virtual int reallocate() = 0;
};

template <typename Sim>
concept HasFull = qsl::HasNonNPGates<Sim>;

template <typename Sim>
concept HasResize = requires(Sim q, unsigned t) {
    q.addQubit(t);
    q.measureOut(t);
    q.postselectOut(t, t);
    q.reallocateState();
};

template <typename Sim>
concept HasSampleAll2 = requires(Sim q, std::size_t n) { q.sampleAll2(n); };

template <typename Sim, typename Fp>
struct Wrap final : Base {
    Sim q;
    template <typename... A>
    explicit Wrap(A... a) : q(a...) {}
    unsigned nqubits() const override { return q.getNumQubits(); }
    void set_state(const void *p) override {
        std::size_t dim = std::size_t(1) << q.getNumQubits();
        std::vector<qsl::complex<Fp>> v(dim);
        std::memcpy(v.data(), p, dim * sizeof(qsl::complex<Fp>));
        q.setState(v);
    }
    void get_state(void *p) const override {
        auto v = q.getState();
        std::memcpy(p, v.data(), v.size() * sizeof(qsl::complex<Fp>));
    }
    int gate1(int kind, unsigned t, double angle) override {
        Fp a = static_cast<Fp>(angle);
        switch (kind) {
        case 2: q.pauliZ(t); return 0;
        case 4: q.phase(t, a); return 0;
        case 7: q.rotateZ(t, a); return 0;
        default: break;
        }
        if constexpr (HasFull<Sim>) {
            switch (kind) {
            case 0: q.pauliX(t); return 0;
            case 1: q.pauliY(t); return 0;
            case 3: q.hadamard(t); return 0;
            case 5: q.rotateX(t, a); return 0;
            case 6: q.rotateY(t, a); return 0;
            default: break;
            }
        }
        return -1;
    }
    int gate2(int kind, unsigned x, unsigned y, double angle) override {
        Fp a = static_cast<Fp>(angle);
        switch (kind) {
        case 18: q.controlZ(x, y); return 0;
        case 21: q.controlRotateZ(x, y, a); return 0;
        case 22: q.controlPhase(x, y, a); return 0;
        case 24: q.swap(x, y); return 0;
        case 25: q.fswap(x, y); return 0;
        case 26: q.npRotateX(x, y, a); return 0;
        case 27: q.npRotateY(x, y, a); return 0;
        case 28: q.npHadamard(x, y); return 0;
        default: break;
        }
        if constexpr (HasFull<Sim>) {
            switch (kind) {
            case 16: q.controlNot(x, y); return 0;
            case 17: q.controlY(x, y); return 0;
            case 19: q.controlRotateX(x, y, a); return 0;
            case 20: q.controlRotateY(x, y, a); return 0;
            case 23: q.controlHadamard(x, y); return 0;
            default: break;
            }
        }
        return -1;
    }
    double prob(unsigned t, unsigned o) const override {
        return static_cast<double>(q.prob(t, o));
    }
    double postselect(unsigned t, unsigned o) override {
        return static_cast<double>(q.postselect(t, o));
    }
    int measure(unsigned t) override { return q.measure(t); }
    std::uint64_t measure_all() override { return q.measureAll(); }
    void sample_all(std::uint64_t n, std::uint64_t *dense) override {
        auto m = q.sampleAll(n);
        for (auto &kv : m) dense[kv.first] += kv.second;
    }
    void sample(unsigned t, std::uint64_t n, std::uint64_t *counts) override {
        auto v = q.sample(t, n);
        counts[0] = v[0];
        counts[1] = v[1];
    }
    int sample_all2(std::uint64_t n, std::uint64_t *dense) override {
        if constexpr (HasSampleAll2<Sim>) {
            auto m = q.sampleAll2(n);
            for (auto &kv : m) dense[kv.first] += kv.second;
            return 0;
        }
        return -1;
    }
    int set_state_n(const void *p, unsigned n) override {
        if constexpr (HasResize<Sim>) {
            std::size_t dim = std::size_t(1) << n;
            std::vector<qsl::complex<Fp>> v(dim);
            std::memcpy(v.data(), p, dim * sizeof(qsl::complex<Fp>));
            q.setState(v);
            return 0;
        }
        return -1;
    }
    int add_qubit(unsigned t) override {
        if constexpr (HasResize<Sim>) {
            q.addQubit(t);
            return 0;
        }
        return -1;
    }
    int measure_out(unsigned t) override {
        if constexpr (HasResize<Sim>) return q.measureOut(t);
        return -1;
    }
    double postselect_out(unsigned t, unsigned o) override {
        if constexpr (HasResize<Sim>) return static_cast<double>(q.postselectOut(t, o));
        return -1;
    }
    int reallocate() override {
        if constexpr (HasResize<Sim>) {
            q.reallocateState();
            return 0;
        }
        return -1;
    }
};

template <qsl::Type T>
Base *make(int prec, unsigned n, unsigned extra)
{
    if constexpr (T == qsl::Type::Default || T == qsl::Type::Resize) {
        if (prec == 4) return new Wrap<qsl::Qubits<T, float>, float>(n);
        return new Wrap<qsl::Qubits<T, double>, double>(n);
    } else {
        if (prec == 4) return new Wrap<qsl::Qubits<T, float>, float>(n, extra);
        return new Wrap<qsl::Qubits<T, double>, double>(n, extra);
    }
}

} // namespace

extern "C" {

// type: 0 Default, 1 Omp (extra = nthreads), 2 NP (extra = nones),
// 3 OmpNP (extra = nones; the reference's default 4 threads), 4 Resize
void *ref_create(int type, int prec, unsigned n, unsigned extra)
{
    try {
        switch (type) {
        case 0: return make<qsl::Type::Default>(prec, n, extra);
        case 1: return make<qsl::Type::Omp>(prec, n, extra);
        case 2: return make<qsl::Type::NP>(prec, n, extra);
        case 3: return make<qsl::Type::OmpNP>(prec, n, extra);
        case 4: return make<qsl::Type::Resize>(prec, n, extra);
        default: return nullptr;
        }
    } catch (...) {
        return nullptr;
    }
}

void ref_destroy(void *h) { delete static_cast<Base *>(h); }

int ref_set_state(void *h, const void *p)
{
    try {
        static_cast<Base *>(h)->set_state(p);
        return 0;
    } catch (...) {
        return -1;
    }
}

void ref_get_state(void *h, void *p) { static_cast<Base *>(h)->get_state(p); }
int ref_gate1(void *h, int kind, unsigned t, double a)
{
    return static_cast<Base *>(h)->gate1(kind, t, a);
}
int ref_gate2(void *h, int kind, unsigned x, unsigned y, double a)
{
    return static_cast<Base *>(h)->gate2(kind, x, y, a);
}
double ref_prob(void *h, unsigned t, unsigned o)
{
    return static_cast<Base *>(h)->prob(t, o);
}
double ref_postselect(void *h, unsigned t, unsigned o)
{
    return static_cast<Base *>(h)->postselect(t, o);
}
int ref_measure(void *h, unsigned t) { return static_cast<Base *>(h)->measure(t); }
std::uint64_t ref_measure_all(void *h)
{
    return static_cast<Base *>(h)->measure_all();
}
void ref_sample_all(void *h, std::uint64_t n, std::uint64_t *dense)
{
    static_cast<Base *>(h)->sample_all(n, dense);
}
void ref_sample(void *h, unsigned t, std::uint64_t n, std::uint64_t *counts)
{
    static_cast<Base *>(h)->sample(t, n, counts);
}
int ref_sample_all2(void *h, std::uint64_t n, std::uint64_t *dense)
{
    return static_cast<Base *>(h)->sample_all2(n, dense);
}
unsigned ref_nqubits(void *h) { return static_cast<Base *>(h)->nqubits(); }
int ref_set_state_n(void *h, const void *p, unsigned n)
{
    try {
        return static_cast<Base *>(h)->set_state_n(p, n);
    } catch (...) {
        return -2;
    }
}
int ref_add_qubit(void *h, unsigned t) { return static_cast<Base *>(h)->add_qubit(t); }
int ref_measure_out(void *h, unsigned t) { return static_cast<Base *>(h)->measure_out(t); }
double ref_postselect_out(void *h, unsigned t, unsigned o)
{
    return static_cast<Base *>(h)->postselect_out(t, o);
}
int ref_reallocate(void *h) { return static_cast<Base *>(h)->reallocate(); }

// Utilities from src/utils/misc.cpp
unsigned ref_choose(unsigned n, unsigned k) { return qsl::choose(n, k); }
unsigned ref_hamming_weight(std::uint64_t x) { return qsl::hammingWeight(x); }
std::uint64_t ref_next(std::uint64_t x)
{
    std::size_t v = x;
    qsl::next(v);
    return v;
}
std::uint64_t ref_generate_lookup(unsigned len, unsigned ones, std::uint64_t *out,
                                  std::uint64_t cap)
{
    auto v = qsl::generateLookup(len, ones);
    std::uint64_t m = v.size() < cap ? v.size() : cap;
    for (std::uint64_t i = 0; i < m; i++) out[i] = v[i];
    return v.size();
}

} // extern "C"
