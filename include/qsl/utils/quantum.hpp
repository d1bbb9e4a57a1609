// qsl/utils/quantum.hpp -- host-side vector utilities of the standalone
// headers: the functions of the reference's src/utils/quantum.cpp and
// src/utils/misc.cpp that the simulator API and its tests use, header-only.
// makeRandomState / makeRandomNPState take an explicit seed (the
// reference's draw from std::random_device, src/utils/random.cpp:48-108).
#ifndef QSL_QUANTUM_HPP
#define QSL_QUANTUM_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "qsl/utils/complex.hpp"

namespace qsl {

// Gosper's next string with the same number of ones (src/utils/misc.cpp:44-52)
inline void next(std::size_t &x)
{
    std::size_t t = x | (x - 1);
    x = (t + 1) | (((~t & (0 - ~t)) - 1) >> (__builtin_ctzll(x) + 1));
}

inline unsigned hammingWeight(std::size_t n)
{
    return static_cast<unsigned>(__builtin_popcountll(n));
}

// Binomial coefficient in 64 bits (the reference's 32-bit version,
// src/utils/misc.cpp:65-73, overflows at C(32,16)).
inline std::uint64_t choose(unsigned n, unsigned k)
{
    if (k > n) return 0;
    std::uint64_t a = 1;
    for (unsigned i = 1; i <= k; i++) {
        a *= (n - k + i);
        a /= i;
    }
    return a;
}

// All len-bit strings with `ones` ones, increasing (src/utils/misc.cpp:75-90)
inline std::vector<std::size_t> generateLookup(unsigned len, unsigned ones)
{
    std::vector<std::size_t> list;
    if (ones == 0) return {0};
    std::size_t x = (std::size_t(1) << ones) - 1;
    std::size_t end = x << (len - ones);
    while (x <= end) {
        list.push_back(x);
        next(x);
    }
    return list;
}

template <std::floating_point Fp>
complex<Fp> innerProduct(const std::vector<complex<Fp>> &v,
                         const std::vector<complex<Fp>> &w)
{
    if (v.size() != w.size())
        throw std::logic_error("Cannot compute inner product between vectors of "
                               "different sizes: " + std::to_string(v.size()) +
                               " and " + std::to_string(w.size()));
    complex<Fp> r{0, 0};
    for (std::size_t n = 0; n < v.size(); n++) {
        r.real += v[n].real * w[n].real + v[n].imag * w[n].imag;
        r.imag += v[n].real * w[n].imag - v[n].imag * w[n].real;
    }
    return r;
}

template <std::floating_point Fp>
Fp norm(const std::vector<complex<Fp>> &v)
{
    Fp ip = 0;
    for (const auto &a : v) ip += a.real * a.real + a.imag * a.imag;
    return std::sqrt(ip);
}

inline double acosSafe(double value)
{
    if (value <= -1.0) return M_PI;
    if (value >= 1.0) return 0;
    return std::acos(value);
}

// Fubini-Study distance acos(|<v,w>| / (|v||w|)) (src/utils/quantum.cpp:84-91)
template <std::floating_point Fp>
Fp fubiniStudy(const std::vector<complex<Fp>> &v, const std::vector<complex<Fp>> &w)
{
    Fp num = abs(innerProduct(v, w));
    Fp den = norm(v) * norm(w);
    return static_cast<Fp>(acosSafe(num / den));
}

template <std::floating_point Fp>
Fp normalise(std::vector<complex<Fp>> &state)
{
    Fp f = norm(state);
    for (auto &a : state) {
        a.real /= f;
        a.imag /= f;
    }
    return f;
}

template <std::floating_point Fp>
unsigned checkStateSize(const std::vector<complex<Fp>> &state)
{
    std::size_t n = state.size();
    if (n == 0 || (n & (n - 1)) != 0)
        throw std::logic_error("State vector size is not a power of two");
    return static_cast<unsigned>(__builtin_ctzll(n));
}

template <std::floating_point Fp>
unsigned checkStateNP(const std::vector<complex<Fp>> &state)
{
    unsigned nones = 0;
    bool found = false;
    for (std::size_t i = 0; i < state.size(); i++) {
        Fp a = state[i].real * state[i].real + state[i].imag * state[i].imag;
        if (a != 0) {
            unsigned w = hammingWeight(i);
            if (!found) {
                nones = w;
                found = true;
            } else if (nones != w) {
                throw std::logic_error("Input state is not number preserving.");
            }
        }
    }
    return nones;
}

// Random state: re, im uniform on (-1, 1), then normalised.
template <std::floating_point Fp>
std::vector<complex<Fp>> makeRandomState(unsigned nqubits, std::uint64_t seed)
{
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> d(-1, 1);
    std::vector<complex<Fp>> st(std::size_t(1) << nqubits);
    for (auto &a : st) {
        a.real = static_cast<Fp>(d(gen));
        a.imag = static_cast<Fp>(d(gen));
    }
    normalise(st);
    return st;
}

template <std::floating_point Fp>
std::vector<complex<Fp>> makeRandomNPState(unsigned nqubits, unsigned nones,
                                           std::uint64_t seed)
{
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> d(-1, 1);
    std::vector<complex<Fp>> st(std::size_t(1) << nqubits);
    for (std::size_t x : generateLookup(nqubits, nones)) {
        st[x].real = static_cast<Fp>(d(gen));
        st[x].imag = static_cast<Fp>(d(gen));
    }
    normalise(st);
    return st;
}

}  // namespace qsl

#endif
