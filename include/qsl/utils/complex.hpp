// qsl/utils/complex.hpp -- amplitude type of the standalone qsl headers.
//
// Layout contract of qsl::complex<Fp> (reference:
// include/qsl/utils/complex.hpp:40-49): two Fp members, real then imag, no
// padding, so a std::vector<complex<Fp>> is the interleaved array that the
// C ABI (include/qslhip.h) copies to and from the device.
#ifndef QSL_COMPLEX_HPP
#define QSL_COMPLEX_HPP

#include <cmath>
#include <concepts>
#include <ostream>

namespace qsl {

template <std::floating_point Fp = double>
struct complex {
    Fp real{0};
    Fp imag{0};
};

static_assert(sizeof(complex<double>) == 16 && sizeof(complex<float>) == 8,
              "qsl::complex must be two packed Fp");

template <std::floating_point Fp>
Fp abs(const complex<Fp> &a)
{
    return std::sqrt(a.real * a.real + a.imag * a.imag);
}

template <std::floating_point Fp>
complex<Fp> operator-(const complex<Fp> &a, const complex<Fp> &b)
{
    return complex<Fp>{a.real - b.real, a.imag - b.imag};
}

}  // namespace qsl

// "(re, im)" in the stream's current format (reference: src/utils/print.cpp:95-100)
template <std::floating_point Fp>
std::ostream &operator<<(std::ostream &os, qsl::complex<Fp> v)
{
    return os << "(" << v.real << ", " << v.imag << ")";
}

#endif
