// cplx.h -- the amplitude type shared by host code, the kernels and the
// run-time compiled passes (no includes: hipRTC builds it as is).
#pragma once

namespace qslhip {

// Complex amplitude of precision T: {re, im}, the layout of qsl::complex<T>
// (include/qsl/utils/complex.hpp in the reference).
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};

}  // namespace qslhip
