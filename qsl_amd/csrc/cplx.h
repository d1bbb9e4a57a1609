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

template <typename T>
struct cvec2;
template <>
struct cvec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct cvec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};

// One amplitude loaded or stored as one vector access; with S, as a
// streaming (nontemporal) access.  States far larger than the caches are
// read and written once per pass, and on MI355X streaming accesses move
// them 4-11% faster than cached ones (in-place pair update of a 16 GiB
// FP64 state: 5.2-5.4 -> 5.5-6.4 TB/s by target).
template <bool S, typename T>
__device__ __forceinline__ cplx<T> amp_ld(const cplx<T> *p)
{
    if constexpr (S) {
        typedef typename cvec2<T>::type V;
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
        cplx<T> a;
        a.re = v.x;
        a.im = v.y;
        return a;
    } else {
        return *p;
    }
}

template <bool S, typename T>
__device__ __forceinline__ void amp_st(cplx<T> *p, const cplx<T> &a)
{
    if constexpr (S) {
        typedef typename cvec2<T>::type V;
        V v;
        v.x = a.re;
        v.y = a.im;
        __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
    } else {
        *p = a;
    }
}

}  // namespace qslhip
