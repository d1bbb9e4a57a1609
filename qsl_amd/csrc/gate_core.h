// gate_core.h -- per-amplitude gate arithmetic shared by every kernel and
// by the passes compiled at run time (device code only; hipRTC builds it).
//
// Each update restates the body of the matching reference gate with the
// same operation order (src/qubits/default/gates.cpp); the library is
// compiled with -ffp-contract=off, so a multiply and an add are never fused
// and the device results are bit-identical to the reference's -O2 x86-64
// build (which has no FMA); the one explicit fma pair (div_sqrt2) computes
// the correctly rounded quotient the reference's division gives.  All
// updates operate on registers only.
#pragma once

#include <stdint.h>

#include "cplx.h"
#include "qslhip.h"

namespace qslhip {

// Pair classes, decided at compile time from the kind.
enum PairClass { PC_PAIR1 = 0, PC_ONE = 1, PC_CTRL = 2, PC_SWAPLIKE = 3,
                 PC_DIAG2 = 4, PC_FSWAP = 5 };

__host__ __device__ constexpr int pair_class(int kind)
{
    return (kind == QSLHIP_PAULI_Z || kind == QSLHIP_PHASE) ? PC_ONE
         : (kind < 16) ? PC_PAIR1
         : (kind == QSLHIP_CZ || kind == QSLHIP_CPHASE) ? PC_DIAG2
         : (kind == QSLHIP_FSWAP) ? PC_FSWAP
         : (kind == QSLHIP_SWAP || kind == QSLHIP_NP_ROT_X ||
            kind == QSLHIP_NP_ROT_Y || kind == QSLHIP_NP_HADAMARD)
             ? PC_SWAPLIKE
             : PC_CTRL;
}

// Fold the controlled / number-preserving variants onto the arithmetic of
// their one-qubit body (controlNot's body is pauliX's swap, etc.).
__host__ __device__ constexpr int body_of(int kind)
{
    switch (kind) {
    case QSLHIP_CNOT: case QSLHIP_SWAP: case QSLHIP_FSWAP:
        return QSLHIP_PAULI_X;
    case QSLHIP_CY: return QSLHIP_PAULI_Y;
    case QSLHIP_CZ: return QSLHIP_PAULI_Z;
    case QSLHIP_CROT_X: case QSLHIP_NP_ROT_X: return QSLHIP_ROT_X;
    case QSLHIP_CROT_Y: case QSLHIP_NP_ROT_Y: return QSLHIP_ROT_Y;
    case QSLHIP_CROT_Z: return QSLHIP_ROT_Z;
    case QSLHIP_CPHASE: return QSLHIP_PHASE;
    case QSLHIP_CHADAMARD: case QSLHIP_NP_HADAMARD: return QSLHIP_HADAMARD;
    default: return kind;
    }
}

// x / d for the Fp-rounded d = sqrt(2) of `constexpr Fp sqrt2 =
// std::sqrt(2)` (gates.cpp:94), correctly rounded like the reference's
// division, in three operations instead of a full IEEE divide:
//     q0 = x * r,  e = fma(q0, d, -x) = q0*d - x  (exact),  q = fma(-e, r, q0),
// with r = RN(1/d).  Markstein's theorem (r within half an ulp of 1/d, q0
// within one ulp of x/d, no underflow in the residual) makes q = RN(x/d).
// The residual's sign convention keeps the sign of zero: x = -0 gives
// q0 = -0, e = +0, q = -0 + -0 = -0.  Tiny nonzero (|x| < 2^-100 FP32,
// 2^-900 FP64), huge, infinite and NaN x take the division itself.
// Checked on the CPU against x / d: every FP32 x, and 2^26 sampled FP64 x,
// bit for bit.
template <typename T>
__host__ __device__ __forceinline__ bool div_sqrt2_fast_ok(T ax)
{
    constexpr bool f64 = sizeof(T) == 8;
    const T lo = f64 ? (T)0x1p-900 : (T)0x1p-100f, hi = f64 ? (T)0x1p+900 : (T)0x1p+100f;
    return (ax <= hi) & ((ax >= lo) | (ax == (T)0));
}

template <typename T>
__host__ __device__ __forceinline__ T div_sqrt2_fast(T x)
{
    constexpr bool f64 = sizeof(T) == 8;
    const T d = f64 ? (T)0x1.6a09e667f3bcdp+0 : (T)0x1.6a09e6p+0f;
    const T r = f64 ? (T)0x1.6a09e667f3bccp-1 : (T)0x1.6a09e6p-1f;
    const T q0 = x * r;
    if constexpr (f64) return __builtin_fma(-__builtin_fma(q0, d, -x), r, q0);
    else return __builtin_fmaf(-__builtin_fmaf(q0, d, -x), r, q0);
}

// Keeps the rarely taken division branch a branch: an empty volatile asm
// on the operand cannot be speculated, so the compiler does not turn the
// branch into a select that evaluates the full IEEE divide on every lane.
template <typename T>
__host__ __device__ __forceinline__ T slow_path_operand(T x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x));
#endif
    return x;
}

template <typename T>
__host__ __device__ __forceinline__ T div_sqrt2(T x)
{
    const T d = sizeof(T) == 8 ? (T)0x1.6a09e667f3bcdp+0 : (T)0x1.6a09e6p+0f;
    T ax;
    if constexpr (sizeof(T) == 8) ax = __builtin_fabs(x);
    else ax = __builtin_fabsf(x);
    T q = div_sqrt2_fast<T>(x);
    if (!div_sqrt2_fast_ok<T>(ax)) q = slow_path_operand(x) / d;
    return q;
}

// The four quotients of one Hadamard pair with a single range test: the
// fast path for all four, and the division itself for all four when any
// of them is outside the safe range (one branch per pair, rarely taken).
template <typename T>
__host__ __device__ __forceinline__ void div_sqrt2_x4(T &x0, T &x1, T &x2, T &x3)
{
    const T d = sizeof(T) == 8 ? (T)0x1.6a09e667f3bcdp+0 : (T)0x1.6a09e6p+0f;
    T *xs[4] = {&x0, &x1, &x2, &x3};
    T q[4];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const T x = *xs[j];
        T ax;
        if constexpr (sizeof(T) == 8) ax = __builtin_fabs(x);
        else ax = __builtin_fabsf(x);
        ok = ok & div_sqrt2_fast_ok<T>(ax);
        q[j] = div_sqrt2_fast<T>(x);
    }
    if (!ok) {
#pragma unroll
        for (int j = 0; j < 4; j++) q[j] = slow_path_operand(*xs[j]) / d;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) *xs[j] = q[j];
}

// a = amplitude with the acted-on bit 0 (|01> for swap-like gates),
// b = amplitude with it 1.
template <typename T>
__device__ __forceinline__ void pair_update(int body, cplx<T> &a, cplx<T> &b,
                                            T c, T s)
{
    T a0r = a.re, a0i = a.im, a1r = b.re, a1i = b.im;
    switch (body) {
    case QSLHIP_PAULI_X:  // gates.cpp:32-44
        a.re = a1r; a.im = a1i; b.re = a0r; b.im = a0i;
        break;
    case QSLHIP_PAULI_Y:  // gates.cpp:46-64
        a.re = a1i; a.im = -a1r; b.re = -a0i; b.im = a0r;
        break;
    case QSLHIP_HADAMARD:  // gates.cpp:83-101 (division, not 1/sqrt2 product)
    {
        T x0 = a0r + a1r, x1 = a0i + a1i, x2 = a0r - a1r, x3 = a0i - a1i;
        div_sqrt2_x4<T>(x0, x1, x2, x3);
        a.re = x0; a.im = x1; b.re = x2; b.im = x3;
        break;
    }
    case QSLHIP_ROT_X:  // gates.cpp:120-148
        a.re = a0r * c + a1i * s; a.im = a0i * c - a1r * s;
        b.re = a1r * c + a0i * s; b.im = a1i * c - a0r * s;
        break;
    case QSLHIP_ROT_Y:  // gates.cpp:150-178
        a.re = a0r * c - a1r * s; a.im = a0i * c - a1i * s;
        b.re = a0r * s + a1r * c; b.im = a0i * s + a1i * c;
        break;
    case QSLHIP_ROT_Z:  // gates.cpp:181-209
        a.re = a0r * c + a0i * s; a.im = a0i * c - a0r * s;
        b.re = a1r * c - a1i * s; b.im = a1i * c + a1r * s;
        break;
    default:
        break;
    }
}

// The same updates for QSLHIP_MODE_REORDER, where results are bound by a
// tolerance instead of bit-identity: products folded into fused
// multiply-adds (one rounding instead of two) and the Hadamard as a product
// with RN(1/sqrt2) (within an ulp of the division).
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c)
{
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}

template <typename T>
__device__ __forceinline__ void pair_update_fast(int body, cplx<T> &a, cplx<T> &b, T c, T s)
{
    const T rsqrt2 = sizeof(T) == 8 ? (T)0x1.6a09e667f3bccp-1 : (T)0x1.6a09e6p-1f;
    T a0r = a.re, a0i = a.im, a1r = b.re, a1i = b.im;
    switch (body) {
    case QSLHIP_HADAMARD:
        a.re = (a0r + a1r) * rsqrt2; a.im = (a0i + a1i) * rsqrt2;
        b.re = (a0r - a1r) * rsqrt2; b.im = (a0i - a1i) * rsqrt2;
        break;
    case QSLHIP_ROT_X:
        a.re = fmadd(a0r, c, a1i * s); a.im = fmadd(a0i, c, -(a1r * s));
        b.re = fmadd(a1r, c, a0i * s); b.im = fmadd(a1i, c, -(a0r * s));
        break;
    case QSLHIP_ROT_Y:
        a.re = fmadd(a0r, c, -(a1r * s)); a.im = fmadd(a0i, c, -(a1i * s));
        b.re = fmadd(a0r, s, a1r * c); b.im = fmadd(a0i, s, a1i * c);
        break;
    default:  // pauliX / pauliY: moves and negations, already exact
        pair_update<T>(body, a, b, c, s);
        break;
    }
}

template <typename T>
__device__ __forceinline__ void one_update_fast(int body, cplx<T> &b, T c, T s)
{
    if (body == QSLHIP_PAULI_Z) {
        b.re = -b.re; b.im = -b.im;
    } else {
        const T tr = b.re, ti = b.im;
        b.re = fmadd(c, tr, -(s * ti));
        b.im = fmadd(c, ti, s * tr);
    }
}

// Update of a single amplitude whose acted-on bit(s) are 1: pauliZ /
// controlZ / fswap's |11> negate (gates.cpp:66-81, 278-301, 536-570) or
// phase / controlPhase (gates.cpp:103-118, 436-465).
template <typename T>
__device__ __forceinline__ void one_update(int body, cplx<T> &b, T c, T s)
{
    if (body == QSLHIP_PAULI_Z) {
        b.re *= (T)-1; b.im *= (T)-1;
    } else {
        T tr = b.re, ti = b.im;
        b.re = c * tr - s * ti;
        b.im = c * ti + s * tr;
    }
}

// rotateZ split into its two diagonal halves (used when an amplitude is
// updated alone, e.g. with the target outside an LDS tile).
template <typename T>
__device__ __forceinline__ void rz_half(bool one, cplx<T> &x, T c, T s)
{
    T r = x.re, i = x.im;
    if (!one) { x.re = r * c + i * s; x.im = i * c - r * s; }
    else      { x.re = r * c - i * s; x.im = i * c + r * s; }
}

__host__ __device__ constexpr bool is_gate2_dev(int kind) { return kind >= 16 && kind <= 28; }

// Deposit the low bits of v into the set bits of mask (ascending).
__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask)
{
    uint64_t r = 0;
    while (mask) {
        uint64_t low = mask & (0 - mask);
        if (v & 1) r |= low;
        v >>= 1;
        mask ^= low;
    }
    return r;
}

// Insert a zero bit at position p.
__device__ __forceinline__ uint64_t ins0(uint64_t x, unsigned p)
{
    uint64_t lo = x & ((1ull << p) - 1);
    return ((x ^ lo) << 1) | lo;
}

// |a|^2 formed in Fp, as the reference's `Fp prob = re*re + im*im`
// (src/qubits/default/measure.cpp:66-67), then widened to FP64 for sums.
template <typename T>
__device__ __forceinline__ double amp_prob(const cplx<T> &a)
{
    T p = a.re * a.re + a.im * a.im;
    return (double)p;
}

// Sum over the 64 lanes of a wavefront, xor butterfly: every lane gets the
// same total, and the association depends only on the lane order.
__device__ __forceinline__ double wave_sum_xor(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Xor butterfly over lane bits below `width` (1 << steps lanes), lowest bit
// first: the sum of 8 lanes is ((l0+l1)+(l2+l3))+((l4+l5)+(l6+l7)), and
// steps 6 extends that tree over the eight 8-lane sums.
template <int STEPS>
__device__ __forceinline__ double lane_sum_up(double v)
{
#pragma unroll
    for (int i = 0; i < STEPS; i++) v += __shfl_xor(v, 1 << i, 64);
    return v;
}

}  // namespace qslhip
