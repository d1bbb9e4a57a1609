// gate_math.h -- per-amplitude gate arithmetic shared by every kernel.
//
// Each update restates the body of the matching reference gate with the
// same operation order (src/qubits/default/gates.cpp); the library is
// compiled with -ffp-contract=off, so a multiply and an add are never fused
// and the device results are bit-identical to the reference's -O2 x86-64
// build (which has no FMA).  Both operate on registers only.
#pragma once

#include "internal.h"

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

// a = amplitude with the acted-on bit 0 (|01> for swap-like gates),
// b = amplitude with it 1.  `sqrt2` is the Fp-rounded double sqrt(2) of
// `constexpr Fp sqrt2 = std::sqrt(2)` (gates.cpp:94).
template <typename T>
__device__ __forceinline__ void pair_update(int body, cplx<T> &a, cplx<T> &b,
                                            T c, T s)
{
    const T sqrt2 = (T)1.4142135623730950488016887242097;
    T a0r = a.re, a0i = a.im, a1r = b.re, a1i = b.im;
    switch (body) {
    case QSLHIP_PAULI_X:  // gates.cpp:32-44
        a.re = a1r; a.im = a1i; b.re = a0r; b.im = a0i;
        break;
    case QSLHIP_PAULI_Y:  // gates.cpp:46-64
        a.re = a1i; a.im = -a1r; b.re = -a0i; b.im = a0r;
        break;
    case QSLHIP_HADAMARD:  // gates.cpp:83-101 (division, not 1/sqrt2 product)
        a.re = (a0r + a1r) / sqrt2; a.im = (a0i + a1i) / sqrt2;
        b.re = (a0r - a1r) / sqrt2; b.im = (a0i - a1i) / sqrt2;
        break;
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

// Insert a zero bit at position p.
__device__ __forceinline__ uint64_t ins0(uint64_t x, unsigned p)
{
    uint64_t lo = x & ((1ull << p) - 1);
    return ((x ^ lo) << 1) | lo;
}

}  // namespace qslhip
