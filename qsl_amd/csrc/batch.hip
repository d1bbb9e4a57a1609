// batch.hip -- energies and adjoint gradients of thousands of small
// variational circuits in one persistent kernel: FP32 states, FP64 energies
// and gradient sums, captured once in a hipGraph and replayed for each new
// set of angles.  (The reference has no gradients; the per-circuit FP64
// path in adjoint.hip is the check.)
//
// One workgroup of 512 threads per circuit, largest circuits first, from an
// atomic queue.  A state of <= 12 qubits lives in LDS; a larger one sits in
// the workgroup's HBM slot and goes through tile passes of 2^12 amplitudes
// (the tile qubits come from the gate path's commutation-aware planner).
// Within a pass the ops are cut into segments whose non-diagonal targets
// lie on at most three tile positions s0 < s1 < s2: thread t holds the eight
// amplitudes base(t) | {subsets of 1<<s0, 1<<s1, 1<<s2} in registers and
// applies every op of the segment to them, with one barrier per segment.
// The host decodes each op into a micro-op for its segment, naming for every
// bit the op tests where that bit comes from: one of the three block
// positions (fixed per register), the thread's base index (one bit for all
// eight registers) or the group's base index (one bit for the workgroup).
// The kernel reads the micro-ops and the cos / sin of the angles with
// scalar loads.
//
// Backward sweep (the adjoint method of adjoint.hip): lambda = H psi, then
// for each op, last to first, the gradient term Im <lambda|P|psi> and the
// inverse op on psi and lambda.  Terms are formed in FP32 (the states'
// precision), summed over a wavefront in FP32 and accumulated over the
// wavefronts and groups in FP64.  The updates are the gate bodies of
// gate_core.h with products folded into fused multiply-adds (results are
// bound by a tolerance against FP64, not by bit-identity).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>

#include "adjoint.h"
#include "tile_ops.h"

namespace qslhip {

namespace {

constexpr int kBBlock = 512;           // threads per circuit workgroup
constexpr int kBWaves = kBBlock / 64;
constexpr int kElems = 8;              // amplitudes per thread: 3 block positions
constexpr int kLdsQubits = 12;         // whole state in LDS up to 12 qubits
constexpr int kTileQubits = 12;        // tile passes of 2^12 amplitudes
constexpr int kMaxPassOps = 64;        // accumulator slots per pass
constexpr int kLowRun = 4;             // lowest tile qubits kept contiguous
static_assert(kBBlock * kElems == 1 << kTileQubits, "one 8-amplitude block per thread");

// Micro-op codes.
enum UCode : uint8_t {
    U_NONE = 0,
    U_PAIR,  // body on the pairs along a block position, optional control
    U_RZ,    // rotateZ halves by a target bit, optional control
    U_ONE,   // pauliZ / phase (and the controlled forms) on the elements
             // whose one or two bits are 1
    U_SWAP   // swap-like body on the {10, 01} pairs of two block positions
};

// Bit sources: 0..2 a block position, kSrcThread + p the thread base's bit
// at tile position p, kSrcGroup + q the group base's bit of qubit q,
// kSrcNone a bit that is always 1.
constexpr uint8_t kSrcThread = 0x10;
constexpr uint8_t kSrcGroup = 0x40;
constexpr uint8_t kSrcNone = 0xFF;

// A micro-op as four words (one 16-byte scalar load):
//   w0 = code | body << 8 | gen << 16 | fswap << 24
//   w1 = ta | tb << 8 | ctl << 16
// ta: U_PAIR the block position of the target, U_SWAP of q0, U_RZ / U_ONE
// the source of the (first) target bit; tb: U_SWAP the block position of
// q1, U_ONE the second bit source; ctl: the control bit source (U_PAIR,
// U_RZ); gop: the global op index (its angle and gradient slot).
struct alignas(16) UOp {
    uint32_t w0, w1, gop, pad;
};

struct UOpD {
    int code, body, gen, fswap, ta, tb, ctl;
    uint32_t gop;
};

__device__ __forceinline__ UOpD uop_at(const UOp *__restrict__ uops, uint32_t i)
{
    const UOp x = uops[i];
    UOpD u;
    u.code = (int)(x.w0 & 255u);
    u.body = (int)((x.w0 >> 8) & 255u);
    u.gen = (int)((x.w0 >> 16) & 255u);
    u.fswap = (int)(x.w0 >> 24);
    u.ta = (int)(x.w1 & 255u);
    u.tb = (int)((x.w1 >> 8) & 255u);
    u.ctl = (int)((x.w1 >> 16) & 255u);
    u.gop = x.gop;
    return u;
}

struct CS {
    float c, s;
};

struct BCirc {
    uint32_t n;
    uint32_t pass_first, npass;
    uint32_t op_first, nop;  // global op range (angles / grads index)
    uint32_t term_first, nterm;
    uint32_t lds;            // 1: whole state in LDS
};

struct BPass {
    uint64_t tile_mask;
    uint32_t top_first, top_count;  // into the micro-op table
    uint32_t crun, pad;
    uint32_t seg_first, seg_count;  // into the segment table
};

// Segment word: s0 | s1 << 4 | s2 << 8 | first op << 12 | op count << 20.
struct Seg {
    int s[3];
    uint32_t first, count;
};

__host__ __device__ inline Seg seg_of(uint32_t w)
{
    Seg g;
    g.s[0] = (int)(w & 15u);
    g.s[1] = (int)((w >> 4) & 15u);
    g.s[2] = (int)((w >> 8) & 15u);
    g.first = (w >> 12) & 255u;
    g.count = (w >> 20) & 255u;
    return g;
}

// ---- per-element arithmetic --------------------------------------------------

__device__ __forceinline__ float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// a: amplitude with the acted-on bit 0 (|10> of a swap-like pair: q0 set),
// b: the other member.
template <int BODY>
__device__ __forceinline__ void pair_op(cplx<float> &a, cplx<float> &b, float c, float s)
{
    const float a0r = a.re, a0i = a.im, a1r = b.re, a1i = b.im;
    if constexpr (BODY == QSLHIP_PAULI_X) {
        a.re = a1r; a.im = a1i; b.re = a0r; b.im = a0i;
    } else if constexpr (BODY == QSLHIP_PAULI_Y) {
        a.re = a1i; a.im = -a1r; b.re = -a0i; b.im = a0r;
    } else if constexpr (BODY == QSLHIP_HADAMARD) {
        const float r = 0x1.6a09e6p-1f;  // RN(1/sqrt2)
        a.re = (a0r + a1r) * r; a.im = (a0i + a1i) * r;
        b.re = (a0r - a1r) * r; b.im = (a0i - a1i) * r;
    } else if constexpr (BODY == QSLHIP_ROT_X) {
        a.re = fmaf_(a0r, c, a1i * s); a.im = fmaf_(a0i, c, -(a1r * s));
        b.re = fmaf_(a1r, c, a0i * s); b.im = fmaf_(a1i, c, -(a0r * s));
    } else {  // QSLHIP_ROT_Y
        a.re = fmaf_(a0r, c, -(a1r * s)); a.im = fmaf_(a0i, c, -(a1i * s));
        b.re = fmaf_(a0r, s, a1r * c); b.im = fmaf_(a0i, s, a1i * c);
    }
}

// rotateZ half: phase exp(-i theta/2) for target bit 0, exp(+i theta/2) for
// 1, written as one form with s' = s (bit 0) or -s (bit 1).
__device__ __forceinline__ void rz_op(cplx<float> &x, float c, float s)
{
    const float r = x.re, i = x.im;
    x.re = fmaf_(r, c, i * s);
    x.im = fmaf_(i, c, -(r * s));
}

template <int BODY>
__device__ __forceinline__ void one_op(cplx<float> &x, float c, float s)
{
    if constexpr (BODY == QSLHIP_PAULI_Z) {
        x.re = -x.re; x.im = -x.im;
    } else {  // QSLHIP_PHASE
        const float tr = x.re, ti = x.im;
        x.re = fmaf_(c, tr, -(s * ti));
        x.im = fmaf_(c, ti, s * tr);
    }
}

// Re / Im of conj(l) * p.
__device__ __forceinline__ float re_cm(const cplx<float> &l, const cplx<float> &p)
{
    return fmaf_(l.re, p.re, l.im * p.im);
}
__device__ __forceinline__ float im_cm(const cplx<float> &l, const cplx<float> &p)
{
    return fmaf_(l.re, p.im, -(l.im * p.re));
}

struct Blk {
    cplx<float> v[kElems];
};

// The bit of an element-uniform source (thread base, group base or none).
__device__ __forceinline__ bool src_bit(uint32_t src, uint32_t base, uint64_t gbase)
{
    if (src == kSrcNone) return true;
    if (src < kSrcGroup) return (base >> (src - kSrcThread)) & 1u;
    return (gbase >> (src - kSrcGroup)) & 1ull;
}

// ---- micro-op application ------------------------------------------------------
// Element e of a block has bit k of e as its bit at block position k, so
// every test on a block position is a compile-time pattern once the
// position is known; the dispatch below turns the run-time position into
// template arguments.

// Pairs (e, e | 1 << K) with bit K of e clear; with KC >= 0 only those whose
// elements have bit KC set (a control on another block position).
template <int BODY, int NB, int K, int KC>
__device__ __forceinline__ void pair_loop(Blk *b, float c, float s)
{
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        if ((e >> K) & 1) continue;
        if (KC >= 0 && !((e >> KC) & 1)) continue;
#pragma unroll
        for (int w = 0; w < NB; w++) pair_op<BODY>(b[w].v[e], b[w].v[e | (1 << K)], c, s);
    }
}

template <int BODY, int NB, int K>
__device__ __forceinline__ void pair_kc(Blk *b, int kc, float c, float s)
{
    constexpr int O1 = K == 0 ? 1 : 0, O2 = K == 2 ? 1 : 2;
    if (kc == O1) pair_loop<BODY, NB, K, O1>(b, c, s);
    else if (kc == O2) pair_loop<BODY, NB, K, O2>(b, c, s);
    else pair_loop<BODY, NB, K, -1>(b, c, s);
}

template <int BODY, int NB>
__device__ __forceinline__ void pair_body(Blk *b, int k, int kc, float c, float s)
{
    if (k == 0) pair_kc<BODY, NB, 0>(b, kc, c, s);
    else if (k == 1) pair_kc<BODY, NB, 1>(b, kc, c, s);
    else pair_kc<BODY, NB, 2>(b, kc, c, s);
}

// Elements e with (KT < 0 or) bit KT of e set take s_hi, the others s_lo;
// with KC >= 0 only elements with bit KC set are updated.
template <int KT, int KC, int NB>
__device__ __forceinline__ void rz_elems(Blk *b, float c, float s_lo, float s_hi)
{
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        if (KC >= 0 && !((e >> KC) & 1)) continue;
        const float se = (KT >= 0 && ((e >> KT) & 1)) ? s_hi : s_lo;
#pragma unroll
        for (int w = 0; w < NB; w++) rz_op(b[w].v[e], c, se);
    }
}

template <int KT, int NB>
__device__ __forceinline__ void rz_kc(Blk *b, int kc, float c, float s_lo, float s_hi)
{
    if (kc == 0 && KT != 0) rz_elems<KT, (KT == 0 ? -1 : 0), NB>(b, c, s_lo, s_hi);
    else if (kc == 1 && KT != 1) rz_elems<KT, (KT == 1 ? -1 : 1), NB>(b, c, s_lo, s_hi);
    else if (kc == 2 && KT != 2) rz_elems<KT, (KT == 2 ? -1 : 2), NB>(b, c, s_lo, s_hi);
    else rz_elems<KT, -1, NB>(b, c, s_lo, s_hi);
}

template <int NB>
__device__ __forceinline__ void rz_apply(Blk *b, int ta, int kc, float c, float s, uint32_t base,
                                         uint64_t gbase)
{
    if (ta == 0) rz_kc<0, NB>(b, kc, c, s, -s);
    else if (ta == 1) rz_kc<1, NB>(b, kc, c, s, -s);
    else if (ta == 2) rz_kc<2, NB>(b, kc, c, s, -s);
    else {
        const float s2 = src_bit(ta, base, gbase) ? -s : s;
        rz_kc<-1, NB>(b, kc, c, s2, s2);
    }
}

template <int KA, int KB, int BODY, int NB>
__device__ __forceinline__ void one_elems(Blk *b, float c, float s)
{
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        if (KA >= 0 && !((e >> KA) & 1)) continue;
        if (KB >= 0 && !((e >> KB) & 1)) continue;
#pragma unroll
        for (int w = 0; w < NB; w++) one_op<BODY>(b[w].v[e], c, s);
    }
}

template <int BODY, int NB>
__device__ __forceinline__ void one_body(Blk *b, int ka, int kb, float c, float s)
{
    // ka <= kb after normalisation; -1: no block-position bit
    if (ka < 0) one_elems<-1, -1, BODY, NB>(b, c, s);
    else if (kb < 0) {
        if (ka == 0) one_elems<0, -1, BODY, NB>(b, c, s);
        else if (ka == 1) one_elems<1, -1, BODY, NB>(b, c, s);
        else one_elems<2, -1, BODY, NB>(b, c, s);
    } else if (ka == 0) {
        if (kb == 1) one_elems<0, 1, BODY, NB>(b, c, s);
        else one_elems<0, 2, BODY, NB>(b, c, s);
    } else {
        one_elems<1, 2, BODY, NB>(b, c, s);
    }
}

template <int KA, int KB, int BODY, bool FSW, int NB>
__device__ __forceinline__ void swap_elems(Blk *b, float c, float s)
{
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        const bool a = (e >> KA) & 1, bb = (e >> KB) & 1;
        if (a && !bb) {
            const int p = e ^ (1 << KA) ^ (1 << KB);
#pragma unroll
            for (int w = 0; w < NB; w++) pair_op<BODY>(b[w].v[e], b[w].v[p], c, s);
        }
        if (FSW && a && bb) {
#pragma unroll
            for (int w = 0; w < NB; w++) one_op<QSLHIP_PAULI_Z>(b[w].v[e], c, s);
        }
    }
}

template <int BODY, bool FSW, int NB>
__device__ __forceinline__ void swap_body(Blk *b, int ka, int kb, float c, float s)
{
    switch (ka * 3 + kb) {
    case 1: swap_elems<0, 1, BODY, FSW, NB>(b, c, s); break;
    case 2: swap_elems<0, 2, BODY, FSW, NB>(b, c, s); break;
    case 3: swap_elems<1, 0, BODY, FSW, NB>(b, c, s); break;
    case 5: swap_elems<1, 2, BODY, FSW, NB>(b, c, s); break;
    case 6: swap_elems<2, 0, BODY, FSW, NB>(b, c, s); break;
    default: swap_elems<2, 1, BODY, FSW, NB>(b, c, s); break;
    }
}

// Apply micro-op u to NB blocks (the forward sweep: psi; the backward
// sweep: psi and lambda, with s negated for the inverse).
template <int NB>
__device__ __forceinline__ void apply_uop(Blk *b, const UOpD &u, float c, float s, uint32_t base,
                                          uint64_t gbase)
{
    const int code = u.code;
    if (code == U_PAIR || code == U_RZ) {
        int kc = -1;
        if (u.ctl < 3) kc = u.ctl;
        else if (!src_bit(u.ctl, base, gbase)) return;  // control bit 0 for the whole block
        if (code == U_RZ) {
            rz_apply<NB>(b, u.ta, kc, c, s, base, gbase);
            return;
        }
        switch (u.body) {
        case QSLHIP_PAULI_X: pair_body<QSLHIP_PAULI_X, NB>(b, u.ta, kc, c, s); break;
        case QSLHIP_PAULI_Y: pair_body<QSLHIP_PAULI_Y, NB>(b, u.ta, kc, c, s); break;
        case QSLHIP_HADAMARD: pair_body<QSLHIP_HADAMARD, NB>(b, u.ta, kc, c, s); break;
        case QSLHIP_ROT_X: pair_body<QSLHIP_ROT_X, NB>(b, u.ta, kc, c, s); break;
        default: pair_body<QSLHIP_ROT_Y, NB>(b, u.ta, kc, c, s); break;
        }
    } else if (code == U_ONE) {
        int ka = u.ta < 3 ? u.ta : -1, kb = u.tb < 3 ? u.tb : -1;
        if (u.ta >= 3 && !src_bit(u.ta, base, gbase)) return;
        if (u.tb >= 3 && !src_bit(u.tb, base, gbase)) return;
        if (ka < 0 || (kb >= 0 && kb < ka)) {
            const int t = ka;
            ka = kb;
            kb = t;
        }
        // pauliZ / controlZ carry c = -1, s = 0: the phase update then
        // negates exactly
        one_body<QSLHIP_PHASE, NB>(b, ka, kb, c, s);
    } else if (code == U_SWAP) {
        if (u.fswap) swap_body<QSLHIP_PAULI_X, true, NB>(b, u.ta, u.tb, c, s);
        else if (u.body == QSLHIP_PAULI_X) swap_body<QSLHIP_PAULI_X, false, NB>(b, u.ta, u.tb, c, s);
        else if (u.body == QSLHIP_ROT_X) swap_body<QSLHIP_ROT_X, false, NB>(b, u.ta, u.tb, c, s);
        else if (u.body == QSLHIP_ROT_Y) swap_body<QSLHIP_ROT_Y, false, NB>(b, u.ta, u.tb, c, s);
        else swap_body<QSLHIP_HADAMARD, false, NB>(b, u.ta, u.tb, c, s);
    }
}

// ---- gradient terms --------------------------------------------------------------
// The lane's sum of Im <lambda|P|psi> over its block for micro-op u's
// generator (0 where the control is off).  No lane leaves early: the
// caller reduces over the whole wavefront.

template <int K, int KC, int GEN>
__device__ __forceinline__ float pair_terms(const Blk &p, const Blk &l)
{
    float v = 0.f;
#pragma unroll
    for (int e0 = 0; e0 < kElems; e0++) {
        if ((e0 >> K) & 1) continue;
        if (KC >= 0 && !((e0 >> KC) & 1)) continue;
        const int e1 = e0 | (1 << K);
        if constexpr (GEN == GEN_X)
            v += im_cm(l.v[e0], p.v[e1]) + im_cm(l.v[e1], p.v[e0]);
        else  // GEN_Y: (Y psi)_i = -i psi_{i^b} for bit 0, +i psi_{i^b} for bit 1
            v += re_cm(l.v[e1], p.v[e0]) - re_cm(l.v[e0], p.v[e1]);
    }
    return v;
}

template <int K, int KC>
__device__ __forceinline__ float pair_terms_g(const Blk &p, const Blk &l, int gen)
{
    return gen == GEN_X ? pair_terms<K, KC, GEN_X>(p, l) : pair_terms<K, KC, GEN_Y>(p, l);
}

template <int KT, int KC>
__device__ __forceinline__ float z_terms(const Blk &p, const Blk &l)
{
    float v = 0.f;
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        if (KC >= 0 && !((e >> KC) & 1)) continue;
        const float t = im_cm(l.v[e], p.v[e]);
        v += (KT >= 0 && ((e >> KT) & 1)) ? -t : t;
    }
    return v;
}

template <int KT>
__device__ __forceinline__ float z_terms_kc(const Blk &p, const Blk &l, int kc)
{
    if (kc == 0 && KT != 0) return z_terms<KT, (KT == 0 ? -1 : 0)>(p, l);
    if (kc == 1 && KT != 1) return z_terms<KT, (KT == 1 ? -1 : 1)>(p, l);
    if (kc == 2 && KT != 2) return z_terms<KT, (KT == 2 ? -1 : 2)>(p, l);
    return z_terms<KT, -1>(p, l);
}

template <int KA, int KB>
__device__ __forceinline__ float p1_terms(const Blk &p, const Blk &l)
{
    float v = 0.f;
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        if (KA >= 0 && !((e >> KA) & 1)) continue;
        if (KB >= 0 && !((e >> KB) & 1)) continue;
        v += im_cm(l.v[e], p.v[e]);
    }
    return -2.f * v;
}

template <int KA, int KB>
__device__ __forceinline__ float np_terms(const Blk &p, const Blk &l, int gen)
{
    float v = 0.f;
#pragma unroll
    for (int e = 0; e < kElems; e++) {
        const bool a = (e >> KA) & 1, bb = (e >> KB) & 1;
        if (a == bb) continue;
        const int q = e ^ (1 << KA) ^ (1 << KB);
        if (gen == GEN_NPX) v += im_cm(l.v[e], p.v[q]);
        else v += a ? -re_cm(l.v[e], p.v[q]) : re_cm(l.v[e], p.v[q]);
    }
    return v;
}

__device__ __forceinline__ float grad_terms(const Blk &p, const Blk &l, const UOpD &u,
                                            uint32_t base, uint64_t gbase)
{
    const int gen = u.gen;
    float v = 0.f;
    if (u.code == U_PAIR || u.code == U_RZ) {
        int kc = -1;
        bool on = true;
        if (u.ctl < 3) kc = u.ctl;
        else on = src_bit(u.ctl, base, gbase);
        if (u.code == U_PAIR) {
            const int k = u.ta;
            if (k == 0) {
                if (kc == 1) v = pair_terms_g<0, 1>(p, l, gen);
                else if (kc == 2) v = pair_terms_g<0, 2>(p, l, gen);
                else v = pair_terms_g<0, -1>(p, l, gen);
            } else if (k == 1) {
                if (kc == 0) v = pair_terms_g<1, 0>(p, l, gen);
                else if (kc == 2) v = pair_terms_g<1, 2>(p, l, gen);
                else v = pair_terms_g<1, -1>(p, l, gen);
            } else {
                if (kc == 0) v = pair_terms_g<2, 0>(p, l, gen);
                else if (kc == 1) v = pair_terms_g<2, 1>(p, l, gen);
                else v = pair_terms_g<2, -1>(p, l, gen);
            }
        } else {  // GEN_Z
            if (u.ta == 0) v = z_terms_kc<0>(p, l, kc);
            else if (u.ta == 1) v = z_terms_kc<1>(p, l, kc);
            else if (u.ta == 2) v = z_terms_kc<2>(p, l, kc);
            else {
                v = z_terms_kc<-1>(p, l, kc);
                if (src_bit(u.ta, base, gbase)) v = -v;
            }
        }
        return on ? v : 0.f;
    }
    if (u.code == U_ONE) {  // GEN_P1
        int ka = u.ta < 3 ? u.ta : -1, kb = u.tb < 3 ? u.tb : -1;
        bool on = true;
        if (u.ta >= 3) on = src_bit(u.ta, base, gbase);
        if (u.tb >= 3) on = on && src_bit(u.tb, base, gbase);
        if (ka < 0 || (kb >= 0 && kb < ka)) {
            const int t = ka;
            ka = kb;
            kb = t;
        }
        if (ka < 0) v = p1_terms<-1, -1>(p, l);
        else if (kb < 0) v = ka == 0 ? p1_terms<0, -1>(p, l)
                           : ka == 1 ? p1_terms<1, -1>(p, l) : p1_terms<2, -1>(p, l);
        else if (ka == 0) v = kb == 1 ? p1_terms<0, 1>(p, l) : p1_terms<0, 2>(p, l);
        else v = p1_terms<1, 2>(p, l);
        return on ? v : 0.f;
    }
    // U_SWAP: GEN_NPX / GEN_NPY
    switch (u.ta * 3 + u.tb) {
    case 1: return np_terms<0, 1>(p, l, gen);
    case 2: return np_terms<0, 2>(p, l, gen);
    case 3: return np_terms<1, 0>(p, l, gen);
    case 5: return np_terms<1, 2>(p, l, gen);
    case 6: return np_terms<2, 0>(p, l, gen);
    default: return np_terms<2, 1>(p, l, gen);
    }
}

// Sum of v over the wavefront (every lane active), by row-local DPP steps
// and the two row broadcasts; the total is in lane 63 and returned to all.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v)
{
    const int x = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false);
    return v + __int_as_float(x);
}

__device__ __forceinline__ float wave_total(float v)
{
    v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]: quad sums
    v = dpp_add<0x141, 0xF>(v);  // row_half_mirror: half-row sums
    v = dpp_add<0x140, 0xF>(v);  // row_mirror: row sums in every lane
    v = dpp_add<0x142, 0xA>(v);  // row_bcast15 into rows 1, 3
    v = dpp_add<0x143, 0xC>(v);  // row_bcast31 into rows 2, 3: lane 63 = total
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_reduce(double v, double *sh)
{
    v = wave_sum_d(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBWaves; w++) r += sh[w];
    __syncthreads();
    return r;
}

// ---- segments and groups -----------------------------------------------------------

__device__ __forceinline__ uint32_t insert0(uint32_t x, int s)
{
    const uint32_t lo = x & ((1u << s) - 1);
    return ((x ^ lo) << 1) | lo;
}

// Element offsets of a segment's block (uniform) and the thread's base.
struct Geo {
    uint32_t off[kElems];
    uint32_t base;
};

__device__ __forceinline__ Geo geo_of(const Seg &g)
{
    Geo o;
    o.base = insert0(insert0(insert0(threadIdx.x, g.s[0]), g.s[1]), g.s[2]);
#pragma unroll
    for (int e = 0; e < kElems; e++)
        o.off[e] = ((e & 1) ? 1u << g.s[0] : 0u) | ((e & 2) ? 1u << g.s[1] : 0u) |
                   ((e & 4) ? 1u << g.s[2] : 0u);
    return o;
}

// Where a group's tile element j lives in the HBM slot.
struct GroupMap {
    uint64_t gbase;
    uint32_t lowm, crun;
    const uint64_t *hoff;
    __device__ __forceinline__ uint64_t at(uint32_t j) const
    {
        return gbase | (j & lowm) | hoff[j >> crun];
    }
};

struct Kargs {
    const BCirc *__restrict__ circ;
    const BPass *__restrict__ pass;
    const UOp *__restrict__ uops;
    const uint32_t *__restrict__ segs;
    const qslhip_pauli *__restrict__ terms;
    const CS *__restrict__ cs;
    const uint32_t *__restrict__ order;
    uint32_t ncirc;
    uint32_t *queue;
    double *energy;
    double *grads;
    cplx<float> *slots;  // 2 * slot_amps per workgroup
    uint64_t slot_amps;
};

// Forward sweep of one group of pass P through its segments.  The first
// segment reads the group from the slot (or |0...0> when init), or from the
// LDS tile when !slot; the last one writes it back the same way.  Elements
// at or past nt (states of fewer than 3 qubits) read as 0 and are not
// written.  Ends with a barrier.
__device__ __forceinline__ void group_fwd(const UOp *__restrict__ uops,
                                          const CS *__restrict__ cs,
                                          const uint32_t *__restrict__ segs, const BPass &P,
                                          const GroupMap &gm, bool slot, bool init,
                                          cplx<float> *gpsi, cplx<float> *tile, uint32_t nt)
{
    for (uint32_t sg = 0; sg < P.seg_count; sg++) {
        const Seg g = seg_of(segs[P.seg_first + sg]);
        const Geo o = geo_of(g);
        Blk b;
        if (sg == 0 && slot) {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint64_t idx = gm.at(o.base | o.off[e]);
                if (init) {
                    b.v[e].re = idx == 0 ? 1.f : 0.f;
                    b.v[e].im = 0.f;
                } else {
                    b.v[e] = gpsi[idx];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint32_t j = o.base | o.off[e];
                b.v[e] = tile[j < nt ? j : 0];
                if (j >= nt) b.v[e].re = b.v[e].im = 0.f;
            }
        }
        for (uint32_t k = g.first; k < g.first + g.count; k++) {
            const UOpD u = uop_at(uops, P.top_first + k);
            const CS a = cs[u.gop];
            apply_uop<1>(&b, u, a.c, a.s, o.base, gm.gbase);
        }
        if (sg + 1 == P.seg_count && slot) {
#pragma unroll
            for (int e = 0; e < kElems; e++) gpsi[gm.at(o.base | o.off[e])] = b.v[e];
        } else {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint32_t j = o.base | o.off[e];
                if (j < nt) tile[j] = b.v[e];
            }
        }
        __syncthreads();
    }
}

// Backward sweep of one group (last segment first): gradient terms, then
// the inverse ops on psi and lambda.  With slot, the first segment
// processed reads both from the slot and the last one writes them back
// when `store` (the final backward pass keeps nothing).  Ends with a
// barrier.
__device__ __forceinline__ void group_bwd(const UOp *__restrict__ uops,
                                          const CS *__restrict__ cs,
                                          const uint32_t *__restrict__ segs, const BPass &P,
                                          const GroupMap &gm, bool slot, bool store,
                                          cplx<float> *gpsi, cplx<float> *glam,
                                          cplx<float> *tpsi, cplx<float> *tlam, uint32_t nt,
                                          double *acc)
{
    for (uint32_t sg = P.seg_count; sg-- > 0;) {
        const Seg g = seg_of(segs[P.seg_first + sg]);
        const Geo o = geo_of(g);
        Blk b[2];
        if (sg + 1 == P.seg_count && slot) {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint64_t idx = gm.at(o.base | o.off[e]);
                b[0].v[e] = gpsi[idx];
                b[1].v[e] = glam[idx];
            }
        } else {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint32_t j = o.base | o.off[e];
                const uint32_t jj = j < nt ? j : 0;
                b[0].v[e] = tpsi[jj];
                b[1].v[e] = tlam[jj];
                if (j >= nt) b[0].v[e].re = b[0].v[e].im = b[1].v[e].re = b[1].v[e].im = 0.f;
            }
        }
        for (uint32_t k = g.first + g.count; k-- > g.first;) {
            const UOpD u = uop_at(uops, P.top_first + k);
            const CS a = cs[u.gop];
            if (u.gen != GEN_NONE) {
                const float t = wave_total(grad_terms(b[0], b[1], u, o.base, gm.gbase));
                if ((threadIdx.x & 63) == 0) acc[k * kBWaves + (threadIdx.x >> 6)] += (double)t;
            }
            apply_uop<2>(b, u, a.c, -a.s, o.base, gm.gbase);
        }
        if (sg == 0 && slot) {
            if (store) {
#pragma unroll
                for (int e = 0; e < kElems; e++) {
                    const uint64_t idx = gm.at(o.base | o.off[e]);
                    gpsi[idx] = b[0].v[e];
                    glam[idx] = b[1].v[e];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < kElems; e++) {
                const uint32_t j = o.base | o.off[e];
                if (j < nt) {
                    tpsi[j] = b[0].v[e];
                    tlam[j] = b[1].v[e];
                }
            }
        }
        __syncthreads();
    }
}

// lambda = H psi over the whole state (the slot or the LDS tile); returns
// the workgroup's energy sum (valid in thread 0).
__device__ __forceinline__ double apply_h(const cplx<float> *psi, cplx<float> *lam, uint64_t namps,
                                          const qslhip_pauli *terms, int nterm, double *sh)
{
    double acc = 0;
    for (uint64_t i = threadIdx.x; i < namps; i += kBBlock) {
        double lr = 0, li = 0;
        for (int t = 0; t < nterm; t++) {
            const qslhip_pauli P = terms[t];
            uint64_t j = i ^ P.xmask;
            cplx<float> v = psi[j];
            int ny = __popcll(P.xmask & P.zmask) & 3;
            double sg = (__popcll(j & P.zmask) & 1) ? -P.coef : P.coef;
            double vr = sg * (double)v.re, vi = sg * (double)v.im;
            double r, m;
            switch (ny) {
            case 0: r = vr; m = vi; break;
            case 1: r = -vi; m = vr; break;
            case 2: r = -vr; m = -vi; break;
            default: r = vi; m = -vr; break;
            }
            lr += r;
            li += m;
        }
        cplx<float> o;
        o.re = (float)lr;
        o.im = (float)li;
        lam[i] = o;
        cplx<float> p = psi[i];
        acc += (double)p.re * lr + (double)p.im * li;
    }
    __syncthreads();
    return block_reduce(acc, sh);
}

__global__ void k_batch_angles(const double *__restrict__ angles,
                               const uint8_t *__restrict__ kinds, CS *__restrict__ cs,
                               uint64_t nops)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nops) return;
    float c = 0.f, s = 0.f;
    switch (kinds[i]) {
    case QSLHIP_PAULI_Z: case QSLHIP_CZ:  // the phase update with c = -1: a negation
        c = -1.f;
        break;
    case QSLHIP_PHASE: case QSLHIP_CPHASE:
        sincosf((float)angles[i], &s, &c);
        break;
    case QSLHIP_ROT_X: case QSLHIP_ROT_Y: case QSLHIP_ROT_Z:
    case QSLHIP_CROT_X: case QSLHIP_CROT_Y: case QSLHIP_CROT_Z:
    case QSLHIP_NP_ROT_X: case QSLHIP_NP_ROT_Y:
        sincosf((float)angles[i] * 0.5f, &s, &c);
        break;
    default:
        break;
    }
    cs[i].c = c;
    cs[i].s = s;
}

// At most 128 VGPRs (4 waves per SIMD): two 512-thread workgroups per CU,
// which the ~70 KiB of LDS per workgroup allows.
// The tables are separate __restrict__ arguments: with nothing in the
// kernel writing them, their uniform reads become scalar loads.
__global__ __launch_bounds__(kBBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(
    const BCirc *__restrict__ circ, const BPass *__restrict__ pass, const UOp *__restrict__ uops,
    const uint32_t *__restrict__ segs, const qslhip_pauli *__restrict__ terms,
    const CS *__restrict__ cs, const uint32_t *__restrict__ order, uint32_t ncirc,
    uint32_t *__restrict__ queue, double *__restrict__ energy, double *__restrict__ grads,
    cplx<float> *__restrict__ slots, uint64_t slot_amps)
{
    const Kargs d{circ, pass, uops, segs, terms, cs, order, ncirc, queue, energy, grads, slots,
                  slot_amps};
    // LDS: psi and lambda tiles of 2^12 FP32 amplitudes (2 x 32 KiB), the
    // offsets of the scattered tile qubits, the gradient accumulators.
    __shared__ cplx<float> spsi[1 << kTileQubits];
    __shared__ cplx<float> slam[1 << kTileQubits];
    __shared__ uint64_t hoff[1 << (kTileQubits - kLowRun)];
    __shared__ double acc[kMaxPassOps * kBWaves];
    __shared__ double sh[kBWaves];
    __shared__ uint32_t s_next;

    cplx<float> *gpsi = d.slots + (uint64_t)blockIdx.x * 2 * d.slot_amps;
    cplx<float> *glam = gpsi + d.slot_amps;

    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(d.queue, 1u);
        __syncthreads();
        const uint32_t qi = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();
        if (qi >= d.ncirc) break;
        const uint32_t ci = d.order[qi];
        const BCirc c = d.circ[ci];
        const uint64_t namps = 1ull << c.n;
        const uint64_t full = namps - 1;
        const bool lds = c.lds != 0;
        const uint32_t nt = lds ? (uint32_t)namps : (1u << kTileQubits);
        const uint64_t ngroups = lds ? 1 : namps >> kTileQubits;
        for (uint32_t g = threadIdx.x; g < c.nop; g += kBBlock) d.grads[c.op_first + g] = 0.0;
        if (lds) {
            for (uint32_t j = threadIdx.x; j < nt; j += kBBlock) {
                spsi[j].re = j == 0 ? 1.f : 0.f;
                spsi[j].im = 0.f;
            }
            __syncthreads();
        }

        // ---- forward -------------------------------------------------------
        // qubits some tile so far held: from |0...0>, an amplitude whose
        // index has a 1 on any other qubit is still zero
        uint64_t support = 0;
        for (uint32_t p = 0; p < c.npass; p++) {
            const BPass P = d.pass[c.pass_first + p];
            GroupMap gm;
            gm.gbase = 0;
            gm.lowm = (1u << P.crun) - 1;
            gm.crun = P.crun;
            gm.hoff = hoff;
            if (!lds) {
                const uint64_t hi_mask = P.tile_mask & ~((1ull << P.crun) - 1);
                for (int m = threadIdx.x; m < (1 << (kTileQubits - (int)P.crun)); m += kBBlock)
                    hoff[m] = deposit((uint64_t)m, hi_mask);
                __syncthreads();
            }
            support |= P.tile_mask;
            for (uint64_t g = 0; g < ngroups; g++) {
                gm.gbase = lds ? 0 : deposit(g, full & ~P.tile_mask);
                if (!lds && (gm.gbase & ~support)) {
                    // A zero group stays zero.  Pass 0 stores its zeros (the
                    // slot still holds the previous circuit); a later pass
                    // stores nothing: each of its indices has a 1 outside
                    // every earlier pass's support too, so it lay in a zero
                    // group of pass 0 and no pass since has written it.
                    if (p == 0) {
#pragma unroll
                        for (int e = 0; e < kElems; e++) {
                            cplx<float> z;
                            z.re = 0.f;
                            z.im = 0.f;
                            gpsi[gm.at(threadIdx.x + (uint32_t)e * kBBlock)] = z;
                        }
                    }
                    continue;
                }
                group_fwd(d.uops, d.cs, d.segs, P, gm, !lds, p == 0, gpsi, spsi, nt);
            }
            // zero groups end without a barrier: every thread's stores (and
            // its reads of hoff) are done before the next pass rewrites hoff
            // and before lambda = H psi reads the slot
            __syncthreads();
        }
        if (c.npass == 0 && !lds) {  // no ops: the state stays |0...0>
            for (uint64_t i = threadIdx.x; i < namps; i += kBBlock) {
                cplx<float> z;
                z.re = (i == 0) ? 1.f : 0.f;
                z.im = 0.f;
                gpsi[i] = z;
            }
            __syncthreads();
        }
        const double e = lds ? apply_h(spsi, slam, namps, d.terms + c.term_first, (int)c.nterm, sh)
                             : apply_h(gpsi, glam, namps, d.terms + c.term_first, (int)c.nterm, sh);
        if (threadIdx.x == 0) d.energy[ci] = e;
        __syncthreads();

        // ---- backward ------------------------------------------------------
        for (uint32_t p = c.npass; p-- > 0;) {
            const BPass P = d.pass[c.pass_first + p];
            GroupMap gm;
            gm.gbase = 0;
            gm.lowm = (1u << P.crun) - 1;
            gm.crun = P.crun;
            gm.hoff = hoff;
            if (!lds) {
                const uint64_t hi_mask = P.tile_mask & ~((1ull << P.crun) - 1);
                for (int m = threadIdx.x; m < (1 << (kTileQubits - (int)P.crun)); m += kBBlock)
                    hoff[m] = deposit((uint64_t)m, hi_mask);
            }
            for (uint32_t i = threadIdx.x; i < kMaxPassOps * kBWaves; i += kBBlock) acc[i] = 0;
            __syncthreads();
            for (uint64_t g = 0; g < ngroups; g++) {
                gm.gbase = lds ? 0 : deposit(g, full & ~P.tile_mask);
                group_bwd(d.uops, d.cs, d.segs, P, gm, !lds, p > 0, gpsi, glam, spsi, slam, nt,
                          acc);
            }
            for (uint32_t k = threadIdx.x; k < P.top_count; k += kBBlock) {
                const UOpD u = uop_at(d.uops, P.top_first + k);
                if (u.gen != GEN_NONE) {
                    double s = 0;
                    for (int w = 0; w < kBWaves; w++) s += acc[k * kBWaves + w];
                    d.grads[u.gop] += s;
                }
            }
            __syncthreads();
        }
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0; }

}  // namespace

// Host-side batch object.
struct Batch {
    int prec = 4;
    uint32_t ncirc = 0;
    uint64_t nops = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf circ, pass, uops, segs, terms, angles, kinds, cs, order, queue, energy, grads, slots;
    PinnedBuf pin_angles, pin_out;
    unsigned grid = 0;
    Kargs dev{};
};

namespace {

// An op of a pass, with its tile positions (-1: qubit outside the tile).
struct PlanOp {
    int kind;
    int pa, pb;
    uint32_t q0, q1;
    uint32_t gop;
};

// Tile positions a micro-op needs as block positions: its non-diagonal
// targets (diagonal gates and controls read their bits from anywhere).
uint32_t block_need(const PlanOp &o)
{
    if (o.kind == QSLHIP_ROT_Z || o.kind == QSLHIP_CROT_Z) return 0;
    switch (pair_class(o.kind)) {
    case PC_PAIR1:
        QH_REQUIRE(o.pa >= 0, QSLHIP_EINVAL, "batch plan: target outside its tile");
        return 1u << o.pa;
    case PC_CTRL:
        QH_REQUIRE(o.pb >= 0, QSLHIP_EINVAL, "batch plan: target outside its tile");
        return 1u << o.pb;
    case PC_SWAPLIKE:
    case PC_FSWAP:
        QH_REQUIRE(o.pa >= 0 && o.pb >= 0, QSLHIP_EINVAL,
                   "batch plan: swap-like op outside its tile");
        return (1u << o.pa) | (1u << o.pb);
    default:
        return 0;
    }
}

// The micro-op of o in a segment on block positions s[0..2].
UOp decode(const PlanOp &o, const int s[3])
{
    auto src = [&](int pos, uint32_t q) -> uint8_t {
        if (pos < 0) return (uint8_t)(kSrcGroup + q);
        for (int k = 0; k < 3; k++)
            if (s[k] == pos) return (uint8_t)k;
        return (uint8_t)(kSrcThread + pos);
    };
    auto blk = [&](int pos) -> uint8_t {
        for (int k = 0; k < 3; k++)
            if (s[k] == pos) return (uint8_t)k;
        throw Error(QSLHIP_EINVAL, "batch plan: target not on a block position");
    };
    struct {
        uint8_t code = U_NONE, body = 0, gen = 0, fswap = 0;
        uint8_t ta = kSrcNone, tb = kSrcNone, ctl = kSrcNone;
    } u;
    u.gen = (uint8_t)gen_of(o.kind);
    const int cls = pair_class(o.kind);
    if (o.kind == QSLHIP_ROT_Z) {
        u.code = U_RZ;
        u.ta = src(o.pa, o.q0);
    } else if (o.kind == QSLHIP_CROT_Z) {
        u.code = U_RZ;
        u.ta = src(o.pb, o.q1);
        u.ctl = src(o.pa, o.q0);
    } else if (cls == PC_ONE) {
        u.code = U_ONE;
        u.body = (uint8_t)o.kind;
        u.ta = src(o.pa, o.q0);
    } else if (cls == PC_DIAG2) {
        u.code = U_ONE;
        u.body = (uint8_t)body_of(o.kind);
        u.ta = src(o.pa, o.q0);
        u.tb = src(o.pb, o.q1);
    } else if (cls == PC_PAIR1) {
        u.code = U_PAIR;
        u.body = (uint8_t)o.kind;
        u.ta = blk(o.pa);
    } else if (cls == PC_CTRL) {
        u.code = U_PAIR;
        u.body = (uint8_t)body_of(o.kind);
        u.ta = blk(o.pb);
        u.ctl = src(o.pa, o.q0);
    } else {  // PC_SWAPLIKE, PC_FSWAP
        u.code = U_SWAP;
        u.body = (uint8_t)body_of(o.kind);
        u.fswap = cls == PC_FSWAP;
        u.ta = blk(o.pa);
        u.tb = blk(o.pb);
    }
    UOp w;
    w.w0 = (uint32_t)u.code | (uint32_t)u.body << 8 | (uint32_t)u.gen << 16 |
           (uint32_t)u.fswap << 24;
    w.w1 = (uint32_t)u.ta | (uint32_t)u.tb << 8 | (uint32_t)u.ctl << 16;
    w.gop = o.gop;
    w.pad = 0;
    return w;
}

// Cut a pass's ops into segments on at most three block positions (a new
// segment starts when the ops' non-diagonal targets would span more; a
// segment with fewer takes the lowest free positions) and decode them.
void add_segments(std::vector<uint32_t> &segs, std::vector<UOp> &uops, const PlanOp *ops,
                  uint32_t count)
{
    uint32_t first = 0, req = 0;
    auto flush = [&](uint32_t end) {
        if (end == first) return;
        uint32_t m = req;
        for (unsigned p = 0; __builtin_popcount(m) < 3; p++) m |= 1u << p;
        int s[3], k = 0;
        for (int p = 0; p < 32 && k < 3; p++)
            if ((m >> p) & 1u) s[k++] = p;
        QH_REQUIRE(s[2] < kTileQubits, QSLHIP_EINVAL, "batch plan: block position out of range");
        segs.push_back((uint32_t)s[0] | (uint32_t)s[1] << 4 | (uint32_t)s[2] << 8 | first << 12 |
                       (end - first) << 20);
        for (uint32_t i = first; i < end; i++) uops.push_back(decode(ops[i], s));
        first = end;
    };
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t nd = block_need(ops[i]);
        if (__builtin_popcount(req | nd) > 3) {
            flush(i);
            req = 0;
        }
        req |= nd;
    }
    flush(count);
}

void batch_plan(Batch &b, const uint32_t *nq, const uint64_t *op_off, const qslhip_op *ops,
                const uint64_t *term_off, const qslhip_pauli *terms)
{
    std::vector<BCirc> circ(b.ncirc);
    std::vector<BPass> passes;
    std::vector<UOp> uops;
    std::vector<uint32_t> segs;
    std::vector<uint8_t> kinds(b.nops);
    uint32_t nmax_tile = 0;
    for (uint32_t c = 0; c < b.ncirc; c++) {
        const unsigned n = nq[c];
        QH_REQUIRE(n >= 1 && n <= 26, QSLHIP_EINVAL, "batch circuits take 1..26 qubits");
        BCirc &C = circ[c];
        C.n = n;
        C.op_first = (uint32_t)op_off[c];
        C.nop = (uint32_t)(op_off[c + 1] - op_off[c]);
        C.term_first = (uint32_t)term_off[c];
        C.nterm = (uint32_t)(term_off[c + 1] - term_off[c]);
        C.lds = n <= (unsigned)kLdsQubits;
        C.pass_first = (uint32_t)passes.size();
        for (uint64_t t = term_off[c]; t < term_off[c + 1]; t++)
            QH_REQUIRE(terms[t].xmask < (1ull << n) && terms[t].zmask < (1ull << n),
                       QSLHIP_EINVAL, "Pauli term on a qubit outside its circuit");
        std::vector<GateDesc> gates;
        for (uint64_t i = op_off[c]; i < op_off[c + 1]; i++) {
            const qslhip_op &o = ops[i];
            bool ok1 = is_gate1(o.kind) && o.q0 < n;
            bool ok2 = is_gate2(o.kind) && o.q0 < n && o.q1 < n && o.q0 != o.q1;
            QH_REQUIRE(ok1 || ok2, QSLHIP_EINVAL, "bad op " + std::to_string(i));
            gates.push_back({o.kind, o.q0, is_gate2(o.kind) ? o.q1 : 0u, 0.0, 0.0});
            kinds[i] = (uint8_t)o.kind;
        }
        // ops first..first+count-1 of the applied order; ord maps that order
        // to the circuit's (null: program order)
        auto add_pass = [&](uint64_t mask, unsigned crun, size_t first, size_t count,
                            const TileOp *tops, const uint32_t *ord) {
            BPass P;
            P.tile_mask = mask;
            P.crun = crun;
            P.pad = 0;
            P.top_first = (uint32_t)uops.size();
            P.top_count = (uint32_t)count;
            std::vector<PlanOp> pops(count);
            for (size_t i = 0; i < count; i++) {
                const size_t gi = ord ? ord[first + i] : first + i;
                const GateDesc &g = gates[gi];
                PlanOp &o = pops[i];
                o.kind = g.kind;
                o.q0 = g.q0;
                o.q1 = g.q1;
                o.gop = C.op_first + (uint32_t)gi;
                if (tops) {
                    QH_REQUIRE(tops[i].kind == g.kind, QSLHIP_EINVAL,
                               "batch plan: tile op out of step with its gate");
                    o.pa = tops[i].pa;
                    o.pb = tops[i].pb;
                } else {  // whole state in LDS: tile index == basis index
                    o.pa = (int)g.q0;
                    o.pb = is_gate2(g.kind) ? (int)g.q1 : -1;
                }
            }
            P.seg_first = (uint32_t)segs.size();
            add_segments(segs, uops, pops.data(), (uint32_t)count);
            P.seg_count = (uint32_t)segs.size() - P.seg_first;
            QH_REQUIRE(uops.size() == P.top_first + count, QSLHIP_EINVAL,
                       "batch plan: micro-ops out of step with the pass");
            passes.push_back(P);
        };
        if (C.lds) {
            for (size_t f = 0; f < gates.size(); f += kMaxPassOps)
                add_pass((1ull << n) - 1, n, f,
                         std::min<size_t>(kMaxPassOps, gates.size() - f), nullptr, nullptr);
        } else {
            nmax_tile = std::max(nmax_tile, n);
            std::vector<TilePass> tp;
            std::vector<TileOp> to;
            // commuting gates regrouped into fewer passes by the gate path's
            // support-aware reordering planner (the unitary, and so every
            // gradient, is unchanged); each op keeps its gate's gradient slot
            std::vector<uint32_t> ord;
            plan_tile_passes_reorder(n, gates, kTileQubits, kLowRun, tp, to, ord);
            QH_REQUIRE(ord.size() == gates.size() && to.size() == gates.size(), QSLHIP_EINVAL,
                       "batch plan: reordering lost ops");
            for (const TilePass &P : tp) {
                unsigned crun = 0;
                while (crun < 64 && ((P.tile_mask >> crun) & 1ull)) crun++;
                if (crun > (unsigned)kTileQubits) crun = kTileQubits;
                QH_REQUIRE(kTileQubits - crun <= kTileQubits - kLowRun, QSLHIP_EINVAL,
                           "tile pass with too many scattered qubits");
                for (uint32_t f = 0; f < P.op_count; f += kMaxPassOps) {
                    uint32_t cnt = std::min<uint32_t>(kMaxPassOps, P.op_count - f);
                    add_pass(P.tile_mask, crun, P.op_first + f, cnt, to.data() + P.op_first + f,
                             ord.data());
                }
            }
        }
        C.npass = (uint32_t)passes.size() - C.pass_first;
    }
    // largest circuits first
    std::vector<uint32_t> order(b.ncirc);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return circ[a].n > circ[c].n; });

    hipDeviceProp_t prop;
    int dev = 0;
    QH_CHECK(hipGetDevice(&dev));
    QH_CHECK(hipGetDeviceProperties(&prop, dev));
    int per_cu = 0;
    QH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_batch,
                                                          kBBlock, 0));
    if (per_cu < 1) per_cu = 1;
    b.grid = std::min<uint32_t>(b.ncirc, (uint32_t)(prop.multiProcessorCount * per_cu));
    if (b.grid < 1) b.grid = 1;
    const uint64_t slot_amps = nmax_tile ? (1ull << nmax_tile) : 1;
    // Each persistent workgroup owns a slot of 2 x slot_amps FP32 amplitudes
    // (psi and lambda).  Fewer workgroups give the same results (circuits
    // come from a queue), so the grid shrinks to what 3/4 of the free device
    // memory holds; with one workgroup the slot itself must fit.
    {
        size_t free_b = 0, total_b = 0;
        QH_CHECK(hipMemGetInfo(&free_b, &total_b));
        const uint64_t slot_bytes = 2 * slot_amps * sizeof(cplx<float>);
        const uint64_t fit = (uint64_t)(free_b / 4 * 3) / slot_bytes;
        QH_REQUIRE(fit >= 1, QSLHIP_ENOMEM,
                   "one batch slot of " + std::to_string(slot_bytes) +
                       " bytes does not fit in free device memory");
        if (fit < b.grid) b.grid = (uint32_t)fit;
    }

    auto up = [&](DevBuf &buf, const void *src, size_t bytes) {
        buf.ensure(bytes ? bytes : 8);
        if (bytes) QH_CHECK(hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice));
    };
    up(b.circ, circ.data(), circ.size() * sizeof(BCirc));
    up(b.pass, passes.data(), passes.size() * sizeof(BPass));
    up(b.uops, uops.data(), uops.size() * sizeof(UOp));
    up(b.segs, segs.data(), segs.size() * sizeof(uint32_t));
    up(b.terms, terms, term_off[b.ncirc] * sizeof(qslhip_pauli));
    up(b.order, order.data(), order.size() * sizeof(uint32_t));
    up(b.kinds, kinds.data(), kinds.size());
    b.angles.ensure(b.nops * sizeof(double) + 8);
    b.cs.ensure(b.nops * sizeof(CS) + 8);
    b.queue.ensure(64);
    b.energy.ensure(b.ncirc * sizeof(double) + 8);
    b.grads.ensure(b.nops * sizeof(double) + 8);
    b.slots.ensure((size_t)b.grid * 2 * slot_amps * sizeof(cplx<float>));
    b.pin_angles.ensure(b.nops * sizeof(double) + 8);
    b.pin_out.ensure((b.ncirc + b.nops) * sizeof(double) + 8);

    Kargs &d = b.dev;
    d.circ = static_cast<const BCirc *>(b.circ.ptr);
    d.pass = static_cast<const BPass *>(b.pass.ptr);
    d.uops = static_cast<const UOp *>(b.uops.ptr);
    d.segs = static_cast<const uint32_t *>(b.segs.ptr);
    d.terms = static_cast<const qslhip_pauli *>(b.terms.ptr);
    d.cs = static_cast<const CS *>(b.cs.ptr);
    d.order = static_cast<const uint32_t *>(b.order.ptr);
    d.ncirc = b.ncirc;
    d.queue = static_cast<uint32_t *>(b.queue.ptr);
    d.energy = static_cast<double *>(b.energy.ptr);
    d.grads = static_cast<double *>(b.grads.ptr);
    d.slots = static_cast<cplx<float> *>(b.slots.ptr);
    d.slot_amps = slot_amps;

    // Capture once: angles in, their cos / sin, queue reset (a kernel, not a
    // memset node), the persistent kernel, results out.
    QH_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    QH_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    if (b.nops) {
        QH_CHECK(hipMemcpyAsync(b.angles.ptr, b.pin_angles.ptr, b.nops * sizeof(double),
                                hipMemcpyHostToDevice, b.stream));
        const unsigned ag = (unsigned)((b.nops + 255) / 256);
        hipLaunchKernelGGL(k_batch_angles, dim3(ag), dim3(256), 0, b.stream,
                           static_cast<const double *>(b.angles.ptr),
                           static_cast<const uint8_t *>(b.kinds.ptr), static_cast<CS *>(b.cs.ptr),
                           b.nops);
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, b.stream, d.queue);
    hipLaunchKernelGGL(k_batch, dim3(b.grid), dim3(kBBlock), 0, b.stream, d.circ, d.pass, d.uops,
                       d.segs, d.terms, d.cs, d.order, d.ncirc, d.queue, d.energy, d.grads,
                       d.slots, d.slot_amps);
    double *po = static_cast<double *>(b.pin_out.ptr);
    QH_CHECK(hipMemcpyAsync(po, b.energy.ptr, b.ncirc * sizeof(double),
                            hipMemcpyDeviceToHost, b.stream));
    if (b.nops)
        QH_CHECK(hipMemcpyAsync(po + b.ncirc, b.grads.ptr, b.nops * sizeof(double),
                                hipMemcpyDeviceToHost, b.stream));
    QH_CHECK(hipStreamEndCapture(b.stream, &b.graph));
    QH_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

}  // namespace

}  // namespace qslhip

using namespace qslhip;

struct qslhip_batch {
    Batch b;
};

extern "C" int qslhip_batch_destroy(qslhip_batch_t bt);

namespace qslhip {
// status wrapper shared with api.cpp
int guarded_call(const std::function<void()> &f);
}

extern "C" {

int qslhip_batch_create(int prec, uint32_t ncirc, const uint32_t *nqubits,
                        const uint64_t *op_offsets, const qslhip_op *ops,
                        const uint64_t *term_offsets, const qslhip_pauli *terms,
                        qslhip_batch_t *out)
{
    return guarded_call([&] {
        QH_REQUIRE(out, QSLHIP_EINVAL, "null output");
        *out = nullptr;
        QH_REQUIRE(prec == 4, QSLHIP_EINVAL, "batched circuits run in FP32 (prec 4)");
        QH_REQUIRE(ncirc >= 1 && nqubits && op_offsets && term_offsets, QSLHIP_EINVAL,
                   "empty batch or null arrays");
        QH_REQUIRE(op_offsets[0] == 0 && term_offsets[0] == 0, QSLHIP_EINVAL,
                   "offsets must start at 0");
        for (uint32_t c = 0; c < ncirc; c++)
            QH_REQUIRE(op_offsets[c + 1] >= op_offsets[c] &&
                           term_offsets[c + 1] >= term_offsets[c],
                       QSLHIP_EINVAL, "offsets must be non-decreasing");
        QH_REQUIRE(op_offsets[ncirc] < (1ull << 31), QSLHIP_EINVAL, "too many ops");
        QH_REQUIRE(ops || op_offsets[ncirc] == 0, QSLHIP_EINVAL, "null ops");
        QH_REQUIRE(terms || term_offsets[ncirc] == 0, QSLHIP_EINVAL, "null terms");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
            throw Error(QSLHIP_ENODEV, "no HIP device");
        auto *bt = new qslhip_batch();
        bt->b.prec = prec;
        bt->b.ncirc = ncirc;
        bt->b.nops = op_offsets[ncirc];
        try {
            batch_plan(bt->b, nqubits, op_offsets, ops, term_offsets, terms);
            // default angles: the ones in the op list
            double *pa = static_cast<double *>(bt->b.pin_angles.ptr);
            for (uint64_t i = 0; i < bt->b.nops; i++) pa[i] = ops[i].angle;
        } catch (...) {
            qslhip_batch_destroy(bt);
            throw;
        }
        *out = bt;
    });
}

int qslhip_batch_run(qslhip_batch_t bt, const double *angles, double *energies,
                     double *grads)
{
    return guarded_call([&] {
        QH_REQUIRE(bt, QSLHIP_EINVAL, "null batch");
        Batch &b = bt->b;
        if (angles)
            std::memcpy(b.pin_angles.ptr, angles, b.nops * sizeof(double));
        QH_CHECK(hipGraphLaunch(b.exec, b.stream));
        QH_CHECK(hipStreamSynchronize(b.stream));
        const double *po = static_cast<const double *>(b.pin_out.ptr);
        if (energies) std::memcpy(energies, po, b.ncirc * sizeof(double));
        if (grads) std::memcpy(grads, po + b.ncirc, b.nops * sizeof(double));
    });
}

int qslhip_batch_stream(qslhip_batch_t bt, void **stream)
{
    return guarded_call([&] {
        QH_REQUIRE(bt && stream, QSLHIP_EINVAL, "null argument");
        *stream = bt->b.stream;
    });
}

int qslhip_batch_destroy(qslhip_batch_t bt)
{
    return guarded_call([&] {
        if (!bt) return;
        Batch &b = bt->b;
        if (b.stream) (void)hipStreamSynchronize(b.stream);
        if (b.exec) (void)hipGraphExecDestroy(b.exec);
        if (b.graph) (void)hipGraphDestroy(b.graph);
        for (DevBuf *d : {&b.circ, &b.pass, &b.uops, &b.segs, &b.terms, &b.angles, &b.kinds,
                          &b.cs, &b.order, &b.queue, &b.energy, &b.grads, &b.slots})
            d->release();
        b.pin_angles.release();
        b.pin_out.release();
        if (b.stream) (void)hipStreamDestroy(b.stream);
        delete bt;
    });
}

}  // extern "C"
