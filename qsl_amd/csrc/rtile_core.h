// rtile_core.h -- device helpers of the register-tile passes, shared by the
// ahead-of-time kernel (regtile.hip) and the passes compiled at run time
// (rtile_jit.cpp).  Device code only, so hipRTC builds it as is.
//
// A workgroup of NT = 2^tb threads owns one group of 2^K amplitudes, R =
// 2^r = 2^(K-tb) per thread, in registers; the layout rho (r of the K tile
// positions) says which tile bits index a thread's registers: the tile index
// of register k of thread t is dep(t, ~rho) | dep(k, rho).
#pragma once

#include "gate_core.h"

namespace qslhip {

constexpr int kRegMaxR = 16;  // registers per thread, at most

// LDS element order of the remap buffer.  A 128-byte bank sweep holds
// B = 128 / A amplitudes (A bytes each); the slot of element e is its low
// log2(B) bits XORed with the fold of its higher bits, so lanes whose
// elements differ only above the slot bits still land in distinct banks.
// The map is linear over XOR: swz(a ^ b) = swz(a) ^ swz(b).
template <int A>
__host__ __device__ constexpr uint32_t lds_swz(uint32_t e)
{
    constexpr uint32_t B = 128 / A;
    constexpr uint32_t L = B == 8 ? 3 : 4;
    const uint32_t x = e >> L;
    return e ^ ((x ^ (x >> L) ^ (x >> (2 * L))) & (B - 1));
}

// Where the tile of one workgroup lives: the tile index j = t | k << tb of
// the load/store layout maps to the state index gbase | dep(t) | kpart(k).
struct RegTileArgs {
    uint64_t tile_mask;          // the K tile qubits
    uint64_t tlow;               // load layout: the qubits on thread bits (ascending)
    uint64_t kq[4];              // load layout: qubit of register bit i, as a one-bit mask
    uint64_t tlow_out;           // the same for the store layout
    uint64_t kq_out[4];
    uint32_t rho_in, rho_out;    // tile positions on register bits: load, store layout
    uint32_t koff0[kRegMaxR];    // swizzled LDS byte offsets of the load layout
    uint64_t basis;              // ~0: load the state; else the state is this
                                 // basis state, synthesized without a read
    double *runs;                // non-null (tile holds qubits 0..2): write the
                                 // |a|^2 sum of every 8-amplitude run here
    uint64_t gvar, gfix;         // group g starts at deposit(g, gvar) | gfix:
                                 // gvar = the outside qubits, gfix = 0 for a
                                 // pass over every group
    uint64_t zmask, zval;        // amplitudes whose bits under zmask differ
                                 // from zval are zero and not in memory yet
    uint32_t stream;             // streaming (nontemporal) accesses, for states far
                                 // larger than the caches: bit 0 loads, bit 1 stores
    uint32_t gorder;             // group order (group_of)
};

// The group that the v-th workgroup slot of a pass takes.  Order 0: group v.
// Order 1: with workgroups dispatched to the eight XCDs in turn (slot v on
// XCD v % 8), XCD x sweeps the groups x * ngroups / 8 .. (x + 1) * ngroups / 8
// in order instead of every eighth group (ngroups and the grid multiples of
// 8).
__device__ __forceinline__ uint64_t group_of(uint64_t v, uint64_t ngroups, uint32_t gorder)
{
    return gorder ? (v & 7u) * (ngroups >> 3) + (v >> 3) : v;
}

// The sampling sums of the amplitudes about to be stored, in the pass's
// group order: with tile positions 0..2 (qubits 0..2) on lane bits 0..2,
// eight lanes of register k hold one 8-amplitude run, tile index
// jt | kt[k] with bits 0..2 clear (jt: the thread's tile-index bits, kt[k]:
// register k's, in the store layout), summed as the sampler sums a run of
// the stored state (amp_prob, lane_sum_up<3>).  Run l of group g goes to
// runs[g * 2^(K-3) + l].
template <typename T, int R, int NT>
__device__ __forceinline__ void reg_run_sums(const cplx<T> (&v)[R], double *runs, uint64_t g,
                                             uint32_t jt, const uint32_t (&kt)[R])
{
    constexpr int tb = NT == 256 ? 8 : 9;
    constexpr int lb = tb - 3 + (R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : 5);  // K - 3
#pragma unroll
    for (int k = 0; k < R; k++) {
        const double s = lane_sum_up<3>(amp_prob(v[k]));
        if ((jt & 7u) == 0) runs[(g << lb) + ((jt | kt[k]) >> 3)] = s;
    }
}

// Tile index bits of thread t in a layout with registers on positions rho:
// t's bits in the other positions, ascending (zeros inserted at rho).
__host__ __device__ inline uint32_t thread_tile_bits(uint32_t t, uint32_t rho)
{
    for (; rho; rho &= rho - 1) {
        const uint32_t lowm = (rho & (0u - rho)) - 1;
        t = (t & lowm) | ((t & ~lowm) << 1);
    }
    return t;
}

// Tile index bits of register k in a layout with registers on positions rho.
__host__ __device__ inline uint32_t reg_tile_bits(uint32_t k, uint32_t rho)
{
    uint32_t out = 0;
    for (int b = 0; rho; b++, rho &= rho - 1)
        if ((k >> b) & 1u) out |= rho & (0u - rho);
    return out;
}

// One thread's R amplitudes of a group: state indices idx0 | kpart[k].  With
// a basis index (ta.basis != ~0) the state is |basis>, never written to
// memory: 1 + 0i at that index and +0 elsewhere, the values a fill of the
// vector would hold.  Otherwise an amplitude whose bits under ta.zmask
// differ from ta.zval is known to be zero and never written yet: it is +0
// and not loaded (zmask = 0: load everything).
template <typename T, int R>
__device__ __forceinline__ void reg_load(cplx<T> (&v)[R], const cplx<T> *st, uint64_t idx0,
                                         const uint64_t (&kpart)[R], const RegTileArgs &ta)
{
    if (ta.basis != ~0ull) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            v[k].re = (idx0 | kpart[k]) == ta.basis ? T(1) : T(0);
            v[k].im = T(0);
        }
    } else if (ta.zmask) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            const uint64_t i = idx0 | kpart[k];
            if ((i & ta.zmask) == ta.zval) {
                v[k] = (ta.stream & 1u) ? amp_ld<true>(st + i) : amp_ld<false>(st + i);
            } else {
                v[k].re = T(0);
                v[k].im = T(0);
            }
        }
    } else if (ta.stream & 1u) {
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = amp_ld<true>(st + (idx0 | kpart[k]));
    } else {
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = amp_ld<false>(st + (idx0 | kpart[k]));
    }
}

// The R amplitudes back to state indices idx0 | kpart[k].
template <typename T, int R>
__device__ __forceinline__ void reg_store(const cplx<T> (&v)[R], cplx<T> *st, uint64_t idx0,
                                          const uint64_t (&kpart)[R], const RegTileArgs &ta)
{
    if (ta.stream & 2u) {
#pragma unroll
        for (int k = 0; k < R; k++) amp_st<true>(st + (idx0 | kpart[k]), v[k]);
    } else {
#pragma unroll
        for (int k = 0; k < R; k++) amp_st<false>(st + (idx0 | kpart[k]), v[k]);
    }
}

// rotateZ's two diagonal halves without a branch on the bit: the |1> half
// is the |0> half with s negated (x + y*(-s) is x - y*s exactly in IEEE
// arithmetic), so the result is bit-identical to gate_math.h's rz_half.
template <typename T>
__device__ __forceinline__ void rz_s(cplx<T> &x, T c, T ss)
{
    const T r = x.re, i = x.im;
    x.re = r * c + i * ss;
    x.im = i * c - r * ss;
}

// x with its sign flipped when neg (an exact negation: one XOR of the sign
// bit, with no select between two negated copies).
__device__ __forceinline__ double flip_sign(double x, bool neg)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)neg << 63));
}
__device__ __forceinline__ float flip_sign(float x, bool neg)
{
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ ((uint32_t)neg << 31));
}

template <typename T>
__device__ __forceinline__ cplx<T> pick(bool p, const cplx<T> &a, const cplx<T> &b)
{
    return cplx<T>{p ? a.re : b.re, p ? a.im : b.im};
}

// The register array must stay in registers: every index into it is a
// compile-time constant.  The empty asm markers (distinct immediates per
// instantiation) keep the optimiser from merging the cases of a runtime
// switch over register bits into one body with a variable index, which
// would move the array to scratch memory.
#define QH_CASE_MARK(A, B) asm volatile("; rtile %0 %1" ::"n"(A), "n"(B))

// Pair op with the target on register bit I; kmask (uniform) selects the
// pairs, tp (per lane, SEL only) the lanes whose control is set.
template <typename T, int R, int I, int BODY, bool SEL, bool FULL, bool EX>
__device__ __forceinline__ void reg_pair(cplx<T> (&v)[R], uint32_t kmask, bool tp, T c, T s)
{
    QH_CASE_MARK(I, BODY + 64 * SEL + 128 * EX + 256 * FULL);
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (k & (1 << I)) continue;
        if (!FULL && !((kmask >> k) & 1u)) continue;
        cplx<T> a = v[k], b = v[k | (1 << I)];
        if constexpr (EX) pair_update<T>(BODY, a, b, c, s);
        else pair_update_fast<T>(BODY, a, b, c, s);
        if (SEL) {
            v[k] = pick(tp, a, v[k]);
            v[k | (1 << I)] = pick(tp, b, v[k | (1 << I)]);
        } else {
            v[k] = a;
            v[k | (1 << I)] = b;
        }
    }
    QH_CASE_MARK(I, BODY + 64 * SEL + 128 * EX + 256 * FULL);
}

// Swap-like op on register bits IA (q0) and IB (q1): |q0=1,q1=0> paired
// with |q0=0,q1=1>; fswap also negates |11>.
template <typename T, int R, int IA, int IB, int BODY, bool FSWAP, bool EX>
__device__ __forceinline__ void reg_swaplike(cplx<T> (&v)[R], T c, T s)
{
    QH_CASE_MARK(IA * 4 + IB, BODY + 64 * FSWAP + 128 * EX);
#pragma unroll
    for (int k = 0; k < R; k++) {
        const bool a = k & (1 << IA), b = k & (1 << IB);
        if (a && !b) {
            if constexpr (EX) pair_update<T>(BODY, v[k], v[k ^ (1 << IA) ^ (1 << IB)], c, s);
            else pair_update_fast<T>(BODY, v[k], v[k ^ (1 << IA) ^ (1 << IB)], c, s);
        }
        if (FSWAP && a && b) one_update<T>(QSLHIP_PAULI_Z, v[k], c, s);
    }
    QH_CASE_MARK(IA * 4 + IB, BODY + 64 * FSWAP + 128 * EX);
}

template <typename T, int R, int BODY, bool SEL, bool FULL, bool EX>
__device__ __forceinline__ void dispatch_pair(cplx<T> (&v)[R], int ia, uint32_t kmask, bool tp,
                                              T c, T s)
{
    switch (ia) {
    case 0: reg_pair<T, R, 0, BODY, SEL, FULL, EX>(v, kmask, tp, c, s); break;
    case 1: reg_pair<T, R, 1, BODY, SEL, FULL, EX>(v, kmask, tp, c, s); break;
    case 2: if constexpr (R > 4) reg_pair<T, R, 2, BODY, SEL, FULL, EX>(v, kmask, tp, c, s); break;
    case 3: if constexpr (R > 8) reg_pair<T, R, 3, BODY, SEL, FULL, EX>(v, kmask, tp, c, s); break;
    default: break;
    }
}

template <typename T, int R, bool SEL, bool FULL, bool EX>
__device__ __forceinline__ void dispatch_body(cplx<T> (&v)[R], int body, int ia, uint32_t kmask,
                                              bool tp, T c, T s)
{
    switch (body) {
    case QSLHIP_PAULI_X: dispatch_pair<T, R, QSLHIP_PAULI_X, SEL, FULL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_PAULI_Y: dispatch_pair<T, R, QSLHIP_PAULI_Y, SEL, FULL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_HADAMARD: dispatch_pair<T, R, QSLHIP_HADAMARD, SEL, FULL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_ROT_X: dispatch_pair<T, R, QSLHIP_ROT_X, SEL, FULL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_ROT_Y: dispatch_pair<T, R, QSLHIP_ROT_Y, SEL, FULL, EX>(v, ia, kmask, tp, c, s); break;
    default: break;
    }
}

template <typename T, int R, int BODY, bool FSWAP, bool EX>
__device__ __forceinline__ void dispatch_swap(cplx<T> (&v)[R], int ia, int ib, T c, T s)
{
#define QH_SW(A, B)                                                                   \
    case (A) * 4 + (B):                                                               \
        if constexpr ((1 << (A)) < R && (1 << (B)) < R)                               \
            reg_swaplike<T, R, A, B, BODY, FSWAP, EX>(v, c, s);                       \
        break;
    switch (ia * 4 + ib) {
        QH_SW(0, 1) QH_SW(0, 2) QH_SW(0, 3) QH_SW(1, 0) QH_SW(1, 2) QH_SW(1, 3)
        QH_SW(2, 0) QH_SW(2, 1) QH_SW(2, 3) QH_SW(3, 0) QH_SW(3, 1) QH_SW(3, 2)
    default: break;
    }
#undef QH_SW
}

template <typename T, int R, int BODY, bool SEL, bool FULL, bool EX>
__device__ __forceinline__ void reg_one(cplx<T> (&v)[R], uint32_t kmask, bool tp, T c, T s)
{
    // tolerance-bound mode: lanes whose predicate is off apply the identity
    // (sign kept; phase 1 + 0i) instead of selecting between two results
    const T cl = (!EX && SEL) ? (tp ? c : (T)1) : c;
    const T sl = (!EX && SEL) ? (tp ? s : (T)0) : s;
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (!FULL && !((kmask >> k) & 1u)) continue;
        cplx<T> x = v[k];
        if constexpr (EX) {
            one_update<T>(BODY, x, c, s);
            v[k] = SEL ? pick(tp, x, v[k]) : x;
        } else if constexpr (BODY == QSLHIP_PAULI_Z) {
            const bool neg = SEL ? tp : true;
            v[k] = cplx<T>{flip_sign(x.re, neg), flip_sign(x.im, neg)};
        } else {
            one_update_fast<T>(BODY, x, cl, sl);
            v[k] = x;
        }
    }
}

// rotateZ (and controlRotateZ): the half of every register follows its
// target bit.  LANE: the target is a thread bit, so the lane's own sign sl
// (with the group's gneg) serves every register; otherwise the sign is
// uniform per register (gneg and kneg), picked on the scalar unit.
template <typename T, int R, bool SEL, bool FULL, bool LANE, bool EX>
__device__ __forceinline__ void reg_rz(cplx<T> (&v)[R], uint32_t kmask, uint32_t kneg, bool gneg,
                                       bool tp, T c, T sl)
{
    const T s_pos = flip_sign(sl, gneg), s_neg = -s_pos;
    // tolerance-bound mode: lanes whose control is off apply c = 1, s = 0
    const T cl = (!EX && SEL) ? (tp ? c : (T)1) : c;
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (!FULL && !((kmask >> k) & 1u)) continue;
        cplx<T> x = v[k];
        T ss = LANE ? s_pos : (((kneg >> k) & 1u) ? s_neg : s_pos);
        if constexpr (EX) {
            rz_s<T>(x, c, ss);
            v[k] = SEL ? pick(tp, x, v[k]) : x;
        } else {
            if (SEL) ss = tp ? ss : (T)0;
            const T r = x.re, i = x.im;
            x.re = fmadd(r, cl, i * ss);
            x.im = fmadd(i, cl, -(r * ss));
            v[k] = x;
        }
    }
}

// Workgroup barrier for the LDS remap buffer: waits for this wave's LDS
// traffic only, so global loads already issued for the next group stay in
// flight across it (__syncthreads would also drain them).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// s_waitcnt vmcnt(N) (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt and
// lgkmcnt left at "no wait"): at most N of this wave's vector memory
// operations still outstanding.  The compiler's wait insertion sees it and
// needs no waits of its own for anything older.
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

}  // namespace qslhip
