// regtile.hip -- LDS tile passes with the amplitudes held in registers.
//
// A workgroup of NT = 2^tb threads owns one group of 2^K amplitudes, R =
// 2^r = 2^(K-tb) per thread.  A "layout" rho (r of the K tile positions)
// says which tile bits index a thread's registers; the other tb index the
// thread:  tile index of register k of thread t = dep(t, ~rho) | dep(k, rho).
//
// A non-diagonal gate needs its target bit(s) among the register bits; the
// host planner (plan_reg_ops) inserts a remap (registers -> LDS ->
// registers in a new layout covering the next targets) only when they are
// not.  Controls and diagonal factors are predicates on single bits, each
// of which is a register bit (uniform per register: decided on the host,
// a scalar branch on the device), a thread bit (one lane test per op) or a
// qubit outside the tile (uniform per workgroup).  The planner decodes every
// gate into that form, so the device loop does the arithmetic and little
// else.  Remaps address LDS through an XOR-linear swizzle: the offset of
// (thread part ^ register part) is the XOR of the two offsets, the register
// parts come precomputed with the op, and one XOR per element remains.
//
// EX selects the arithmetic: the reference's exact per-gate formulas
// (bit-identical), or for QSLHIP_MODE_REORDER the tolerance-bound forms with
// fused multiply-adds (gate_math.h pair_update_fast).
//
// Loads and stores use the layout with registers on the top r positions,
// so threads sit on the lowest tile qubits (coalesced).  A workgroup may
// take several groups and prefetch the next while computing one (k_rtile).  The per-amplitude
// arithmetic is gate_math.h's in program order: results are bit-identical
// to the one-gate kernels.
#include <cstring>

#include "tile_ops.h"

namespace qslhip {

enum RegCode : int32_t { RC_REMAP = 0, RC_PAIR = 1, RC_SWAP = 2, RC_ONE = 3, RC_RZ = 4 };

constexpr int kRegMaxR = 16;

// One decoded op of a register-tile pass.
struct RegOp {
    int32_t code;    // RegCode
    int32_t body;    // gate_math body (RC_PAIR, RC_SWAP, RC_ONE)
    int32_t ia, ib;  // RC_PAIR: register bit of the target; RC_SWAP: of q0 and q1
    uint32_t kmask;  // registers taking part (RC_PAIR: the bit-ia-clear one of each pair)
    uint32_t kneg;   // RC_RZ: registers whose target bit is 1
    uint32_t tmask;  // thread predicate (jt & tmask) == tmask
    uint32_t tneg;   // RC_RZ: thread-bit target, as a tile-position mask
    uint64_t gmask;  // group predicate (gbase & gmask) == gmask
    uint64_t gneg;   // RC_RZ: target qubit outside the tile, as a mask
    double c, s;
    uint32_t rho;    // RC_REMAP: register positions of the new layout
    uint32_t fswap;  // RC_SWAP: also negate |11>
    uint32_t koff[kRegMaxR];  // RC_REMAP: swizzled LDS byte offset of register k
};

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

namespace {

// Where the tile of one workgroup lives: the tile index j = t | k << tb of
// the load/store layout maps to the state index gbase | dep(t) | kpart(k).
struct RegTileArgs {
    uint64_t tile_mask;          // the K tile qubits
    uint64_t tlow;               // the tb lowest of them (thread bits)
    uint64_t kq[4];              // qubit of tile position tb + i, as a one-bit mask
    uint32_t koff0[kRegMaxR];    // swizzled LDS byte offsets of the load layout
};

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
template <typename T, int R, int I, int BODY, bool SEL, bool EX>
__device__ __forceinline__ void reg_pair(cplx<T> (&v)[R], uint32_t kmask, bool tp, T c, T s)
{
    QH_CASE_MARK(I, BODY + 64 * SEL + 128 * EX);
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (k & (1 << I)) continue;
        if (!((kmask >> k) & 1u)) continue;
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
    QH_CASE_MARK(I, BODY + 64 * SEL + 128 * EX);
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

template <typename T, int R, int BODY, bool SEL, bool EX>
__device__ __forceinline__ void dispatch_pair(cplx<T> (&v)[R], int ia, uint32_t kmask, bool tp,
                                              T c, T s)
{
    switch (ia) {
    case 0: reg_pair<T, R, 0, BODY, SEL, EX>(v, kmask, tp, c, s); break;
    case 1: reg_pair<T, R, 1, BODY, SEL, EX>(v, kmask, tp, c, s); break;
    case 2: if constexpr (R > 4) reg_pair<T, R, 2, BODY, SEL, EX>(v, kmask, tp, c, s); break;
    case 3: if constexpr (R > 8) reg_pair<T, R, 3, BODY, SEL, EX>(v, kmask, tp, c, s); break;
    default: break;
    }
}

template <typename T, int R, bool SEL, bool EX>
__device__ __forceinline__ void dispatch_body(cplx<T> (&v)[R], int body, int ia, uint32_t kmask,
                                              bool tp, T c, T s)
{
    switch (body) {
    case QSLHIP_PAULI_X: dispatch_pair<T, R, QSLHIP_PAULI_X, SEL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_PAULI_Y: dispatch_pair<T, R, QSLHIP_PAULI_Y, SEL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_HADAMARD: dispatch_pair<T, R, QSLHIP_HADAMARD, SEL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_ROT_X: dispatch_pair<T, R, QSLHIP_ROT_X, SEL, EX>(v, ia, kmask, tp, c, s); break;
    case QSLHIP_ROT_Y: dispatch_pair<T, R, QSLHIP_ROT_Y, SEL, EX>(v, ia, kmask, tp, c, s); break;
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

template <typename T, int R, int BODY, bool SEL, bool EX>
__device__ __forceinline__ void reg_one(cplx<T> (&v)[R], uint32_t kmask, bool tp, T c, T s)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (!((kmask >> k) & 1u)) continue;
        cplx<T> x = v[k];
        if constexpr (EX) one_update<T>(BODY, x, c, s);
        else one_update_fast<T>(BODY, x, c, s);
        v[k] = SEL ? pick(tp, x, v[k]) : x;
    }
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

// rotateZ (and controlRotateZ): the half of every register follows its
// target bit: gneg (group), kneg (register k) and the lane's own sl.
template <typename T, int R, bool SEL, bool EX>
__device__ __forceinline__ void reg_rz(cplx<T> (&v)[R], uint32_t kmask, uint32_t kneg, bool gneg,
                                       bool tp, T c, T sl)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (!((kmask >> k) & 1u)) continue;
        cplx<T> x = v[k];
        const T ss = flip_sign(sl, gneg != (bool)((kneg >> k) & 1u));
        if constexpr (EX) {
            rz_s<T>(x, c, ss);
        } else {
            const T r = x.re, i = x.im;
            x.re = fmadd(r, c, i * ss);
            x.im = fmadd(i, c, -(r * ss));
        }
        v[k] = SEL ? pick(tp, x, v[k]) : x;
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

// Workgroup b takes groups b, b + gridDim.x, ... (one group when the grid
// covers them all; persistent workgroups when it is the device's resident
// count) and loads its next group into a second register set while it
// computes and stores the current one, so HBM traffic overlaps the gate
// arithmetic.
template <typename T, int NT, int R, bool EX>
__global__ __launch_bounds__(NT) void k_rtile(cplx<T> *__restrict__ st, unsigned n,
                                              RegTileArgs ta, const RegOp *__restrict__ ops,
                                              int nops, uint64_t ngroups)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int A = (int)sizeof(cplx<T>);
    constexpr int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    const uint32_t t = threadIdx.x;
    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const uint64_t outside = full & ~ta.tile_mask;
    const uint64_t tthread = deposit((uint64_t)t, ta.tlow);
    uint64_t kpart[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        uint64_t kp = 0;
#pragma unroll
        for (int i = 0; i < r; i++)
            if ((k >> i) & 1) kp |= ta.kq[i];
        kpart[k] = kp;
    }
    uint64_t g = blockIdx.x;
    if (g >= ngroups) return;
    uint64_t gbase_n = deposit(g, outside);
    cplx<T> v[R], vn[R];
#pragma unroll
    for (int k = 0; k < R; k++) vn[k] = st[gbase_n | tthread | kpart[k]];
    wait_vmcnt<0>();
    for (;;) {
        const uint64_t gbase = gbase_n;
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = vn[k];
        const uint64_t gn = g + gridDim.x;
        const bool more = gn < ngroups;
        if (more) {
            gbase_n = deposit(gn, outside);
#pragma unroll
            for (int k = 0; k < R; k++) vn[k] = st[gbase_n | tthread | kpart[k]];
            // in flight: the previous group's R stores and these R loads;
            // everything older (this group's loads) is complete
            wait_vmcnt<2 * R>();
        } else {
            wait_vmcnt<R>();  // only the previous group's stores may remain
        }
        uint32_t jt = t;  // this thread's tile-index bits (zeros at register positions)
        uint32_t base = lds_swz<A>(jt) * A;
        uint32_t koff[R];
#pragma unroll
        for (int k = 0; k < R; k++) koff[k] = ta.koff0[k];
        for (int i = 0; i < nops; i++) {
            const RegOp &o = ops[i];
            const int code = o.code;
            if (code == RC_REMAP) {
#pragma unroll
                for (int k = 0; k < R; k++)
                    *reinterpret_cast<cplx<T> *>(smem + (base ^ koff[k])) = v[k];
                lds_barrier();
                jt = t;
                for (uint32_t rho = o.rho; rho; rho &= rho - 1) {  // insert zeros at rho
                    const uint32_t lowm = (rho & (0u - rho)) - 1;
                    jt = (jt & lowm) | ((jt & ~lowm) << 1);
                }
                base = lds_swz<A>(jt) * A;
#pragma unroll
                for (int k = 0; k < R; k++) koff[k] = o.koff[k];
#pragma unroll
                for (int k = 0; k < R; k++)
                    v[k] = *reinterpret_cast<const cplx<T> *>(smem + (base ^ koff[k]));
                lds_barrier();
                continue;
            }
            if ((gbase & o.gmask) != o.gmask) continue;  // a control outside the tile is 0
            const T c = (T)o.c, s = (T)o.s;
            const uint32_t tmask = o.tmask;
            const bool sel = tmask != 0;
            const bool tp = (jt & tmask) == tmask;
            const uint32_t kmask = o.kmask;
            switch (code) {
            case RC_PAIR:
                if (sel) dispatch_body<T, R, true, EX>(v, o.body, o.ia, kmask, tp, c, s);
                else dispatch_body<T, R, false, EX>(v, o.body, o.ia, kmask, tp, c, s);
                break;
            case RC_SWAP:
                switch (o.body) {
                case QSLHIP_PAULI_X:
                    if (o.fswap) dispatch_swap<T, R, QSLHIP_PAULI_X, true, EX>(v, o.ia, o.ib, c, s);
                    else dispatch_swap<T, R, QSLHIP_PAULI_X, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                case QSLHIP_ROT_X:
                    dispatch_swap<T, R, QSLHIP_ROT_X, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                case QSLHIP_ROT_Y:
                    dispatch_swap<T, R, QSLHIP_ROT_Y, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                default:
                    dispatch_swap<T, R, QSLHIP_HADAMARD, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                }
                break;
            case RC_ONE:
                if (o.body == QSLHIP_PAULI_Z) {
                    if (sel) reg_one<T, R, QSLHIP_PAULI_Z, true, EX>(v, kmask, tp, c, s);
                    else reg_one<T, R, QSLHIP_PAULI_Z, false, EX>(v, kmask, tp, c, s);
                } else {
                    if (sel) reg_one<T, R, QSLHIP_PHASE, true, EX>(v, kmask, tp, c, s);
                    else reg_one<T, R, QSLHIP_PHASE, false, EX>(v, kmask, tp, c, s);
                }
                break;
            default: {  // RC_RZ
                const bool gneg = (gbase & o.gneg) != 0;
                const T sl = flip_sign(s, (jt & o.tneg) != 0);
                if (sel) reg_rz<T, R, true, EX>(v, kmask, o.kneg, gneg, tp, c, sl);
                else reg_rz<T, R, false, EX>(v, kmask, o.kneg, gneg, tp, c, sl);
                break;
            }
            }
        }
#pragma unroll
        for (int k = 0; k < R; k++) st[gbase | tthread | kpart[k]] = v[k];
        if (!more) break;
        g = gn;
    }
}

// Tile positions a non-diagonal body moves amplitudes across.
void targets_of(const TileOp &t, int out[2], int &cnt)
{
    cnt = 0;
    const int cls = pair_class(t.kind), body = body_of(t.kind);
    if (cls == PC_PAIR1 && body != QSLHIP_ROT_Z) out[cnt++] = t.pa;
    else if (cls == PC_CTRL && body != QSLHIP_ROT_Z) out[cnt++] = t.pb;
    else if (cls == PC_SWAPLIKE || cls == PC_FSWAP) {
        out[cnt++] = t.pa;
        out[cnt++] = t.pb;
    }
}

uint32_t dep_host(uint32_t v, uint32_t mask)
{
    uint32_t out = 0;
    for (int b = 0; mask; b++, mask &= mask - 1)
        if ((v >> b) & 1u) out |= mask & (0u - mask);
    return out;
}

uint32_t swz_bytes(uint32_t e, int A)
{
    return A == 16 ? lds_swz<16>(e) * 16 : lds_swz<8>(e) * 8;
}

}  // namespace

// Host planner: decode the TileOps of one pass (tile width K, 2^tb
// threads, A bytes per amplitude) into register ops.  The pass starts and
// ends in the load/store layout (registers on the top r positions); a remap
// goes before a gate whose targets are not register bits, to the gate's
// targets plus the next gates' targets in order (the set that postpones the
// following remap longest), topped up from the current layout.
void plan_reg_ops(const TileOp *ops, uint32_t count, int K, int tb, int A,
                  std::vector<RegOp> &out)
{
    const int r = K - tb, R = 1 << r;
    const uint32_t all_k = (R >= 32) ? ~0u : ((1u << R) - 1);
    const uint32_t rho0 = ((1u << K) - 1) & ~((1u << tb) - 1);
    uint32_t cur = rho0;
    auto slot = [&](int pos) { return __builtin_popcount(cur & ((1u << pos) - 1)); };
    auto kset = [&](int pos) {  // registers whose bit at tile position pos is 1
        const int sl = slot(pos);
        uint32_t m = 0;
        for (int k = 0; k < R; k++)
            if ((k >> sl) & 1) m |= 1u << k;
        return m;
    };
    auto remap = [&](uint32_t rho) {
        RegOp m{};
        m.code = RC_REMAP;
        m.rho = rho;
        for (int k = 0; k < R; k++) m.koff[k] = swz_bytes(dep_host((uint32_t)k, rho), A);
        out.push_back(m);
        cur = rho;
    };
    auto pred = [&](RegOp &o, int pos, uint32_t q) {  // require qubit q (tile pos) == 1
        if (pos < 0) o.gmask |= 1ull << q;
        else if ((cur >> pos) & 1u) o.kmask &= kset(pos);
        else o.tmask |= 1u << pos;
    };
    auto neg = [&](RegOp &o, int pos, uint32_t q) {  // rotateZ half follows qubit q
        if (pos < 0) o.gneg = 1ull << q;
        else if ((cur >> pos) & 1u) o.kneg = kset(pos);
        else o.tneg = 1u << pos;
    };
    for (uint32_t i = 0; i < count; i++) {
        const TileOp &t = ops[i];
        int tg[2], nt;
        targets_of(t, tg, nt);
        uint32_t need = 0;
        for (int k = 0; k < nt; k++) need |= 1u << tg[k];
        if ((need & ~cur) != 0) {
            uint32_t rho = need;
            for (uint32_t j = i + 1; j < count && __builtin_popcount(rho) < r; j++) {
                int tj[2], nj;
                targets_of(ops[j], tj, nj);
                for (int k = 0; k < nj && __builtin_popcount(rho) < r; k++) rho |= 1u << tj[k];
            }
            for (int p = K - 1; p >= 0 && __builtin_popcount(rho) < r; p--)
                if ((cur >> p) & 1u) rho |= 1u << p;
            for (int p = K - 1; p >= 0 && __builtin_popcount(rho) < r; p--) rho |= 1u << p;
            remap(rho);
        }
        RegOp o{};
        o.c = t.c;
        o.s = t.s;
        o.kmask = all_k;
        const int body = body_of(t.kind);
        switch (t.kind) {
        case QSLHIP_PAULI_X: case QSLHIP_PAULI_Y: case QSLHIP_HADAMARD:
        case QSLHIP_ROT_X: case QSLHIP_ROT_Y:
            o.code = RC_PAIR;
            o.body = body;
            o.ia = slot(t.pa);
            o.kmask = all_k & ~kset(t.pa);
            break;
        case QSLHIP_ROT_Z:
            o.code = RC_RZ;
            neg(o, t.pa, t.q0);
            break;
        case QSLHIP_PAULI_Z: case QSLHIP_PHASE:
            o.code = RC_ONE;
            o.body = body;
            pred(o, t.pa, t.q0);
            break;
        case QSLHIP_CNOT: case QSLHIP_CY: case QSLHIP_CROT_X: case QSLHIP_CROT_Y:
        case QSLHIP_CHADAMARD:
            o.code = RC_PAIR;
            o.body = body;
            o.ia = slot(t.pb);
            o.kmask = all_k & ~kset(t.pb);
            pred(o, t.pa, t.q0);
            break;
        case QSLHIP_CROT_Z:
            o.code = RC_RZ;
            pred(o, t.pa, t.q0);
            neg(o, t.pb, t.q1);
            break;
        case QSLHIP_CZ: case QSLHIP_CPHASE:
            o.code = RC_ONE;
            o.body = body;
            pred(o, t.pa, t.q0);
            pred(o, t.pb, t.q1);
            break;
        default:  // swap-like
            o.code = RC_SWAP;
            o.body = body;
            o.ia = slot(t.pa);
            o.ib = slot(t.pb);
            o.fswap = t.kind == QSLHIP_FSWAP;
            break;
        }
        out.push_back(o);
    }
    if (cur != rho0) remap(rho0);
}

namespace {

template <typename T, int NT, int R, bool EX>
void launch_one(qslhip_state *h, void *amps, const TilePass &p, const RegOp *dops, uint32_t nops,
                int gpw)
{
    constexpr int A = (int)sizeof(cplx<T>);
    constexpr int tb = NT == 256 ? 8 : 9;
    constexpr int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    const int K = tb + r;
    RegTileArgs ta{};
    ta.tile_mask = p.tile_mask;
    int pos = 0;
    for (unsigned q = 0; q < 64; q++) {
        if (!((p.tile_mask >> q) & 1ull)) continue;
        if (pos < tb) ta.tlow |= 1ull << q;
        else ta.kq[pos - tb] = 1ull << q;
        pos++;
    }
    QH_REQUIRE(pos == K && p.logt == K, QSLHIP_EINVAL, "tile mask does not hold the tile width");
    for (int k = 0; k < R; k++) ta.koff0[k] = swz_bytes((uint32_t)k << tb, A);
    const size_t shmem = (size_t)A << K;
    QH_REQUIRE(shmem <= 64 * 1024, QSLHIP_EINVAL, "register tile needs more than 64 KiB of LDS");
    QH_REQUIRE(h->n >= (unsigned)K, QSLHIP_EINVAL, "tile wider than the state");
    const uint64_t groups = 1ull << (h->n - K);
    // gpw groups per workgroup (0: persistent, as many workgroups as the
    // device holds at once)
    uint64_t grid;
    if (gpw > 0) {
        // several groups per workgroup only when groups are plentiful
        if (groups < (uint64_t)gpw * 8192) gpw = 1;
        grid = (groups + (uint64_t)gpw - 1) / (uint64_t)gpw;
    } else {
        static int resident = 0;  // per instantiation
        if (!resident) {
            int per_cu = 0, dev = 0, cus = 0;
            QH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rtile<T, NT, R, EX>,
                                                                  NT, shmem));
            QH_CHECK(hipGetDevice(&dev));
            QH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            resident = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        }
        grid = groups < (uint64_t)resident ? groups : (uint64_t)resident;
    }
    hipLaunchKernelGGL((k_rtile<T, NT, R, EX>), dim3((unsigned)grid), dim3(NT), shmem, h->stream,
                       static_cast<cplx<T> *>(amps), h->n, ta, dops, (int)nops, groups);
    QH_CHECK(hipGetLastError());
}

template <typename T, bool EX>
void launch_t(qslhip_state *h, void *amps, const TilePass &p, int nt, const RegOp *dops,
              uint32_t nops, int gpw)
{
    const int r = p.logt - (nt == 512 ? 9 : 8);
    if (nt == 256 && r == 2) launch_one<T, 256, 4, EX>(h, amps, p, dops, nops, gpw);
    else if (nt == 256 && r == 3) launch_one<T, 256, 8, EX>(h, amps, p, dops, nops, gpw);
    else if (nt == 256 && r == 4) launch_one<T, 256, 16, EX>(h, amps, p, dops, nops, gpw);
    else if (nt == 512 && r == 3) launch_one<T, 512, 8, EX>(h, amps, p, dops, nops, gpw);
    else if (nt == 512 && r == 4) launch_one<T, 512, 16, EX>(h, amps, p, dops, nops, gpw);
    else throw Error(QSLHIP_EINVAL, "no register-tile kernel for this tile width");
}

}  // namespace

bool reg_tile_ok(int logt, int threads, int prec)
{
    const int r = logt - (threads == 512 ? 9 : 8);
    if (threads != 256 && threads != 512) return false;
    if (r < 2 || r > 4 || (threads == 512 && r == 2)) return false;
    return ((size_t)(prec == 8 ? 16 : 8) << logt) <= 64 * 1024;
}

void launch_reg_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                       const std::vector<TileOp> &ops, void *amps, int threads, int gpw,
                       bool exact)
{
    const int tb = threads == 512 ? 9 : 8;
    const int A = h->prec == 8 ? 16 : 8;
    std::vector<RegOp> rops;
    std::vector<uint32_t> first(passes.size() + 1);
    for (size_t i = 0; i < passes.size(); i++) {
        QH_REQUIRE(reg_tile_ok(passes[i].logt, threads, h->prec), QSLHIP_EINVAL,
                   "no register-tile kernel for this tile width");
        first[i] = (uint32_t)rops.size();
        plan_reg_ops(ops.data() + passes[i].op_first, passes[i].op_count, passes[i].logt, tb, A,
                     rops);
    }
    first[passes.size()] = (uint32_t)rops.size();
    const size_t bytes = rops.size() * sizeof(RegOp);
    // The pinned staging buffer may still be read by the previous copy.
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    h->pin.ensure(bytes);
    h->optab.ensure(bytes);
    std::memcpy(h->pin.ptr, rops.data(), bytes);
    QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, bytes, hipMemcpyHostToDevice, h->stream));
    QH_CHECK(hipEventRecord(h->staged, h->stream));
    const RegOp *dops = static_cast<const RegOp *>(h->optab.ptr);
    for (size_t i = 0; i < passes.size(); i++) {
        const uint32_t n_ops = first[i + 1] - first[i];
        const TilePass &p = passes[i];
        const RegOp *d = dops + first[i];
        if (h->prec == 8) {
            if (exact) launch_t<double, true>(h, amps, p, threads, d, n_ops, gpw);
            else launch_t<double, false>(h, amps, p, threads, d, n_ops, gpw);
        } else {
            if (exact) launch_t<float, true>(h, amps, p, threads, d, n_ops, gpw);
            else launch_t<float, false>(h, amps, p, threads, d, n_ops, gpw);
        }
    }
}

}  // namespace qslhip
