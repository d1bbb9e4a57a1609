// regtile.hip -- LDS tile passes with the amplitudes held in registers.
//
// A workgroup of 256 threads owns one group of 2^K amplitudes (K = 10..12),
// R = 2^(K-8) per thread.  A "layout" says which r = K-8 tile bits index a
// thread's registers (the other 8 index the thread):
//     tile index of register k of thread t = dep(t, ~rho) | dep(k, rho).
// Diagonal gates, and the controls of controlled gates, only read the bits
// of each amplitude's tile index, so they run in registers in any layout.
// A non-diagonal gate needs its target bit(s) among the register bits; the
// host planner inserts a remap op (registers -> LDS -> registers in a new
// layout chosen to cover the next targets) only when they are not.  A pass
// of 13 gates thus touches LDS a few times instead of 26, and the load and
// store are straight from/to registers with the thread bits on the lowest
// tile qubits (coalesced).  The per-amplitude arithmetic is gate_math.h's,
// in program order, so results are bit-identical to the one-gate kernels.
#include <cstring>

#include "tile_ops.h"

namespace qslhip {

// One op of a register-tile pass (host: plan_reg_ops).
struct RegOp {
    TileOp t;        // t.kind < 0: remap to layout `rho`
    int32_t ia, ib;  // register-bit indices of the non-diagonal target(s)
    uint32_t rho;    // remap: tile-bit mask of the new register bits
    uint32_t pad;
};

namespace {

constexpr int kBlock = 256;

// LDS element order of the remap buffer.  A 128-byte bank sweep holds
// B = 128 / sizeof(cplx<T>) elements (8 FP64 or 16 FP32 amplitudes); the
// slot of element e is its low log2(B) bits XORed with the fold of its
// higher bits, so lanes whose elements differ only above the slot bits
// (register bits among the low tile qubits) still land in distinct banks.
template <typename T>
struct LdsSwizzle {
    static constexpr uint32_t kB = 128 / (uint32_t)sizeof(cplx<T>);
    static constexpr uint32_t kL = kB == 8 ? 3 : 4;
    __device__ __forceinline__ uint32_t operator()(uint32_t e) const
    {
        const uint32_t x = e >> kL;
        return e ^ ((x ^ (x >> kL) ^ (x >> (2 * kL))) & (kB - 1));
    }
};

// Where the tile of one workgroup lives: the tile index j = t | k << 8 of
// the load/store layout (thread t on the 8 lowest tile qubits, register k
// on the top r) maps to the state index gbase | dep(t) | kpart(k).
struct RegTileArgs {
    uint64_t tile_mask;  // the K tile qubits
    uint64_t tlow;       // the 8 lowest of them (thread bits)
    uint64_t kq[4];      // qubit of tile position 8 + i, as a one-bit mask
};

__device__ __forceinline__ uint32_t dep32(uint32_t v, uint32_t mask)
{
    uint32_t r = 0;
    while (mask) {
        uint32_t low = mask & (0u - mask);
        if (v & 1u) r |= low;
        v >>= 1;
        mask ^= low;
    }
    return r;
}

// tile index of every register of this thread in layout rho
template <int R>
__device__ __forceinline__ void layout_index(uint32_t (&jk)[R], uint32_t rho, uint32_t kmask)
{
    const uint32_t tb = dep32(threadIdx.x, kmask & ~rho);
#pragma unroll
    for (int k = 0; k < R; k++) jk[k] = tb | dep32((uint32_t)k, rho);
}

// rotateZ's two diagonal halves without a branch: the |1> half is the |0>
// half with s negated (x + y*(-s) is x - y*s exactly in IEEE arithmetic),
// so the result is bit-identical to gate_math.h's rz_half.
template <typename T>
__device__ __forceinline__ void rz_sel(bool one, cplx<T> &x, T c, T s)
{
    const T ss = one ? -s : s;
    const T r = x.re, i = x.im;
    x.re = r * c + i * ss;
    x.im = i * c - r * ss;
}

// Pair op with the target on register bit I: registers k and k | 2^I; the
// control (a tile bit cpos >= 0, or none) is evaluated per element.
// The register array must stay in registers: every index into it is a
// compile-time constant.  The empty asm markers (distinct immediates per
// instantiation) keep the optimiser from merging the cases of a runtime
// switch over register bits into one body with a variable index, which
// would move the array to scratch memory.
#define QH_CASE_MARK(A, B) asm volatile("; rtile %0 %1" ::"n"(A), "n"(B))

template <typename T, int R, int I, int BODY>
__device__ __forceinline__ void reg_pair(cplx<T> (&v)[R], const uint32_t (&jk)[R], int cpos,
                                         T c, T s)
{
    QH_CASE_MARK(I, BODY);
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (k & (1 << I)) continue;
        cplx<T> a = v[k], b = v[k | (1 << I)];
        pair_update<T>(BODY, a, b, c, s);
        const bool on = cpos < 0 || ((jk[k] >> cpos) & 1u);
        if (on) {
            v[k] = a;
            v[k | (1 << I)] = b;
        }
    }
    QH_CASE_MARK(I, BODY);
}

// Swap-like op on register bits IA (q0) and IB (q1): |q0=1,q1=0> paired
// with |q0=0,q1=1>; fswap also negates |11>.
template <typename T, int R, int IA, int IB, int BODY, bool FSWAP>
__device__ __forceinline__ void reg_swaplike(cplx<T> (&v)[R], T c, T s)
{
    QH_CASE_MARK(IA * 4 + IB, BODY + 64 * FSWAP);
#pragma unroll
    for (int k = 0; k < R; k++) {
        const bool a = k & (1 << IA), b = k & (1 << IB);
        if (a && !b) pair_update<T>(BODY, v[k], v[k ^ (1 << IA) ^ (1 << IB)], c, s);
        if (FSWAP && a && b) one_update<T>(QSLHIP_PAULI_Z, v[k], c, s);
    }
    QH_CASE_MARK(IA * 4 + IB, BODY + 64 * FSWAP);
}

template <typename T, int R, int BODY>
__device__ __forceinline__ void dispatch_pair(cplx<T> (&v)[R], const uint32_t (&jk)[R], int ia,
                                              int cpos, T c, T s)
{
    switch (ia) {
    case 0: reg_pair<T, R, 0, BODY>(v, jk, cpos, c, s); break;
    case 1: reg_pair<T, R, 1, BODY>(v, jk, cpos, c, s); break;
    case 2: if constexpr (R > 4) reg_pair<T, R, 2, BODY>(v, jk, cpos, c, s); break;
    case 3: if constexpr (R > 8) reg_pair<T, R, 3, BODY>(v, jk, cpos, c, s); break;
    default: break;
    }
}

template <typename T, int R, int BODY, bool FSWAP>
__device__ __forceinline__ void dispatch_swap(cplx<T> (&v)[R], int ia, int ib, T c, T s)
{
#define QH_SW(A, B)                                                                   \
    case (A) * 4 + (B):                                                               \
        if constexpr ((1 << (A)) < R && (1 << (B)) < R)                               \
            reg_swaplike<T, R, A, B, BODY, FSWAP>(v, c, s);                           \
        break;
    switch (ia * 4 + ib) {
        QH_SW(0, 1) QH_SW(0, 2) QH_SW(0, 3) QH_SW(1, 0) QH_SW(1, 2) QH_SW(1, 3)
        QH_SW(2, 0) QH_SW(2, 1) QH_SW(2, 3) QH_SW(3, 0) QH_SW(3, 1) QH_SW(3, 2)
    default: break;
    }
#undef QH_SW
}

// Apply one gate to the registers.  bit(q) of an amplitude: a tile bit
// (position >= 0) of its index, or the group's value of a non-tile qubit.
template <typename T, int R>
__device__ __forceinline__ void reg_op(cplx<T> (&v)[R], const uint32_t (&jk)[R],
                                       uint64_t gbase, const RegOp &o)
{
    const TileOp &op = o.t;
    const T c = (T)op.c, s = (T)op.s;
    const int pa = op.pa, pb = op.pb;
    auto bit = [&](int pos, uint32_t q, uint32_t j) -> bool {
        return pos >= 0 ? ((j >> pos) & 1u) : ((gbase >> q) & 1ull);
    };
    switch (op.kind) {
    case QSLHIP_PAULI_X: dispatch_pair<T, R, QSLHIP_PAULI_X>(v, jk, o.ia, -1, c, s); break;
    case QSLHIP_PAULI_Y: dispatch_pair<T, R, QSLHIP_PAULI_Y>(v, jk, o.ia, -1, c, s); break;
    case QSLHIP_HADAMARD: dispatch_pair<T, R, QSLHIP_HADAMARD>(v, jk, o.ia, -1, c, s); break;
    case QSLHIP_ROT_X: dispatch_pair<T, R, QSLHIP_ROT_X>(v, jk, o.ia, -1, c, s); break;
    case QSLHIP_ROT_Y: dispatch_pair<T, R, QSLHIP_ROT_Y>(v, jk, o.ia, -1, c, s); break;
    case QSLHIP_ROT_Z:
#pragma unroll
        for (int k = 0; k < R; k++) rz_sel<T>(bit(pa, op.q0, jk[k]), v[k], c, s);
        break;
    case QSLHIP_PAULI_Z:
    case QSLHIP_PHASE: {
        const int body = op.kind;
#pragma unroll
        for (int k = 0; k < R; k++) {
            cplx<T> x = v[k];
            one_update<T>(body, x, c, s);
            if (bit(pa, op.q0, jk[k])) v[k] = x;
        }
        break;
    }
    case QSLHIP_CZ:
    case QSLHIP_CPHASE: {
        const int body = body_of(op.kind);
#pragma unroll
        for (int k = 0; k < R; k++) {
            cplx<T> x = v[k];
            one_update<T>(body, x, c, s);
            if (bit(pa, op.q0, jk[k]) && bit(pb, op.q1, jk[k])) v[k] = x;
        }
        break;
    }
    case QSLHIP_CROT_Z:
        if (pa < 0 && !((gbase >> op.q0) & 1ull)) break;
#pragma unroll
        for (int k = 0; k < R; k++) {
            cplx<T> x = v[k];
            rz_sel<T>(bit(pb, op.q1, jk[k]), x, c, s);
            if (pa < 0 || ((jk[k] >> pa) & 1u)) v[k] = x;
        }
        break;
    case QSLHIP_CNOT:
    case QSLHIP_CY:
    case QSLHIP_CROT_X:
    case QSLHIP_CROT_Y:
    case QSLHIP_CHADAMARD: {
        if (pa < 0 && !((gbase >> op.q0) & 1ull)) break;  // control off for the group
        const int cpos = pa;
        switch (body_of(op.kind)) {
        case QSLHIP_PAULI_X: dispatch_pair<T, R, QSLHIP_PAULI_X>(v, jk, o.ia, cpos, c, s); break;
        case QSLHIP_PAULI_Y: dispatch_pair<T, R, QSLHIP_PAULI_Y>(v, jk, o.ia, cpos, c, s); break;
        case QSLHIP_ROT_X: dispatch_pair<T, R, QSLHIP_ROT_X>(v, jk, o.ia, cpos, c, s); break;
        case QSLHIP_ROT_Y: dispatch_pair<T, R, QSLHIP_ROT_Y>(v, jk, o.ia, cpos, c, s); break;
        default: dispatch_pair<T, R, QSLHIP_HADAMARD>(v, jk, o.ia, cpos, c, s); break;
        }
        break;
    }
    case QSLHIP_SWAP:
        dispatch_swap<T, R, QSLHIP_PAULI_X, false>(v, o.ia, o.ib, c, s);
        break;
    case QSLHIP_FSWAP:
        dispatch_swap<T, R, QSLHIP_PAULI_X, true>(v, o.ia, o.ib, c, s);
        break;
    case QSLHIP_NP_ROT_X:
        dispatch_swap<T, R, QSLHIP_ROT_X, false>(v, o.ia, o.ib, c, s);
        break;
    case QSLHIP_NP_ROT_Y:
        dispatch_swap<T, R, QSLHIP_ROT_Y, false>(v, o.ia, o.ib, c, s);
        break;
    case QSLHIP_NP_HADAMARD:
        dispatch_swap<T, R, QSLHIP_HADAMARD, false>(v, o.ia, o.ib, c, s);
        break;
    default:
        break;
    }
}

template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_rtile(cplx<T> *__restrict__ st, unsigned n,
                                                  RegTileArgs ta,
                                                  const RegOp *__restrict__ ops, int nops)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem);
    const LdsSwizzle<T> sw;
    constexpr int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    constexpr int K = 8 + r;
    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const uint64_t gbase = deposit((uint64_t)blockIdx.x, full & ~ta.tile_mask);
    const uint64_t tpart = gbase | deposit((uint64_t)threadIdx.x, ta.tlow);
    uint64_t addr[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        uint64_t kp = 0;
#pragma unroll
        for (int i = 0; i < r; i++)
            if ((k >> i) & 1) kp |= ta.kq[i];
        addr[k] = tpart | kp;
    }
    const uint32_t kmask = (1u << K) - 1;
    const uint32_t rho0 = kmask & ~0xFFu;  // registers on the top tile bits
    uint32_t jk[R];
    layout_index<R>(jk, rho0, kmask);
    cplx<T> v[R];
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = st[addr[k]];
    for (int i = 0; i < nops; i++) {
        const RegOp o = ops[i];
        if (o.t.kind < 0) {
#pragma unroll
            for (int k = 0; k < R; k++) lds[sw(jk[k])] = v[k];
            __syncthreads();
            layout_index<R>(jk, o.rho, kmask);
#pragma unroll
            for (int k = 0; k < R; k++) v[k] = lds[sw(jk[k])];
            __syncthreads();
        } else {
            reg_op<T, R>(v, jk, gbase, o);
        }
    }
#pragma unroll
    for (int k = 0; k < R; k++) st[addr[k]] = v[k];
}

// Tile qubits a non-diagonal body moves amplitudes across (tile positions).
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

}  // namespace

// Host planner: the register layouts of one pass.  Starts and ends in the
// load/store layout (registers on the top r tile bits); a remap is inserted
// before a gate whose targets are not register bits, choosing the gate's
// targets plus the targets of the gates that follow, in order.
void plan_reg_ops(const TileOp *ops, uint32_t count, int K, std::vector<RegOp> &out)
{
    const int r = K - 8;
    const uint32_t kmask = (1u << K) - 1;
    const uint32_t rho0 = kmask & ~0xFFu;
    uint32_t cur = rho0;
    auto slot = [](uint32_t rho, int pos) {
        return __builtin_popcount(rho & ((1u << pos) - 1));
    };
    for (uint32_t i = 0; i < count; i++) {
        int tg[2], nt;
        targets_of(ops[i], tg, nt);
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
                if (cur & (1u << p)) rho |= 1u << p;
            for (int p = K - 1; p >= 0 && __builtin_popcount(rho) < r; p--) rho |= 1u << p;
            RegOp m{};
            m.t.kind = -1;
            m.rho = rho;
            out.push_back(m);
            cur = rho;
        }
        RegOp o{};
        o.t = ops[i];
        o.ia = nt > 0 ? slot(cur, tg[0]) : 0;
        o.ib = nt > 1 ? slot(cur, tg[1]) : 0;
        out.push_back(o);
    }
    if (cur != rho0) {
        RegOp m{};
        m.t.kind = -1;
        m.rho = rho0;
        out.push_back(m);
    }
}

template <typename T>
static void launch_rtile_t(qslhip_state *h, void *amps, const TilePass &p, const RegOp *dops,
                           uint32_t nops)
{
    const int K = p.logt;
    RegTileArgs ta{};
    ta.tile_mask = p.tile_mask;
    int pos = 0;
    for (unsigned q = 0; q < 64; q++) {
        if (!((p.tile_mask >> q) & 1ull)) continue;
        if (pos < 8) ta.tlow |= 1ull << q;
        else ta.kq[pos - 8] = 1ull << q;
        pos++;
    }
    QH_REQUIRE(pos == K, QSLHIP_EINVAL, "tile mask does not hold the tile width");
    const size_t shmem = sizeof(cplx<T>) << K;
    QH_REQUIRE(shmem <= 64 * 1024, QSLHIP_EINVAL, "register tile needs more than 64 KiB of LDS");
    const unsigned groups = (unsigned)(1ull << (h->n - K));
    cplx<T> *st = static_cast<cplx<T> *>(amps);
    switch (K) {
    case 10:
        hipLaunchKernelGGL((k_rtile<T, 4>), dim3(groups), dim3(kBlock), shmem, h->stream, st,
                           h->n, ta, dops, (int)nops);
        break;
    case 11:
        hipLaunchKernelGGL((k_rtile<T, 8>), dim3(groups), dim3(kBlock), shmem, h->stream, st,
                           h->n, ta, dops, (int)nops);
        break;
    case 12:
        hipLaunchKernelGGL((k_rtile<T, 16>), dim3(groups), dim3(kBlock), shmem, h->stream, st,
                           h->n, ta, dops, (int)nops);
        break;
    default:
        throw Error(QSLHIP_EINVAL, "register tiles take 10..12 tile qubits");
    }
    QH_CHECK(hipGetLastError());
}

bool reg_tile_ok(const TilePass &p) { return p.logt >= 10 && p.logt <= 12; }

void launch_reg_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                       const std::vector<TileOp> &ops, void *amps)
{
    std::vector<RegOp> rops;
    std::vector<uint32_t> first(passes.size() + 1);
    for (size_t i = 0; i < passes.size(); i++) {
        QH_REQUIRE(reg_tile_ok(passes[i]), QSLHIP_EINVAL, "register tiles take 10..12 tile qubits");
        first[i] = (uint32_t)rops.size();
        plan_reg_ops(ops.data() + passes[i].op_first, passes[i].op_count, passes[i].logt, rops);
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
        const TilePass &p = passes[i];
        QH_REQUIRE(h->n >= (unsigned)p.logt, QSLHIP_EINVAL, "tile wider than the state");
        const uint32_t n_ops = first[i + 1] - first[i];
        if (h->prec == 8) launch_rtile_t<double>(h, amps, p, dops + first[i], n_ops);
        else launch_rtile_t<float>(h, amps, p, dops + first[i], n_ops);
    }
}

}  // namespace qslhip
