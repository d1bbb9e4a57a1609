// gates.hip -- one kernel per gate: the streaming path.
//
// A gate on qubit t pairs amplitude i0 = ins0(f, t) with i0 | 2^t for the
// flat pair id f in [0, 2^(n-1)) -- the mask form of the reference's s/r
// loop nest (src/qubits/default/gates.cpp:35-41).  Consecutive lanes take
// consecutive f, so for t >= 2 (FP64) both streams are contiguous runs of
// whole cache lines; targets below the LDS threshold go through a one-op
// LDS tile pass instead (fused.hip), which loads the tile with fully
// contiguous 16-byte lanes.  Two-qubit gates insert two zero bits
// (f in [0, 2^(n-2))).  Each lane handles kItems pairs, issuing all loads
// before any arithmetic so that every wave keeps 2*kItems dwordx4 (FP64)
// loads in flight.
#include "gate_math.h"

namespace qslhip {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 4;

// Which flat ids a lane takes: lane l of workgroup b takes g = b * 256 + l
// with two zero bits inserted at flat-id positions u < v, and item j sets
// them to j.  With (u, v) = (8, 9) this is f = b * 1024 + j * 256 + l, four
// consecutive runs of 256 ids.  The spread map puts u, v on the positions of
// state-index bits 11 and 12 instead: a lane's items are FP64 amplitudes 32
// and 64 KiB apart, which HBM streams faster than adjacent runs (in place on
// MI355X with streaming accesses: 6.2-6.3 against 5.4 TB/s for targets
// 4-10).  Either way the map has no branch, so all of a lane's loads issue
// back to back.
struct ItemMap {
    unsigned u, v;
};

__device__ __forceinline__ uint64_t item_f(const ItemMap &m, int j)
{
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    return ins0(ins0(g, m.u), m.v) | ((uint64_t)(j & 1) << m.u) | ((uint64_t)(j >> 1) << m.v);
}

// Whether every item of this workgroup is in range.  A state of 2^10 or
// more pairs (quads) fills every workgroup, with either item map; the
// guard-free body then issues all loads back to back.
__device__ __forceinline__ bool full_block(uint64_t nitems)
{
    return (uint64_t)(blockIdx.x + 1) * (kBlock * kItems) <= nitems;
}

// One-qubit gate acting on both amplitudes of each pair.
template <typename T, int BODY, bool S, bool G>
__device__ __forceinline__ void pair1_body(cplx<T> *__restrict__ st, uint64_t npairs,
                                           unsigned t, T c, T s, const ItemMap &m)
{
    const uint64_t tb = 1ull << t;
    cplx<T> a[kItems], b[kItems];
    uint64_t i0[kItems];
    bool in[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const uint64_t f = item_f(m, j);
        i0[j] = ins0(f, t);
        in[j] = !G || f < npairs;
        if (in[j]) {
            a[j] = amp_ld<S>(st + i0[j]);
            b[j] = amp_ld<S>(st + (i0[j] | tb));
        }
    }
    // every load is in flight before the first use (the scheduler would
    // otherwise issue the last item's loads after waiting for the others)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kItems; j++) pair_update<T>(BODY, a[j], b[j], c, s);
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        if (in[j]) {
            amp_st<S>(st + i0[j], a[j]);
            amp_st<S>(st + (i0[j] | tb), b[j]);
        }
    }
}

template <typename T, int BODY, bool S>
__global__ __launch_bounds__(kBlock) void k_pair1(cplx<T> *__restrict__ st,
                                                  uint64_t npairs, unsigned t,
                                                  T c, T s, ItemMap m)
{
    if (full_block(npairs)) pair1_body<T, BODY, S, false>(st, npairs, t, c, s, m);
    else pair1_body<T, BODY, S, true>(st, npairs, t, c, s, m);
}

// One-qubit gate on a target t < 6 (inside one wavefront's 64 lanes): lane l
// of the grid owns amplitude l, so every load and store instruction covers
// 64 consecutive amplitudes, and takes its pair partner, amplitude l ^ 2^t,
// from lane l ^ 2^t by a cross-lane exchange.  Each lane applies the pair's
// update (the reference's arithmetic for both outputs) and keeps its own
// output, so results are those of k_pair1.  (In place on MI355X, FP64,
// streaming: 6.7 TB/s at every target 0-5.)
template <typename T, int BODY, bool S>
__global__ __launch_bounds__(kBlock) void k_pair1_lane(cplx<T> *__restrict__ st, uint64_t namps,
                                                       unsigned t, T c, T s)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= namps) return;
    const cplx<T> own = amp_ld<S>(st + i);
    const int m = 1 << t;
    cplx<T> other;
    other.re = __shfl_xor(own.re, m);
    other.im = __shfl_xor(own.im, m);
    const bool one = (i >> t) & 1ull;
    // component selects (a select of whole structs goes through scratch)
    cplx<T> a, b, out;
    a.re = one ? other.re : own.re;
    a.im = one ? other.im : own.im;
    b.re = one ? own.re : other.re;
    b.im = one ? own.im : other.im;
    pair_update<T>(BODY, a, b, c, s);
    out.re = one ? b.re : a.re;
    out.im = one ? b.im : a.im;
    amp_st<S>(st + i, out);
}

// One-qubit gate touching only the |1> half (pauliZ, phase).
template <typename T, int BODY, bool S, bool G>
__device__ __forceinline__ void one1_body(cplx<T> *__restrict__ st, uint64_t npairs, unsigned t,
                                          T c, T s, const ItemMap &m)
{
    const uint64_t tb = 1ull << t;
    cplx<T> b[kItems];
    uint64_t i1[kItems];
    bool in[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const uint64_t f = item_f(m, j);
        i1[j] = ins0(f, t) | tb;
        in[j] = !G || f < npairs;
        if (in[j]) b[j] = amp_ld<S>(st + i1[j]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kItems; j++) one_update<T>(BODY, b[j], c, s);
#pragma unroll
    for (int j = 0; j < kItems; j++)
        if (in[j]) amp_st<S>(st + i1[j], b[j]);
}

template <typename T, int BODY, bool S>
__global__ __launch_bounds__(kBlock) void k_one1(cplx<T> *__restrict__ st,
                                                 uint64_t npairs, unsigned t,
                                                 T c, T s, ItemMap m)
{
    if (full_block(npairs)) one1_body<T, BODY, S, false>(st, npairs, t, c, s, m);
    else one1_body<T, BODY, S, true>(st, npairs, t, c, s, m);
}

// Two-qubit gates: f in [0, 2^(n-2)) -> base with zeros at lo < hi.
//   CLASS PC_CTRL:     pair (base|abit, base|abit|bbit)     a = ctrl
//   CLASS PC_SWAPLIKE: pair (base|abit, base|bbit)
//   CLASS PC_FSWAP:    swap (base|abit, base|bbit), negate base|abit|bbit
//   CLASS PC_DIAG2:    one (base|abit|bbit)
template <typename T, int CLASS, int BODY, bool S, bool G>
__device__ __forceinline__ void gate2_body(cplx<T> *__restrict__ st, uint64_t nquads,
                                           unsigned lo, unsigned hi, uint64_t abit,
                                           uint64_t bbit, T c, T s, const ItemMap &m)
{
    uint64_t base[kItems];
    bool in[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const uint64_t f = item_f(m, j);
        base[j] = ins0(ins0(f, lo), hi);
        in[j] = !G || f < nquads;
    }
    if constexpr (CLASS == PC_DIAG2) {
        cplx<T> x[kItems];
#pragma unroll
        for (int j = 0; j < kItems; j++)
            if (in[j]) x[j] = amp_ld<S>(st + (base[j] | abit | bbit));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kItems; j++) one_update<T>(BODY, x[j], c, s);
#pragma unroll
        for (int j = 0; j < kItems; j++)
            if (in[j]) amp_st<S>(st + (base[j] | abit | bbit), x[j]);
    } else {
        const uint64_t off_a = abit;
        const uint64_t off_b = (CLASS == PC_CTRL) ? (abit | bbit) : bbit;
        cplx<T> a[kItems], b[kItems], z[kItems];
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            if (in[j]) {
                a[j] = amp_ld<S>(st + (base[j] | off_a));
                b[j] = amp_ld<S>(st + (base[j] | off_b));
                if constexpr (CLASS == PC_FSWAP) z[j] = amp_ld<S>(st + (base[j] | abit | bbit));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            pair_update<T>(BODY, a[j], b[j], c, s);
            if constexpr (CLASS == PC_FSWAP)
                one_update<T>(QSLHIP_PAULI_Z, z[j], c, s);
        }
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            if (in[j]) {
                amp_st<S>(st + (base[j] | off_a), a[j]);
                amp_st<S>(st + (base[j] | off_b), b[j]);
                if constexpr (CLASS == PC_FSWAP) amp_st<S>(st + (base[j] | abit | bbit), z[j]);
            }
        }
    }
}

template <typename T, int CLASS, int BODY, bool S>
__global__ __launch_bounds__(kBlock) void k_gate2(cplx<T> *__restrict__ st,
                                                  uint64_t nquads, unsigned lo,
                                                  unsigned hi, uint64_t abit,
                                                  uint64_t bbit, T c, T s, ItemMap m)
{
    if (full_block(nquads))
        gate2_body<T, CLASS, BODY, S, false>(st, nquads, lo, hi, abit, bbit, c, s, m);
    else
        gate2_body<T, CLASS, BODY, S, true>(st, nquads, lo, hi, abit, bbit, c, s, m);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_basis(cplx<T> *__restrict__ st,
                                                       uint64_t namps,
                                                       uint64_t index)
{
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i < namps; i += (uint64_t)gridDim.x * kBlock) {
        cplx<T> v;
        v.re = (i == index) ? (T)1 : (T)0;
        v.im = (T)0;
        st[i] = v;
    }
}

// The spread item map for a flat id space of 2^nflat ids that skips the
// state-index bits in `skip` (the gate's qubits), or the default one when
// a gate qubit is 11 or 12 or the space is too small.
ItemMap item_map(unsigned nflat, uint64_t skip)
{
    ItemMap m{8, 9};  // four consecutive runs of 256 ids
    const unsigned a = 11, b = 12;
    if ((skip >> a) & 1ull || (skip >> b) & 1ull || nflat < 13) return m;
    const unsigned u = a - (unsigned)__builtin_popcountll(skip & ((1ull << a) - 1));
    const unsigned v = b - (unsigned)__builtin_popcountll(skip & ((1ull << b) - 1));
    if (v < nflat) m = ItemMap{u, v};
    return m;
}

template <typename T, bool S>
void dispatch(qslhip_state *h, const GateDesc &g, void *amps)
{
    cplx<T> *st = static_cast<cplx<T> *>(amps);
    const T c = (T)g.c, s = (T)g.s;
    const unsigned n = h->n;
    const unsigned per = kBlock * kItems;
    if (is_gate1(g.kind)) {
        uint64_t np = 1ull << (n - 1);
        dim3 grid(grid_for(np, per)), blk(kBlock);
        const ItemMap m = item_map(n - 1, 1ull << g.q0);
        switch (g.kind) {
#define P1(K)                                                                  \
    case K:                                                                    \
        hipLaunchKernelGGL((k_pair1<T, body_of(K), S>), grid, blk, 0, h->stream, \
                           st, np, g.q0, c, s, m);                             \
        break;
            P1(QSLHIP_PAULI_X) P1(QSLHIP_PAULI_Y) P1(QSLHIP_HADAMARD)
            P1(QSLHIP_ROT_X) P1(QSLHIP_ROT_Y) P1(QSLHIP_ROT_Z)
#undef P1
        case QSLHIP_PAULI_Z:
            hipLaunchKernelGGL((k_one1<T, QSLHIP_PAULI_Z, S>), grid, blk, 0,
                               h->stream, st, np, g.q0, c, s, m);
            break;
        case QSLHIP_PHASE:
            hipLaunchKernelGGL((k_one1<T, QSLHIP_PHASE, S>), grid, blk, 0,
                               h->stream, st, np, g.q0, c, s, m);
            break;
        default:
            throw Error(QSLHIP_EINVAL, "unknown one-qubit gate kind");
        }
    } else {
        uint64_t nq = 1ull << (n - 2);
        unsigned lo = g.q0 < g.q1 ? g.q0 : g.q1, hi = g.q0 < g.q1 ? g.q1 : g.q0;
        uint64_t abit = 1ull << g.q0, bbit = 1ull << g.q1;
        dim3 grid(grid_for(nq, per)), blk(kBlock);
        const ItemMap m = item_map(n - 2, abit | bbit);
        switch (g.kind) {
#define P2(K)                                                                  \
    case K:                                                                    \
        hipLaunchKernelGGL((k_gate2<T, pair_class(K), body_of(K), S>), grid, blk, \
                           0, h->stream, st, nq, lo, hi, abit, bbit, c, s, m); \
        break;
            P2(QSLHIP_CNOT) P2(QSLHIP_CY) P2(QSLHIP_CZ) P2(QSLHIP_CROT_X)
            P2(QSLHIP_CROT_Y) P2(QSLHIP_CROT_Z) P2(QSLHIP_CPHASE)
            P2(QSLHIP_CHADAMARD) P2(QSLHIP_SWAP) P2(QSLHIP_FSWAP)
            P2(QSLHIP_NP_ROT_X) P2(QSLHIP_NP_ROT_Y) P2(QSLHIP_NP_HADAMARD)
#undef P2
        default:
            throw Error(QSLHIP_EINVAL, "unknown two-qubit gate kind");
        }
    }
    QH_CHECK(hipGetLastError());
}

}  // namespace

template <typename T, bool S>
void dispatch_lane(qslhip_state *h, const GateDesc &g, void *amps)
{
    cplx<T> *st = static_cast<cplx<T> *>(amps);
    const T c = (T)g.c, s = (T)g.s;
    dim3 grid(grid_for(h->namps, kBlock)), blk(kBlock);
    switch (g.kind) {
#define PL(K)                                                                      \
    case K:                                                                        \
        hipLaunchKernelGGL((k_pair1_lane<T, body_of(K), S>), grid, blk, 0, h->stream, \
                           st, h->namps, g.q0, c, s);                              \
        break;
        PL(QSLHIP_PAULI_X) PL(QSLHIP_PAULI_Y) PL(QSLHIP_HADAMARD)
        PL(QSLHIP_ROT_X) PL(QSLHIP_ROT_Y) PL(QSLHIP_ROT_Z)
#undef PL
    default:
        throw Error(QSLHIP_EINVAL, "not a pair-updating one-qubit gate");
    }
    QH_CHECK(hipGetLastError());
}

bool lane_kind(int kind)
{
    return kind == QSLHIP_PAULI_X || kind == QSLHIP_PAULI_Y || kind == QSLHIP_HADAMARD ||
           kind == QSLHIP_ROT_X || kind == QSLHIP_ROT_Y || kind == QSLHIP_ROT_Z;
}

// One-qubit gates on targets 0..5 exchange pair partners across lanes;
// other low targets go through a one-op LDS tile; the rest stream.
void launch_gate(qslhip_state *h, const GateDesc &g, void *amps)
{
    if (!amps) amps = h->amps;
    const Tuning &tu = tuning();
    if (lane_kind(g.kind) && g.q0 < 6 && h->n >= 6 && tu.lane_gates) {
        const bool stream = (int)h->n >= tu.stream_min_qubits;
        if (h->prec == 8) {
            if (stream) dispatch_lane<double, true>(h, g, amps);
            else dispatch_lane<double, false>(h, g, amps);
        } else {
            if (stream) dispatch_lane<float, true>(h, g, amps);
            else dispatch_lane<float, false>(h, g, amps);
        }
        return;
    }
    const int lim = h->prec == 8 ? tu.lds_max_target_f64 : tu.lds_max_target_f32;
    const int logt = h->prec == 8 ? tu.logt_f64 : tu.logt_f32;
    unsigned top = g.q0;
    if (is_gate2(g.kind)) top = g.q0 > g.q1 ? g.q0 : g.q1;
    if ((int)top <= lim && (int)h->n >= logt) {
        std::vector<GateDesc> one{g};
        std::vector<TilePass> passes;
        std::vector<TileOp> ops;
        plan_tile_passes(h->n, h->prec, one, logt, logt, passes, ops);
        launch_tile_passes(h, passes, ops, amps);
        return;
    }
    if (h->n < 2 && is_gate2(g.kind)) throw Error(QSLHIP_EINVAL, "need 2 qubits");
    // Streaming accesses pay when each load instruction covers whole 128-byte
    // lines: the lowest gate qubit must leave runs of at least 128 bytes of
    // touched amplitudes (with a control or target on qubit 0, a lane's
    // loads are 2^lo amplitudes apart and a cached access fetches each line
    // once for both of its halves).
    unsigned low = g.q0;
    if (is_gate2(g.kind)) low = g.q0 < g.q1 ? g.q0 : g.q1;
    const size_t amp = h->prec == 8 ? 16 : 8;
    const bool stream = (int)h->n >= tu.stream_min_qubits && (amp << low) >= 128;
    if (h->prec == 8) {
        if (stream) dispatch<double, true>(h, g, amps);
        else dispatch<double, false>(h, g, amps);
    } else {
        if (stream) dispatch<float, true>(h, g, amps);
        else dispatch<float, false>(h, g, amps);
    }
}

void launch_fill_basis(qslhip_state *h, uint64_t index)
{
    unsigned grid = grid_for(h->namps, kBlock);
    if (grid > 65536u * 4) grid = 65536u * 4;
    if (h->prec == 8)
        hipLaunchKernelGGL(k_fill_basis<double>, dim3(grid), dim3(kBlock), 0,
                           h->stream, static_cast<cplx<double> *>(h->amps),
                           h->namps, index);
    else
        hipLaunchKernelGGL(k_fill_basis<float>, dim3(grid), dim3(kBlock), 0,
                           h->stream, static_cast<cplx<float> *>(h->amps),
                           h->namps, index);
    QH_CHECK(hipGetLastError());
}

}  // namespace qslhip
