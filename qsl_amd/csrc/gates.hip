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

// Which flat ids a lane takes.  By default item j of lane l in workgroup b
// is f = b * 1024 + j * 256 + l.  With `spread`, the lane's four items
// instead differ in flat-id bits u < v, the positions of state-index bits
// 11 and 12: FP64 amplitudes 32 and 64 KiB apart, which HBM streams 2-9%
// faster than 4 KiB steps (measured in place on MI355X, every target but
// 11 and 12, whose pairs already sit that far apart).
struct ItemMap {
    unsigned u, v;
    int spread;
};

__device__ __forceinline__ uint64_t item_f(const ItemMap &m, int j)
{
    if (!m.spread) return (uint64_t)blockIdx.x * (kBlock * kItems) + (uint64_t)j * kBlock + threadIdx.x;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;  // id without bits u, v
    return ins0(ins0(g, m.u), m.v) | ((uint64_t)(j & 1) << m.u) | ((uint64_t)(j >> 1) << m.v);
}

// One-qubit gate acting on both amplitudes of each pair.
template <typename T, int BODY>
__global__ __launch_bounds__(kBlock) void k_pair1(cplx<T> *__restrict__ st,
                                                  uint64_t npairs, unsigned t,
                                                  T c, T s, ItemMap m)
{
    const uint64_t tb = 1ull << t;
    cplx<T> a[kItems], b[kItems];
    uint64_t i0[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint64_t f = item_f(m, j);
        i0[j] = ins0(f, t);
        if (f < npairs) {
            a[j] = st[i0[j]];
            b[j] = st[i0[j] | tb];
        }
    }
#pragma unroll
    for (int j = 0; j < kItems; j++) pair_update<T>(BODY, a[j], b[j], c, s);
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint64_t f = item_f(m, j);
        if (f < npairs) {
            st[i0[j]] = a[j];
            st[i0[j] | tb] = b[j];
        }
    }
}

// One-qubit gate touching only the |1> half (pauliZ, phase).
template <typename T, int BODY>
__global__ __launch_bounds__(kBlock) void k_one1(cplx<T> *__restrict__ st,
                                                 uint64_t npairs, unsigned t,
                                                 T c, T s, ItemMap m)
{
    const uint64_t tb = 1ull << t;
    cplx<T> b[kItems];
    uint64_t i1[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint64_t f = item_f(m, j);
        i1[j] = ins0(f, t) | tb;
        if (f < npairs) b[j] = st[i1[j]];
    }
#pragma unroll
    for (int j = 0; j < kItems; j++) one_update<T>(BODY, b[j], c, s);
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint64_t f = item_f(m, j);
        if (f < npairs) st[i1[j]] = b[j];
    }
}

// Two-qubit gates: f in [0, 2^(n-2)) -> base with zeros at lo < hi.
//   CLASS PC_CTRL:     pair (base|abit, base|abit|bbit)     a = ctrl
//   CLASS PC_SWAPLIKE: pair (base|abit, base|bbit)
//   CLASS PC_FSWAP:    swap (base|abit, base|bbit), negate base|abit|bbit
//   CLASS PC_DIAG2:    one (base|abit|bbit)
template <typename T, int CLASS, int BODY>
__global__ __launch_bounds__(kBlock) void k_gate2(cplx<T> *__restrict__ st,
                                                  uint64_t nquads, unsigned lo,
                                                  unsigned hi, uint64_t abit,
                                                  uint64_t bbit, T c, T s, ItemMap m)
{
    uint64_t fj[kItems], base[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        fj[j] = item_f(m, j);
        base[j] = ins0(ins0(fj[j], lo), hi);
    }
    if constexpr (CLASS == PC_DIAG2) {
        cplx<T> x[kItems];
#pragma unroll
        for (int j = 0; j < kItems; j++)
            if (fj[j] < nquads) x[j] = st[base[j] | abit | bbit];
#pragma unroll
        for (int j = 0; j < kItems; j++) one_update<T>(BODY, x[j], c, s);
#pragma unroll
        for (int j = 0; j < kItems; j++)
            if (fj[j] < nquads) st[base[j] | abit | bbit] = x[j];
    } else {
        const uint64_t off_a = abit;
        const uint64_t off_b = (CLASS == PC_CTRL) ? (abit | bbit) : bbit;
        cplx<T> a[kItems], b[kItems], z[kItems];
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            if (fj[j] < nquads) {
                a[j] = st[base[j] | off_a];
                b[j] = st[base[j] | off_b];
                if constexpr (CLASS == PC_FSWAP) z[j] = st[base[j] | abit | bbit];
            }
        }
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            pair_update<T>(BODY, a[j], b[j], c, s);
            if constexpr (CLASS == PC_FSWAP)
                one_update<T>(QSLHIP_PAULI_Z, z[j], c, s);
        }
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            if (fj[j] < nquads) {
                st[base[j] | off_a] = a[j];
                st[base[j] | off_b] = b[j];
                if constexpr (CLASS == PC_FSWAP) st[base[j] | abit | bbit] = z[j];
            }
        }
    }
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
    ItemMap m{0, 0, 0};
    const unsigned a = 11, b = 12;
    if ((skip >> a) & 1ull || (skip >> b) & 1ull || nflat < 13) return m;
    m.u = a - (unsigned)__builtin_popcountll(skip & ((1ull << a) - 1));
    m.v = b - (unsigned)__builtin_popcountll(skip & ((1ull << b) - 1));
    m.spread = m.v < nflat;
    return m;
}

template <typename T>
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
        hipLaunchKernelGGL((k_pair1<T, body_of(K)>), grid, blk, 0, h->stream,  \
                           st, np, g.q0, c, s, m);                             \
        break;
            P1(QSLHIP_PAULI_X) P1(QSLHIP_PAULI_Y) P1(QSLHIP_HADAMARD)
            P1(QSLHIP_ROT_X) P1(QSLHIP_ROT_Y) P1(QSLHIP_ROT_Z)
#undef P1
        case QSLHIP_PAULI_Z:
            hipLaunchKernelGGL((k_one1<T, QSLHIP_PAULI_Z>), grid, blk, 0,
                               h->stream, st, np, g.q0, c, s, m);
            break;
        case QSLHIP_PHASE:
            hipLaunchKernelGGL((k_one1<T, QSLHIP_PHASE>), grid, blk, 0,
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
        hipLaunchKernelGGL((k_gate2<T, pair_class(K), body_of(K)>), grid, blk, \
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

// Low targets go through a one-op LDS tile; the rest stream.
void launch_gate(qslhip_state *h, const GateDesc &g, void *amps)
{
    if (!amps) amps = h->amps;
    const Tuning &tu = tuning();
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
    if (h->prec == 8) dispatch<double>(h, g, amps);
    else dispatch<float>(h, g, amps);
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
