// np_compact.hip -- number-preserving backend on the compact colex layout.
//
// The reference's NP backend keeps a full 2^n vector and walks fixed-weight
// lookup lists (src/qubits/np/np.cpp:168-181).  Here the device holds only
// the C(n,k) amplitudes of weight k, at index rank(x) where
//     rank(x) = sum over set bits p_1 < ... < p_k of C(p_i, i),
// the position of x in the reference's own enumeration order (Gosper's next
// visits weight-k strings in increasing numeric order, which is colex
// order; src/utils/misc.cpp:44-90).
//
// Two-level rank: x = (hi << L) | lo with L = min(n, max(16, n - 16)) low
// bits.  Strings sharing `hi` are contiguous ("segment hi"), ordered by lo
// within it, so
//     rank(x) = seg_off[hi] + rank_lo[lo],   unrank = table lookups,
// with rank_lo / lo_list (the L-bit strings of each weight, in increasing
// order) and seg_off (prefix sums of C(L, k - popcount(hi))) built exactly
// on the host in u64.  A workgroup handles up to kChunk consecutive ranks of
// one segment (the block map), so every element's string costs one
// coalesced table load and the amplitude loads are contiguous.  Gates whose
// qubits are all in `hi` are constant per segment: whole workgroups skip,
// and a swap-like pair in `hi` pairs element i of segment hi with element i
// of segment hi', two contiguous streams.
#include <cstring>

#include "gate_math.h"

namespace qslhip {

struct NPBlk {
    uint32_t hi;  // segment (high bits of the strings)
    uint32_t i0;  // first index inside the segment
};

// A block of k_np_lowbonds: the ranks [first, end) of segment hi, whole
// runs of strings that share their bits from bit 12 up.
struct NPRunBlk {
    uint32_t hi, first, end;
};

struct NPTables {
    unsigned L = 0, hib = 0;
    uint64_t *seg_off = nullptr;      // [2^hib + 1] rank of each segment's start
    uint32_t *lo_list = nullptr;      // [2^L] strings grouped by weight, ascending
    uint32_t *rank_lo = nullptr;      // [2^L] index of a string inside its weight
    uint32_t *class_start = nullptr;  // [L + 2] offset of each weight in lo_list
    uint64_t *binom = nullptr;        // [64 * 64] C(q, m) at q * 64 + m
    NPBlk *blk = nullptr;             // block map
    uint32_t nblk = 0;
    NPRunBlk *runblk = nullptr;       // whole runs below bit 12 per block
    uint32_t nrunblk = 0;             // (k_np_lowbonds; built on first use)
    std::vector<uint64_t> h_seg_off;  // host copies
    std::vector<uint32_t> h_lo_list, h_rank_lo, h_class_start;
};

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 4;
constexpr int kChunk = kBlock * kItems;

uint64_t choose_host(unsigned n, unsigned k)
{
    if (k > n) return 0;
    uint64_t a = 1;
    for (unsigned i = 1; i <= k; i++) {
        a *= (uint64_t)(n - k + i);
        a /= i;
    }
    return a;
}

// Device view passed by value to every kernel.
struct NPDev {
    const uint64_t *seg_off;
    const uint32_t *lo_list;
    const uint32_t *rank_lo;
    const uint32_t *class_start;
    const NPBlk *blk;
    const uint64_t *binom;
    unsigned L, k;
    uint64_t lomask;
};

NPDev dev_view(const qslhip_state *h)
{
    const NPTables *t = h->npt;
    NPDev d;
    d.seg_off = t->seg_off;
    d.lo_list = t->lo_list;
    d.rank_lo = t->rank_lo;
    d.class_start = t->class_start;
    d.blk = t->blk;
    d.binom = t->binom;
    d.L = t->L;
    d.k = h->k;
    d.lomask = (1ull << t->L) - 1;
    return d;
}

__device__ __forceinline__ uint64_t np_rank(const NPDev &d, uint64_t x)
{
    return d.seg_off[x >> d.L] + d.rank_lo[x & d.lomask];
}

// The rank of x with the one of the adjacent pair (q0, q1) moved from q0 to
// q1 (x has bit q0 set and q1 clear), from the rank r of x.  With q the
// lower of the two and j - 1 ones of x below q, the moved one is the j-th
// one; its colex term C(q, j) becomes C(q + 1, j), a change of C(q, j - 1)
// (Pascal's rule): up by it when q1 = q0 + 1, down when q1 = q0 - 1.  No
// rank table is read.
__device__ __forceinline__ uint64_t adj_partner(const NPDev &d, uint64_t x, uint64_t r,
                                                unsigned q0, unsigned q1)
{
    const unsigned q = q0 < q1 ? q0 : q1;
    const uint64_t m = d.binom[q * 64 + __popcll(x & ((1ull << q) - 1))];
    return q1 > q0 ? r + m : r - m;
}

// Per-block segment context.
struct Seg {
    uint32_t hi, i0, size, cstart;
    uint64_t off;
};

__device__ __forceinline__ Seg seg_of(const NPDev &d)
{
    Seg s;
    NPBlk b = d.blk[blockIdx.x];
    s.hi = b.hi;
    s.i0 = b.i0;
    unsigned w = d.k - __popc(b.hi);
    s.cstart = d.class_start[w];
    s.size = d.class_start[w + 1] - s.cstart;
    s.off = d.seg_off[b.hi];
    return s;
}

// Element-wise gate classes on the compact vector (the same per-amplitude
// arithmetic as the reference NP gates, src/qubits/np/gates.cpp):
//   PC_ONE   phase / pauliZ on t       (np/gates.cpp:33-81)
//   PC_PAIR1 rotateZ on t, per-element halves (np/gates.cpp:83-113)
//   PC_DIAG2 controlZ / controlPhase   (np/gates.cpp:118-150, 195-233)
//   PC_CTRL  controlRotateZ            (np/gates.cpp:153-192)
//   PC_SWAPLIKE / PC_FSWAP  owner x (bit q0 = 1, bit q1 = 0) updates itself
//            and its partner x ^ bit q0 ^ bit q1 (np/gates.cpp:236-415);
//            fswap also negates the |11> strings (swap then controlZ).
template <typename T, int CLASS, bool S>
__global__ __launch_bounds__(kBlock) void k_np_gate(cplx<T> *__restrict__ amp, NPDev d,
                                                    int body, unsigned q0, unsigned q1,
                                                    T c, T s)
{
    const Seg g = seg_of(d);
    const uint64_t b0 = 1ull << q0, b1 = 1ull << q1;
    const uint64_t hix = (uint64_t)g.hi << d.L;
    const bool adjacent = q0 + 1 == q1 || q1 + 1 == q0;
    // Whole-segment decisions when the qubits are in the high part.
    if (CLASS == PC_ONE && q0 >= d.L && !(hix & b0)) return;
    if (CLASS == PC_DIAG2 && ((q0 >= d.L && !(hix & b0)) || (q1 >= d.L && !(hix & b1))))
        return;
    if (CLASS == PC_CTRL && q0 >= d.L && !(hix & b0)) return;
    if ((CLASS == PC_SWAPLIKE || CLASS == PC_FSWAP) && q0 >= d.L && q1 >= d.L) {
        // both qubits in the high part: segment hi pairs element i with
        // element i of segment hi ^ b0 ^ b1 (same weight, same low strings),
        // two contiguous streams with no per-element table reads
        const bool own = (hix & b0) && !(hix & b1);
        if (CLASS == PC_FSWAP && (hix & b0) && (hix & b1)) {
#pragma unroll
            for (int e = 0; e < kItems; e++) {
                const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
                if (i >= g.size || i >= g.i0 + kChunk) continue;
                cplx<T> a = amp_ld<S>(amp + g.off + i);
                one_update<T>(QSLHIP_PAULI_Z, a, c, s);
                amp_st<S>(amp + g.off + i, a);
            }
            return;
        }
        if (!own) return;
        const uint64_t offp = d.seg_off[(hix ^ b0 ^ b1) >> d.L];
#pragma unroll
        for (int e = 0; e < kItems; e++) {
            const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
            if (i >= g.size || i >= g.i0 + kChunk) continue;
            cplx<T> a = amp_ld<S>(amp + g.off + i), b = amp_ld<S>(amp + offp + i);
            pair_update<T>(body, a, b, c, s);
            amp_st<S>(amp + g.off + i, a);
            amp_st<S>(amp + offp + i, b);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < kItems; e++) {
        uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        if (i >= g.size || i >= g.i0 + kChunk) continue;
        uint64_t x = hix | d.lo_list[g.cstart + i];
        uint64_t r = g.off + i;
        if constexpr (CLASS == PC_ONE) {
            if (x & b0) {
                cplx<T> a = amp_ld<S>(amp + r);
                one_update<T>(body, a, c, s);
                amp_st<S>(amp + r, a);
            }
        } else if constexpr (CLASS == PC_PAIR1) {
            cplx<T> a = amp_ld<S>(amp + r);
            rz_half<T>((x & b0) != 0, a, c, s);
            amp_st<S>(amp + r, a);
        } else if constexpr (CLASS == PC_DIAG2) {
            if ((x & b0) && (x & b1)) {
                cplx<T> a = amp_ld<S>(amp + r);
                one_update<T>(body, a, c, s);
                amp_st<S>(amp + r, a);
            }
        } else if constexpr (CLASS == PC_CTRL) {
            if (x & b0) {
                cplx<T> a = amp_ld<S>(amp + r);
                rz_half<T>((x & b1) != 0, a, c, s);
                amp_st<S>(amp + r, a);
            }
        } else {  // PC_SWAPLIKE, PC_FSWAP
            if ((x & b0) && !(x & b1)) {
                uint64_t rp = adjacent ? adj_partner(d, x, r, q0, q1) : np_rank(d, x ^ b0 ^ b1);
                cplx<T> a = amp_ld<S>(amp + r), b = amp_ld<S>(amp + rp);
                pair_update<T>(body, a, b, c, s);
                amp_st<S>(amp + r, a);
                amp_st<S>(amp + rp, b);
            } else if (CLASS == PC_FSWAP && (x & b0) && (x & b1)) {
                cplx<T> a = amp_ld<S>(amp + r);
                one_update<T>(QSLHIP_PAULI_Z, a, c, s);
                amp_st<S>(amp + r, a);
            }
        }
    }
}

// A swap-like gate on an adjacent pair inside the low part of the strings
// (q_lo + 1 < L), one element per lane.  Within a segment the strings are in
// numeric order, so the two members of a pair (bits (q+1, q) = 01 and 10,
// the same other bits) sit C(q, ones below q) ranks apart; for low q that is
// a few ranks, inside one wavefront.  Every lane loads its own amplitude
// (contiguous loads across the wave), takes its pair member's amplitude from
// that member's lane by a cross-lane exchange (or from memory when the
// member is in another wavefront), applies the pair update of the gate
// (k_np_gate's arithmetic for both members) and stores its own result; a
// pair split over two wavefronts is done by its owner alone.  Strings with
// bits 00 or 11 at (q+1, q) are neither read nor written.
template <typename T, bool S, bool ALL>
__global__ __launch_bounds__(kBlock) void k_np_adj(cplx<T> *__restrict__ amp, NPDev d, int body,
                                                   unsigned q0, unsigned q1, T c, T s)
{
    __shared__ uint64_t crow[64];  // C(q, m)
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
    const unsigned q = q0 < q1 ? q0 : q1;
    if (threadIdx.x < 64) crow[threadIdx.x] = d.binom[q * 64 + threadIdx.x];
    __syncthreads();
    const uint64_t bq = 1ull << q, bq1 = bq << 1, b0 = 1ull << q0;
    const uint64_t lowm = bq - 1;
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int e = 0; e < kItems; e++) {
        const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        const bool in = i < g.size && i < g.i0 + kChunk;
        uint64_t x = 0;
        cplx<T> own;
        own.re = own.im = (T)0;
        // ALL: every element of the chunk is loaded and stored (whole lines,
        // no wait for the string before the amplitude load)
        if (ALL && in) own = amp_ld<S>(amp + g.off + i);
        if (in) x = hix | d.lo_list[g.cstart + i];
        // member of a pair: exactly one of bits q, q + 1 set
        const bool mem = in && (((x & bq) != 0) != ((x & bq1) != 0));
        if (!ALL && mem) own = amp_ld<S>(amp + g.off + i);
        // offset to the other member: + m from the 01 string, - m from 10
        const uint64_t m = mem ? crow[__popcll(x & lowm)] : 0;
        const bool up = (x & bq) != 0;  // holds the 01 string
        const int64_t delta = up ? (int64_t)m : -(int64_t)m;
        const int64_t src = (int64_t)lane + delta;
        // the other member is in this wavefront's lanes (same item, same
        // block) when its segment index lies in the chunk and its lane in
        // [0, 64)
        const int64_t io = (int64_t)i + delta;
        const bool local = mem && src >= 0 && src < 64 && io >= (int64_t)g.i0 &&
                           io < (int64_t)(g.i0 + kChunk) && io < (int64_t)g.size;
        const int srcl = local ? (int)src : (int)lane;
        cplx<T> other;
        other.re = __shfl(own.re, srcl);
        other.im = __shfl(own.im, srcl);
        // owner: bit q0 set (and q1 clear); the pair update takes it first
        const bool owner = mem && (x & b0) != 0;
        // A pair split over two wavefronts is the owner's alone: it reads
        // and writes both members, and the other member's lane leaves it
        // (its lane could otherwise read the owner after the owner's store).
        if (mem && !local && owner) other = amp[g.off + (uint64_t)io];
        if (mem && (local || owner)) {
            cplx<T> a, b;
            a.re = owner ? own.re : other.re;
            a.im = owner ? own.im : other.im;
            b.re = owner ? other.re : own.re;
            b.im = owner ? other.im : own.im;
            pair_update<T>(body, a, b, c, s);
            cplx<T> out;
            out.re = owner ? a.re : b.re;
            out.im = owner ? a.im : b.im;
            own = out;
            if (!ALL) amp_st<S>(amp + g.off + i, out);
            if (!local) amp[g.off + (uint64_t)io] = b;
        }
        // the untouched strings are written back unchanged; a split pair's
        // non-owner member is the owner's to write
        if (ALL && in && !(mem && !local && !owner)) amp_st<S>(amp + g.off + i, own);
    }
}

// One diagonal gate of a fused run (k_np_diag).
struct NPDiagOp {
    int32_t cls, body;  // PC_ONE, PC_PAIR1, PC_DIAG2 or PC_CTRL; gate body
    uint32_t q0, q1;
    double c, s;        // coefficients, already in the handle's precision
};

// A run of diagonal gates in one pass over the compact vector: each
// amplitude is loaded once and takes the run's factors in program order,
// each with k_np_gate's arithmetic for its class, so the result is
// bit-identical to one launch per gate.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_diag(cplx<T> *__restrict__ amp, NPDev d,
                                                    const NPDiagOp *__restrict__ ops, int nops)
{
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
    // all items loaded first, then the ops outer and the items inner: each
    // op is decoded once per thread, and each amplitude still takes the
    // run's factors in program order
    uint64_t x[kItems];
    cplx<T> a[kItems];
    bool in[kItems];
#pragma unroll
    for (int e = 0; e < kItems; e++) {
        const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        in[e] = i < g.size && i < g.i0 + kChunk;
        x[e] = 0;
        a[e].re = a[e].im = (T)0;
        if (in[e]) {
            x[e] = hix | d.lo_list[g.cstart + i];
            a[e] = amp[g.off + i];
        }
    }
    for (int j = 0; j < nops; j++) {
        const NPDiagOp o = ops[j];
        const T c = (T)o.c, s = (T)o.s;
        switch (o.cls) {
        case PC_ONE:
            if (o.q0 >= d.L) {  // a bit of the segment: the same for every item
                if ((hix >> o.q0) & 1ull) {
#pragma unroll
                    for (int e = 0; e < kItems; e++) one_update<T>(o.body, a[e], c, s);
                }
                break;
            }
#pragma unroll
            for (int e = 0; e < kItems; e++)
                if ((x[e] >> o.q0) & 1ull) one_update<T>(o.body, a[e], c, s);
            break;
        case PC_PAIR1:
#pragma unroll
            for (int e = 0; e < kItems; e++) rz_half<T>((x[e] >> o.q0) & 1ull, a[e], c, s);
            break;
        case PC_DIAG2:
#pragma unroll
            for (int e = 0; e < kItems; e++)
                if (((x[e] >> o.q0) & 1ull) && ((x[e] >> o.q1) & 1ull))
                    one_update<T>(o.body, a[e], c, s);
            break;
        default:  // PC_CTRL
#pragma unroll
            for (int e = 0; e < kItems; e++)
                if ((x[e] >> o.q0) & 1ull) rz_half<T>((x[e] >> o.q1) & 1ull, a[e], c, s);
            break;
        }
    }
#pragma unroll
    for (int e = 0; e < kItems; e++)
        if (in[e]) amp[g.off + g.i0 + e * kBlock + threadIdx.x] = a[e];
}

// A run of up to kPhaseMax phase gates (np/gates.cpp:33-81) in one pass.
// A phase gate changes only the amplitudes whose string has its qubit set,
// so each amplitude takes just the run indices acting on it: per string the
// mask of them is one table word per byte of the low part (lut[b][v] holds
// the gates on qubit 8b + i for each bit i set in v) OR'd with the gates on
// qubits of the segment that are 1 in it (the same for the whole
// workgroup), and its set bits are taken lowest first, i.e. in program
// order, each with one_update's arithmetic: bit-identical to one launch per
// gate.  For a run over distinct qubits the strings of one segment have
// equal counts, so the lanes iterate in lockstep.
constexpr int kPhaseMax = 32;
constexpr int kPhaseBytes = 3;  // low parts of up to 24 bits
struct NPPhaseRun {
    int nops;                // gates in the run, at most kPhaseMax
    int nbytes;              // bytes of the low part the table covers
    uint32_t lowmask;        // run indices of the gates on qubits below L
    uint32_t q[kPhaseMax];   // qubit of each gate
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_phases(cplx<T> *__restrict__ amp, NPDev d,
                                                      const uint32_t *__restrict__ lut,
                                                      const double *__restrict__ cs,
                                                      NPPhaseRun run)
{
    __shared__ uint32_t s_lut[kPhaseBytes * 256];
    __shared__ cplx<T> s_cs[kPhaseMax];
    for (int i = threadIdx.x; i < run.nbytes * 256; i += kBlock) s_lut[i] = lut[i];
    if (threadIdx.x < (unsigned)run.nops) {
        s_cs[threadIdx.x].re = (T)cs[2 * threadIdx.x];
        s_cs[threadIdx.x].im = (T)cs[2 * threadIdx.x + 1];
    }
    const Seg g = seg_of(d);
    uint32_t lo[kItems];
    cplx<T> a[kItems];
    bool in[kItems];
#pragma unroll
    for (int e = 0; e < kItems; e++) {
        const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        in[e] = i < g.size && i < g.i0 + kChunk;
        lo[e] = 0;
        a[e].re = a[e].im = (T)0;
        if (in[e]) {
            lo[e] = d.lo_list[g.cstart + i];
            a[e] = amp[g.off + i];
        }
    }
    __syncthreads();
    uint32_t m[kItems];  // the low gates acting on each string
#pragma unroll
    for (int e = 0; e < kItems; e++) {
        uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < kPhaseBytes; b++)
            if (b < run.nbytes) w |= s_lut[b * 256 + ((lo[e] >> (8 * b)) & 255u)];
        m[e] = w;
    }
    // the gates on qubits of the segment that are 1 in it act on every
    // string of the workgroup
    uint32_t mh = 0;
    for (int j = 0; j < run.nops; j++)
        if (!((run.lowmask >> j) & 1u) && ((g.hi >> (run.q[j] - d.L)) & 1u)) mh |= 1u << j;
#pragma unroll
    for (int e = 0; e < kItems; e++) m[e] = in[e] ? (m[e] | mh) : 0u;
    // the usual case (a run over distinct qubits): every string of the
    // wavefront has the same number of acting gates, and the loop needs no
    // per-lane selects
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(__popc(m[0]));
    bool same = true;
#pragma unroll
    for (int e = 0; e < kItems; e++) same = same && (uint32_t)__popc(m[e]) == cnt;
    if (__all(same)) {
        for (uint32_t it = 0; it < cnt; it++) {
            cplx<T> f[kItems];
#pragma unroll
            for (int e = 0; e < kItems; e++)  // cnt more set bits: m[e] != 0
                f[e] = s_cs[__builtin_ctz(m[e])];
#pragma unroll
            for (int e = 0; e < kItems; e++) {
                one_update<T>(QSLHIP_PHASE, a[e], f[e].re, f[e].im);
                m[e] &= m[e] - 1;
            }
        }
    } else {
        while (m[0] | m[1] | m[2] | m[3]) {
            cplx<T> f[kItems];
#pragma unroll
            for (int e = 0; e < kItems; e++) f[e] = s_cs[(__ffs(m[e]) - 1) & (kPhaseMax - 1)];
#pragma unroll
            for (int e = 0; e < kItems; e++) {
                cplx<T> b = a[e];
                one_update<T>(QSLHIP_PHASE, b, f[e].re, f[e].im);
                const bool act = m[e] != 0;
                a[e].re = act ? b.re : a[e].re;
                a[e].im = act ? b.im : a[e].im;
                m[e] &= m[e] - 1;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kItems; e++)
        if (in[e]) amp[g.off + g.i0 + e * kBlock + threadIdx.x] = a[e];
}

// Up to kBondMax swap-like gates (npRotateX/Y, npHadamard, swap) on
// pairwise disjoint qubit pairs, in program order.
constexpr int kBondMax = 4;
struct NPBonds {
    int m;
    int body[kBondMax];
    uint32_t q0[kBondMax], q1[kBondMax];
    double c[kBondMax], s[kBondMax];
};

// The gates of NPBonds in one pass.  A gate on (q0, q1) pairs each string
// with bit q0 set and q1 clear (its owner) with the string that has the two
// bits exchanged (k_np_gate's PC_SWAPLIKE).  With disjoint pairs the strings
// reachable from x by exchanging the bits of any subset of its "active"
// pairs (the two bits differ) form an orbit of 2^a strings that no other
// gate of the run leaves.  The thread whose string owns every active pair
// gathers the orbit, applies the gates in program order with the same
// pair_update as one launch per gate, and scatters it back: bit-identical.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_bonds(cplx<T> *__restrict__ amp, NPDev d,
                                                     NPBonds B)
{
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
#pragma unroll 1
    for (int e = 0; e < kItems; e++) {
        uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        if (i >= g.size || i >= g.i0 + kChunk) continue;
        const uint64_t x = hix | d.lo_list[g.cstart + i];
        unsigned active = 0;
        bool rep = true;
#pragma unroll
        for (int j = 0; j < kBondMax; j++) {
            if (j >= B.m) break;
            const bool a0 = (x >> B.q0[j]) & 1ull, a1 = (x >> B.q1[j]) & 1ull;
            if (a0 != a1) {
                active |= 1u << j;
                if (!a0) rep = false;  // another orbit member owns this pair
            }
        }
        if (!rep || !active) continue;
        cplx<T> v[1 << kBondMax];
        uint64_t rk[1 << kBondMax];
        const int na = 1 << B.m;
#pragma unroll
        for (int sub = 0; sub < (1 << kBondMax); sub++) {
            if (sub >= na || (sub & ~active)) continue;
            uint64_t y = x;
#pragma unroll
            for (int j = 0; j < kBondMax; j++)
                if ((sub >> j) & 1) y ^= (1ull << B.q0[j]) | (1ull << B.q1[j]);
            rk[sub] = sub ? np_rank(d, y) : (g.off + i);
            v[sub] = amp[rk[sub]];
        }
#pragma unroll
        for (int j = 0; j < kBondMax; j++) {
            if (j >= B.m || !((active >> j) & 1)) continue;
            const T c = (T)B.c[j], s = (T)B.s[j];
#pragma unroll
            for (int sub = 0; sub < (1 << kBondMax); sub++) {
                if (sub >= na || (sub & ~active) || ((sub >> j) & 1)) continue;
                // v[sub] owns pair j (bits q0 = 1, q1 = 0), its partner is sub | bit j
                const int par = sub | (1 << j);
                cplx<T> a = v[sub], b = v[par];
                pair_update<T>(B.body[j], a, b, c, s);
                v[sub] = a;
                v[par] = b;
            }
        }
#pragma unroll
        for (int sub = 0; sub < (1 << kBondMax); sub++) {
            if (sub >= na || (sub & ~active)) continue;
            amp[rk[sub]] = v[sub];
        }
    }
}

// Within a segment the strings are in numeric order, so those sharing bits
// kLowBits and up form a contiguous run (at most C(12, 6) = 924 ranks) that
// every adjacent swap-like gate below kLowBits maps to itself.
constexpr int kLowBits = 12;
constexpr int kLowBondMax = 8;
constexpr int kLowSpan = 1024;  // amplitudes per block: whole runs, at most C(12, 6) = 924 each
struct NPLowBonds {
    int m;
    int body[kLowBondMax];
    uint32_t q0[kLowBondMax], q1[kLowBondMax];
    double c[kLowBondMax], s[kLowBondMax];
};

// Adjacent swap-like gates on disjoint pairs below kLowBits, in one pass.
// Such a gate pairs strings that share every bit from kLowBits up, so the
// pairs stay inside a run; each block stages whole runs (the host's
// run-aligned block map) in LDS and applies the gates in program order,
// one barrier apart, each owner (bit q0 set, q1 clear) updating itself and
// its partner at rank i +- C(q, j - 1) (adj_partner) with pair_update: the
// same updates in the same order as one launch per gate, bit-identical.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_lowbonds(cplx<T> *__restrict__ amp, NPDev d,
                                                        const NPRunBlk *__restrict__ rb,
                                                        NPLowBonds B)
{
    __shared__ cplx<T> sv[kLowSpan];
    __shared__ uint32_t sbin[kLowBits][kLowBits + 1];  // C(q, m) for q < kLowBits
    const NPRunBlk b = rb[blockIdx.x];
    const uint32_t w = d.k - __popc(b.hi);
    const uint32_t *lo = d.lo_list + d.class_start[w];
    const uint64_t off = d.seg_off[b.hi];
    for (int t = threadIdx.x; t < kLowBits * (kLowBits + 1); t += kBlock)
        sbin[t / (kLowBits + 1)][t % (kLowBits + 1)] =
            (uint32_t)d.binom[(t / (kLowBits + 1)) * 64 + t % (kLowBits + 1)];
    constexpr int kPer = kLowSpan / kBlock;
    uint32_t xs[kPer];  // each thread's strings, for every gate
#pragma unroll
    for (int e = 0; e < kPer; e++) {
        const uint32_t i = b.first + threadIdx.x + e * kBlock;
        xs[e] = 0;
        if (i < b.end) {
            sv[i - b.first] = amp[off + i];
            xs[e] = lo[i];
        }
    }
    __syncthreads();
    for (int j = 0; j < B.m; j++) {
        const unsigned q0 = B.q0[j], q1 = B.q1[j], q = q0 < q1 ? q0 : q1;
        const T c = (T)B.c[j], s = (T)B.s[j];
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const uint32_t i = b.first + threadIdx.x + e * kBlock;
            const uint32_t x = xs[e];
            // owners: bit q0 set, q1 clear (slots past `end` hold 0: never owners)
            if (!((x >> q0) & 1u) || ((x >> q1) & 1u)) continue;
            const uint32_t m = sbin[q][__popc(x & ((1u << q) - 1))];
            const uint32_t p = q1 > q0 ? i + m : i - m;
            cplx<T> a = sv[i - b.first], bb = sv[p - b.first];
            pair_update<T>(B.body[j], a, bb, c, s);
            sv[i - b.first] = a;
            sv[p - b.first] = bb;
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < kPer; e++) {
        const uint32_t i = b.first + threadIdx.x + e * kBlock;
        if (i < b.end) amp[off + i] = sv[i - b.first];
    }
}

// Up to kBondMax swap-like gates on pairwise disjoint qubit pairs that all
// lie in the high part of the strings, in one pass.  Such a gate maps
// segment hi onto segment hi ^ b0 ^ b1 element for element, so the segments
// reachable from hi by exchanging the bits of its active pairs (the two
// bits differ) form an orbit of 2^a segments of the same weight.  The
// blocks of the orbit's representative (every active pair holds bit q0 = 1,
// q1 = 0) gather element i of each member, apply the gates in program
// order with pair_update and scatter them back: the same updates in the same
// order as one launch per gate, so the result is bit-identical.  Other
// blocks return at once.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_hibonds(cplx<T> *__restrict__ amp, NPDev d,
                                                       NPBonds B)
{
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
    unsigned active = 0;
#pragma unroll
    for (int j = 0; j < kBondMax; j++) {
        if (j >= B.m) break;
        const bool a0 = (hix >> B.q0[j]) & 1ull, a1 = (hix >> B.q1[j]) & 1ull;
        if (a0 != a1) {
            if (!a0) return;  // another member of the orbit represents it
            active |= 1u << j;
        }
    }
    if (!active) return;
    uint64_t off[1 << kBondMax];
#pragma unroll
    for (int sub = 0; sub < (1 << kBondMax); sub++) {
        off[sub] = 0;
        if (sub & ~active) continue;
        uint64_t y = hix;
#pragma unroll
        for (int j = 0; j < kBondMax; j++)
            if ((sub >> j) & 1) y ^= (1ull << B.q0[j]) | (1ull << B.q1[j]);
        off[sub] = sub ? d.seg_off[y >> d.L] : g.off;
    }
#pragma unroll 1
    for (int e = 0; e < kItems; e++) {
        const uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        if (i >= g.size || i >= g.i0 + kChunk) continue;
        cplx<T> v[1 << kBondMax];
#pragma unroll
        for (int sub = 0; sub < (1 << kBondMax); sub++)
            if (!(sub & ~active)) v[sub] = amp[off[sub] + i];
#pragma unroll
        for (int j = 0; j < kBondMax; j++) {
            if (j >= B.m || !((active >> j) & 1)) continue;
            const T c = (T)B.c[j], s = (T)B.s[j];
#pragma unroll
            for (int sub = 0; sub < (1 << kBondMax); sub++) {
                if ((sub & ~active) || ((sub >> j) & 1)) continue;
                // v[sub] holds q0 = 1, q1 = 0 for pair j; its partner is sub | bit j
                const int par = sub | (1 << j);
                cplx<T> a = v[sub], b = v[par];
                pair_update<T>(B.body[j], a, b, c, s);
                v[sub] = a;
                v[par] = b;
            }
        }
#pragma unroll
        for (int sub = 0; sub < (1 << kBondMax); sub++)
            if (!(sub & ~active)) amp[off[sub] + i] = v[sub];
    }
}

// Up to kHiBondMax swap-like gates on pairwise disjoint pairs in the high
// part (tuning np_hibonds 2), in one pass: the orbit of 2^a segments (a
// active pairs) of k_np_hibonds, staged through LDS instead of registers,
// X = kHiLds / 2^a elements of every member per round.  The gates run as
// butterflies over the staged orbit in program order, a barrier between
// gates: the same updates in the same order as k_np_hibonds, so the same
// bits as one launch per gate.  The next round's elements are loaded into
// registers while a round's bonds run.
constexpr int kHiBondMax = 8;
struct NPHiBonds {
    int m;
    int body[kHiBondMax];
    uint32_t q0[kHiBondMax], q1[kHiBondMax];
    double c[kHiBondMax], s[kHiBondMax];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_hibonds_lds(cplx<T> *__restrict__ amp, NPDev d,
                                                           NPHiBonds B)
{
    constexpr uint32_t kHiLds = 32768 / sizeof(cplx<T>);  // 2^11 FP64, 2^12 FP32 amplitudes
    constexpr int kHiLog = sizeof(cplx<T>) == 16 ? 11 : 12;
    static_assert(kHiLds == (1u << kHiLog) && kHiLog > kHiBondMax, "staging size");
    __shared__ cplx<T> sv[kHiLds];
    __shared__ uint64_t soff[1 << kHiBondMax];
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
    uint32_t active = 0;
    for (int j = 0; j < B.m; j++) {
        const bool a0 = (hix >> B.q0[j]) & 1ull, a1 = (hix >> B.q1[j]) & 1ull;
        if (a0 != a1) {
            if (!a0) return;  // another member of the orbit represents it
            active |= 1u << j;
        }
    }
    if (!active) return;
    const int a = __popc(active);
    // member t of the orbit: the segment with the pairs of t's set bits
    // exchanged (bit k of t: the k-th active pair in program order)
    if (threadIdx.x < (1u << a)) {
        uint64_t y = hix;
        int k = 0;
        for (int j = 0; j < B.m; j++) {
            if (!((active >> j) & 1u)) continue;
            if ((threadIdx.x >> k) & 1u) y ^= (1ull << B.q0[j]) | (1ull << B.q1[j]);
            k++;
        }
        soff[threadIdx.x] = threadIdx.x ? d.seg_off[y >> d.L] : g.off;
    }
    __syncthreads();
    const int xb = kHiLog - a;  // log2 X
    const uint32_t xm = (1u << xb) - 1;
    const uint32_t end = g.i0 + kChunk < g.size ? g.i0 + kChunk : g.size;
    // staged element idx = threadIdx.x + r * kBlock: member idx >> xb,
    // element e0 + (idx & xm).  The next round's loads are issued once this
    // round is in LDS, so they land during its bonds and stores.
    constexpr int kPer = kHiLds / kBlock;
    cplx<T> nx[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const uint32_t idx = threadIdx.x + r * kBlock, i = g.i0 + (idx & xm);
        nx[r] = i < end ? amp[soff[idx >> xb] + i] : cplx<T>{};
    }
    for (uint32_t e0 = g.i0; e0 < end; e0 += xm + 1) {
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const uint32_t idx = threadIdx.x + r * kBlock;
            if (e0 + (idx & xm) < end) sv[idx] = nx[r];
        }
        __syncthreads();
        const uint32_t e1 = e0 + xm + 1;
        if (e1 < end) {
#pragma unroll
            for (int r = 0; r < kPer; r++) {
                const uint32_t idx = threadIdx.x + r * kBlock, i = e1 + (idx & xm);
                nx[r] = i < end ? amp[soff[idx >> xb] + i] : cplx<T>{};
            }
        }
        for (int j = 0; j < B.m; j++) {
            if (!((active >> j) & 1u)) continue;
            const int k = __popc(active & ((1u << j) - 1));
            const T c = (T)B.c[j], s = (T)B.s[j];
            const int body = B.body[j];
            for (uint32_t p = threadIdx.x; p < kHiLds / 2; p += kBlock) {
                const uint32_t el = p & xm;
                if (e0 + el >= end) continue;
                // member with bit k clear: pair j holds q0 = 1, q1 = 0 (the
                // owner); its partner has bit k set
                const uint32_t m0 = (uint32_t)ins0(p >> xb, (unsigned)k);
                const uint32_t i0 = (m0 << xb) | el, i1 = ((m0 | (1u << k)) << xb) | el;
                cplx<T> x0 = sv[i0], x1 = sv[i1];
                pair_update<T>(body, x0, x1, c, s);
                sv[i0] = x0;
                sv[i1] = x1;
            }
            __syncthreads();
        }
        for (uint32_t idx = threadIdx.x; idx < kHiLds; idx += kBlock) {
            const uint32_t i = e0 + (idx & xm);
            if (i < end) amp[soff[idx >> xb] + i] = sv[idx];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Per-block partial of sum |a|^2 over strings with bit t == outcome.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_prob(const cplx<T> *__restrict__ amp,
                                                    NPDev d, unsigned t,
                                                    unsigned outcome,
                                                    double *__restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    const Seg g = seg_of(d);
    const uint64_t tb = 1ull << t;
    const uint64_t hix = (uint64_t)g.hi << d.L;
    double acc = 0;
    bool seg_skip = (t >= d.L) && (((hix & tb) != 0) != (outcome != 0));
    if (!seg_skip) {
        for (int e = 0; e < kItems; e++) {
            uint32_t i = g.i0 + e * kBlock + threadIdx.x;
            if (i >= g.size || i >= g.i0 + kChunk) continue;
            uint64_t x = hix | d.lo_list[g.cstart + i];
            if (((x & tb) != 0) == (outcome != 0)) {
                cplx<T> a = amp[g.off + i];
                T p = a.re * a.re + a.im * a.im;
                acc += (double)p;
            }
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < kBlock / 64; w++) r += sh[w];
        part[blockIdx.x] = r;
    }
}

// Deterministic sum of n doubles into *out (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void k_sum_all(const double *__restrict__ v, uint64_t n,
                                                  double *__restrict__ out)
{
    __shared__ double sh[1024];
    double acc = 0;
    for (uint64_t i = threadIdx.x; i < n; i += 1024) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 512; off >= 1; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_collapse(cplx<T> *__restrict__ amp, NPDev d,
                                                        unsigned t, unsigned outcome,
                                                        double factor)
{
    const Seg g = seg_of(d);
    const uint64_t tb = 1ull << t;
    const uint64_t hix = (uint64_t)g.hi << d.L;
    for (int e = 0; e < kItems; e++) {
        uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        if (i >= g.size || i >= g.i0 + kChunk) continue;
        uint64_t x = hix | d.lo_list[g.cstart + i];
        cplx<T> a = amp[g.off + i];
        if (((x & tb) != 0) == (outcome != 0)) {
            a.re = (T)((double)a.re * factor);
            a.im = (T)((double)a.im * factor);
        } else {
            a.re = 0;
            a.im = 0;
        }
        amp[g.off + i] = a;
    }
}

// full[x] <-> amp[rank(x)]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_np_scatter(const cplx<T> *__restrict__ amp,
                                                       NPDev d, cplx<T> *__restrict__ full,
                                                       int to_full)
{
    const Seg g = seg_of(d);
    const uint64_t hix = (uint64_t)g.hi << d.L;
    for (int e = 0; e < kItems; e++) {
        uint32_t i = g.i0 + e * kBlock + threadIdx.x;
        if (i >= g.size || i >= g.i0 + kChunk) continue;
        uint64_t x = hix | d.lo_list[g.cstart + i];
        if (to_full) full[x] = amp[g.off + i];
        else const_cast<cplx<T> *>(amp)[g.off + i] = full[x];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero(cplx<T> *__restrict__ v, uint64_t n)
{
    uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        cplx<T> z;
        z.re = 0;
        z.im = 0;
        v[i] = z;
    }
}

// CDF search over ranks, then unrank to the basis index.
__global__ __launch_bounds__(kBlock) void k_np_search(const double *__restrict__ cdf,
                                                      uint64_t namps, NPDev d,
                                                      uint64_t nseg,
                                                      const double *__restrict__ u,
                                                      uint64_t nshots, int strict,
                                                      uint64_t *__restrict__ out)
{
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nshots) return;
    double uu = u[i];
    bool zero = strict || !(uu > 0.0);
    double key = strict ? uu : 0.0;
    uint64_t lo = 0, hi = namps;
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        double v = cdf[mid];
        bool ge = zero ? (v > key) : (v >= uu);
        if (ge) hi = mid;
        else lo = mid + 1;
    }
    if (lo >= namps) {
        out[i] = strict ? ~0ull : 0ull;
        return;
    }
    // last segment whose start is <= lo
    uint64_t a = 0, b = nseg;  // seg_off has nseg + 1 entries
    while (b - a > 1) {
        uint64_t m = (a + b) >> 1;
        if (d.seg_off[m] <= lo) a = m;
        else b = m;
    }
    uint64_t idx = lo - d.seg_off[a];
    unsigned w = d.k - __popcll(a);
    out[i] = (a << d.L) | d.lo_list[d.class_start[w] + idx];
}

template <typename T, bool S>
void gate_impl(qslhip_state *h, const GateDesc &g)
{
    NPDev d = dev_view(h);
    cplx<T> *amp = static_cast<cplx<T> *>(h->amps);
    dim3 grid(h->npt->nblk), blk(kBlock);
    const T c = (T)g.c, s = (T)g.s;
    const int body = body_of(g.kind);
    switch (g.kind) {
    case QSLHIP_PAULI_Z: case QSLHIP_PHASE:
        hipLaunchKernelGGL((k_np_gate<T, PC_ONE, S>), grid, blk, 0, h->stream, amp, d, body,
                           g.q0, 0u, c, s);
        break;
    case QSLHIP_ROT_Z:
        hipLaunchKernelGGL((k_np_gate<T, PC_PAIR1, S>), grid, blk, 0, h->stream, amp, d, body,
                           g.q0, 0u, c, s);
        break;
    case QSLHIP_CZ: case QSLHIP_CPHASE:
        hipLaunchKernelGGL((k_np_gate<T, PC_DIAG2, S>), grid, blk, 0, h->stream, amp, d, body,
                           g.q0, g.q1, c, s);
        break;
    case QSLHIP_CROT_Z:
        hipLaunchKernelGGL((k_np_gate<T, PC_CTRL, S>), grid, blk, 0, h->stream, amp, d, body,
                           g.q0, g.q1, c, s);
        break;
    case QSLHIP_SWAP: case QSLHIP_NP_ROT_X: case QSLHIP_NP_ROT_Y:
    case QSLHIP_NP_HADAMARD: {
        // adjacent pairs low in the strings: members a few ranks apart, one
        // element per lane with cross-lane exchange
        const unsigned ql = g.q0 < g.q1 ? g.q0 : g.q1, qh = g.q0 < g.q1 ? g.q1 : g.q0;
        if (qh == ql + 1 && qh < h->npt->L && (int)ql <= tuning().np_adj_max) {
            const int pol = tuning().np_adj_policy;
#define ADJ(SV, AV)                                                                     \
    hipLaunchKernelGGL((k_np_adj<T, SV, AV>), grid, blk, 0, h->stream, amp, d, body, g.q0, \
                       g.q1, c, s)
            if (pol == 0) ADJ(false, false);
            else if (pol == 1) ADJ(true, false);
            else if (pol == 2) ADJ(false, true);
            else ADJ(true, true);
#undef ADJ
            break;
        }
        hipLaunchKernelGGL((k_np_gate<T, PC_SWAPLIKE, S>), grid, blk, 0, h->stream, amp, d,
                           body, g.q0, g.q1, c, s);
        break;
    }
    case QSLHIP_FSWAP:
        hipLaunchKernelGGL((k_np_gate<T, PC_FSWAP, S>), grid, blk, 0, h->stream, amp, d, body,
                           g.q0, g.q1, c, s);
        break;
    default:
        throw Error(QSLHIP_EKIND, "gate does not preserve the number of ones");
    }
    QH_CHECK(hipGetLastError());
}

template <typename T>
double prob_impl(qslhip_state *h, unsigned t, unsigned outcome)
{
    NPDev d = dev_view(h);
    uint32_t nb = h->npt->nblk;
    h->red.ensure(((size_t)nb + 8) * sizeof(double));
    double *part = static_cast<double *>(h->red.ptr);
    double *res = part + nb;
    hipLaunchKernelGGL(k_np_prob<T>, dim3(nb), dim3(kBlock), 0, h->stream,
                       static_cast<const cplx<T> *>(h->amps), d, t, outcome, part);
    hipLaunchKernelGGL(k_sum_all, dim3(1), dim3(1024), 0, h->stream, part, (uint64_t)nb,
                       res);
    QH_CHECK(hipGetLastError());
    h->pin.ensure(64);
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    QH_CHECK(hipMemcpyAsync(h->pin.ptr, res, sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    double v;
    std::memcpy(&v, h->pin.ptr, sizeof(double));
    return v;
}

}  // namespace

uint64_t np_rank_host(uint64_t x)
{
    uint64_t r = 0;
    unsigned i = 0;
    while (x) {
        unsigned p = (unsigned)__builtin_ctzll(x);
        i++;
        r += choose_host(p, i);
        x &= x - 1;
    }
    return r;
}

void np_init(qslhip_state *h)
{
    auto *t = new NPTables();
    h->npt = t;
    const unsigned n = h->n, k = h->k;
    unsigned L = n <= 16 ? n : (n - 16 > 16 ? n - 16 : 16);
    if (L > n) L = n;
    const unsigned hib = n - L;
    t->L = L;
    t->hib = hib;
    QH_REQUIRE(L <= 24 && hib <= 24, QSLHIP_EINVAL, "NP handle too large");
    // Low-part tables.
    const uint64_t nlo = 1ull << L;
    t->h_class_start.assign(L + 2, 0);
    for (uint64_t p = 0; p < nlo; p++) t->h_class_start[__builtin_popcountll(p) + 1]++;
    for (unsigned w = 1; w < L + 2; w++) t->h_class_start[w] += t->h_class_start[w - 1];
    t->h_lo_list.assign(nlo, 0);
    t->h_rank_lo.assign(nlo, 0);
    std::vector<uint32_t> fill(t->h_class_start.begin(), t->h_class_start.end() - 1);
    for (uint64_t p = 0; p < nlo; p++) {  // increasing p: colex order per weight
        unsigned w = (unsigned)__builtin_popcountll(p);
        uint32_t pos = fill[w]++;
        t->h_lo_list[pos] = (uint32_t)p;
        t->h_rank_lo[p] = pos - t->h_class_start[w];
    }
    // Segment offsets: segment hi holds C(L, k - popcount(hi)) strings.
    const uint64_t nseg = 1ull << hib;
    t->h_seg_off.assign(nseg + 1, 0);
    std::vector<NPBlk> blks;
    for (uint64_t s = 0; s < nseg; s++) {
        unsigned wh = (unsigned)__builtin_popcountll(s);
        uint64_t size = (wh <= k && k - wh <= L) ? choose_host(L, k - wh) : 0;
        t->h_seg_off[s + 1] = t->h_seg_off[s] + size;
        for (uint64_t i0 = 0; i0 < size; i0 += kChunk) blks.push_back({(uint32_t)s, (uint32_t)i0});
    }
    QH_REQUIRE(t->h_seg_off[nseg] == h->namps, QSLHIP_EDEVICE,
               "NP segment table does not add up to C(n,k)");
    t->nblk = (uint32_t)blks.size();
    QH_CHECK(hipMalloc(&t->seg_off, (nseg + 1) * sizeof(uint64_t)));
    QH_CHECK(hipMalloc(&t->lo_list, nlo * sizeof(uint32_t)));
    QH_CHECK(hipMalloc(&t->rank_lo, nlo * sizeof(uint32_t)));
    QH_CHECK(hipMalloc(&t->class_start, (L + 2) * sizeof(uint32_t)));
    QH_CHECK(hipMalloc(&t->blk, blks.size() * sizeof(NPBlk)));
    std::vector<uint64_t> binom(64 * 64);
    for (unsigned q = 0; q < 64; q++)
        for (unsigned m = 0; m < 64; m++) binom[q * 64 + m] = choose_host(q, m);
    QH_CHECK(hipMalloc(&t->binom, binom.size() * sizeof(uint64_t)));
    QH_CHECK(hipMemcpy(t->binom, binom.data(), binom.size() * sizeof(uint64_t),
                       hipMemcpyHostToDevice));
    QH_CHECK(hipMemcpy(t->seg_off, t->h_seg_off.data(), (nseg + 1) * sizeof(uint64_t),
                       hipMemcpyHostToDevice));
    QH_CHECK(hipMemcpy(t->lo_list, t->h_lo_list.data(), nlo * sizeof(uint32_t),
                       hipMemcpyHostToDevice));
    QH_CHECK(hipMemcpy(t->rank_lo, t->h_rank_lo.data(), nlo * sizeof(uint32_t),
                       hipMemcpyHostToDevice));
    QH_CHECK(hipMemcpy(t->class_start, t->h_class_start.data(), (L + 2) * sizeof(uint32_t),
                       hipMemcpyHostToDevice));
    QH_CHECK(hipMemcpy(t->blk, blks.data(), blks.size() * sizeof(NPBlk),
                       hipMemcpyHostToDevice));
}

void np_free(qslhip_state *h)
{
    NPTables *t = h->npt;
    if (!t) return;
    (void)hipDeviceSynchronize();
    if (t->seg_off) (void)hipFree(t->seg_off);
    if (t->lo_list) (void)hipFree(t->lo_list);
    if (t->rank_lo) (void)hipFree(t->rank_lo);
    if (t->class_start) (void)hipFree(t->class_start);
    if (t->blk) (void)hipFree(t->blk);
    if (t->runblk) (void)hipFree(t->runblk);
    if (t->binom) (void)hipFree(t->binom);
    delete t;
    h->npt = nullptr;
}

void np_launch_gate(qslhip_state *h, const GateDesc &g)
{
    // Streaming accesses where each lane's run is long: a gate whose lowest
    // qubit is q touches runs of C(q, .) consecutive ranks, whole segments
    // when both qubits are in the high part of the strings.
    unsigned ql = g.q0;
    if (is_gate2(g.kind) && g.q1 < ql) ql = g.q1;
    const bool stream = (int)h->n >= tuning().stream_min_qubits &&
                        (int)ql >= tuning().np_stream_min_q;
    if (h->prec == 8) {
        if (stream) gate_impl<double, true>(h, g);
        else gate_impl<double, false>(h, g);
    } else {
        if (stream) gate_impl<float, true>(h, g);
        else gate_impl<float, false>(h, g);
    }
}

// The class of an elementwise (diagonal) NP gate, or -1.
static int diag_class(int kind)
{
    switch (kind) {
    case QSLHIP_PAULI_Z: case QSLHIP_PHASE: return PC_ONE;
    case QSLHIP_ROT_Z: return PC_PAIR1;
    case QSLHIP_CZ: case QSLHIP_CPHASE: return PC_DIAG2;
    case QSLHIP_CROT_Z: return PC_CTRL;
    default: return -1;
    }
}

static bool bond_kind(int kind)
{
    return kind == QSLHIP_SWAP || kind == QSLHIP_NP_ROT_X || kind == QSLHIP_NP_ROT_Y ||
           kind == QSLHIP_NP_HADAMARD;
}

// The run-aligned block map of k_np_lowbonds, built once per handle: in
// each segment, the strings sharing their low-part bits from bit 12 up form
// runs of C(12, w') consecutive ranks (in increasing order of those bits),
// packed whole into blocks of at most kLowSpan ranks.
static void np_run_blocks(qslhip_state *h)
{
    NPTables *t = h->npt;
    if (t->runblk) return;
    const unsigned L = t->L, lb = L < (unsigned)kLowBits ? L : (unsigned)kLowBits, hb = L - lb;
    std::vector<NPRunBlk> blks;
    const uint64_t nseg = 1ull << t->hib;
    for (uint64_t sg = 0; sg < nseg; sg++) {
        const unsigned wh = (unsigned)__builtin_popcountll(sg);
        if (wh > h->k || h->k - wh > L) continue;
        const unsigned wl = h->k - wh;
        NPRunBlk cur{(uint32_t)sg, 0, 0};
        for (uint32_t v = 0; v < (1u << hb); v++) {
            const unsigned pv = (unsigned)__builtin_popcount(v);
            if (pv > wl || wl - pv > lb) continue;
            const uint32_t len = (uint32_t)choose_host(lb, wl - pv);
            if (cur.end - cur.first + len > (uint32_t)kLowSpan) {
                blks.push_back(cur);
                cur.first = cur.end;
            }
            cur.end += len;
        }
        QH_REQUIRE(cur.end == t->h_seg_off[sg + 1] - t->h_seg_off[sg], QSLHIP_EDEVICE,
                   "NP runs do not add up to the segment");
        if (cur.end > cur.first) blks.push_back(cur);
    }
    QH_REQUIRE(!blks.empty() && blks.size() < (1ull << 31), QSLHIP_EINVAL,
               "NP run-aligned block map out of range");
    QH_CHECK(hipMalloc(&t->runblk, blks.size() * sizeof(NPRunBlk)));
    QH_CHECK(hipMemcpy(t->runblk, blks.data(), blks.size() * sizeof(NPRunBlk),
                       hipMemcpyHostToDevice));
    t->nrunblk = (uint32_t)blks.size();
}

void np_apply_gates(qslhip_state *h, const std::vector<GateDesc> &gates)
{
    // Segments of the list: runs of two or more diagonal gates (one
    // k_np_diag pass each, their ops staged in one copy), groups of two to
    // kBondMax swap-like gates on disjoint qubit pairs (one k_np_bonds pass
    // each), and single gates.
    enum { SINGLE, DIAG, PHASES, BONDS, LOWBONDS, HIBONDS };
    struct Part { int type; size_t first, end; };
    std::vector<Part> parts;
    for (size_t i = 0; i < gates.size();) {
        size_t j = i;
        bool phases = tuning().np_phase_lut != 0;
        while (j < gates.size() && diag_class(gates[j].kind) >= 0) {
            phases = phases && gates[j].kind == QSLHIP_PHASE;
            j++;
        }
        if (j - i >= 2 && phases) {  // phase gates only: k_np_phases, kPhaseMax a pass
            for (size_t f = i; f < j; f += kPhaseMax)
                parts.push_back({PHASES, f, f + kPhaseMax < j ? f + kPhaseMax : j});
            i = j;
            continue;
        }
        if (j - i >= 2) {
            parts.push_back({DIAG, i, j});
            i = j;
            continue;
        }
        uint64_t used = 0;
        // adjacent bonds below kLowBits (and inside the low part of the
        // strings), pairwise disjoint: one k_np_lowbonds pass
        const unsigned lowq = h->npt->L < (unsigned)kLowBits ? h->npt->L : (unsigned)kLowBits;
        j = i;
        while (tuning().np_lowbonds && j < gates.size() && j - i < (size_t)kLowBondMax &&
               bond_kind(gates[j].kind)) {
            const GateDesc &b = gates[j];
            const unsigned lo = b.q0 < b.q1 ? b.q0 : b.q1, hi = b.q0 < b.q1 ? b.q1 : b.q0;
            if (hi != lo + 1 || hi >= lowq) break;
            const uint64_t m = (1ull << b.q0) | (1ull << b.q1);
            if (used & m) break;
            used |= m;
            j++;
        }
        if (j - i >= 2) {
            parts.push_back({LOWBONDS, i, j});
            i = j;
            continue;
        }
        // swap-like gates on disjoint pairs in the high part: one
        // k_np_hibonds pass per group of up to kBondMax
        used = 0;
        j = i;
        const size_t hmax = tuning().np_hibonds == 2 ? (size_t)kHiBondMax : (size_t)kBondMax;
        while (tuning().np_hibonds && j < gates.size() && j - i < hmax &&
               bond_kind(gates[j].kind) && gates[j].q0 >= h->npt->L &&
               gates[j].q1 >= h->npt->L) {
            const uint64_t m = (1ull << gates[j].q0) | (1ull << gates[j].q1);
            if (used & m) break;
            used |= m;
            j++;
        }
        if (j - i >= 2) {
            parts.push_back({HIBONDS, i, j});
            i = j;
            continue;
        }
        used = 0;
        j = i;
        const size_t gmax = (size_t)tuning().np_bonds;
        while (j < gates.size() && j - i < gmax && bond_kind(gates[j].kind)) {
            const uint64_t m = (1ull << gates[j].q0) | (1ull << gates[j].q1);
            if (used & m) break;
            used |= m;
            j++;
        }
        if (j - i >= 2) {
            parts.push_back({BONDS, i, j});
            i = j;
            continue;
        }
        parts.push_back({SINGLE, i, i + 1});
        i++;
    }
    std::vector<NPDiagOp> dops;
    for (const Part &pt : parts) {
        if (pt.type != DIAG) continue;
        for (size_t k = pt.first; k < pt.end; k++) {
            const GateDesc &g = gates[k];
            NPDiagOp o;
            o.cls = diag_class(g.kind);
            o.body = body_of(g.kind);
            o.q0 = g.q0;
            o.q1 = is_gate2(g.kind) ? g.q1 : g.q0;
            o.c = g.c;
            o.s = g.s;
            dops.push_back(o);
        }
    }
    // the phase runs' tables, after the diagonal ops in the same staging
    // copy: per run its lut (nbytes x 256 words) and its (c, s) pairs
    const unsigned L = h->npt->L;
    const int nbytes = (int)((L + 7) / 8);
    QH_REQUIRE(nbytes <= kPhaseBytes, QSLHIP_EINVAL, "NP low part wider than the phase table");
    const size_t dbytes = (dops.size() * sizeof(NPDiagOp) + 15) & ~(size_t)15;
    std::vector<unsigned char> ptab;
    for (const Part &pt : parts) {
        if (pt.type != PHASES) continue;
        std::vector<uint32_t> lut((size_t)nbytes * 256, 0u);
        std::vector<double> cs;
        for (size_t k = pt.first; k < pt.end; k++) {
            const unsigned q = gates[k].q0, bit = (unsigned)(k - pt.first);
            if (q < L)
                for (unsigned v = 0; v < 256; v++)
                    if ((v >> (q & 7)) & 1u) lut[(q >> 3) * 256 + v] |= 1u << bit;
            cs.push_back(gates[k].c);
            cs.push_back(gates[k].s);
        }
        const size_t at = ptab.size();
        ptab.resize(at + lut.size() * 4 + cs.size() * 8);
        std::memcpy(ptab.data() + at, lut.data(), lut.size() * 4);
        std::memcpy(ptab.data() + at + lut.size() * 4, cs.data(), cs.size() * 8);
    }
    const NPDiagOp *dev = nullptr;
    const unsigned char *dev_ptab = nullptr;
    if (!dops.empty() || !ptab.empty()) {
        const size_t bytes = dbytes + ptab.size();
        if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));  // pin may still be read
        h->pin.ensure(bytes);
        h->optab.ensure(bytes);
        std::memcpy(h->pin.ptr, dops.data(), dops.size() * sizeof(NPDiagOp));
        std::memcpy(static_cast<unsigned char *>(h->pin.ptr) + dbytes, ptab.data(), ptab.size());
        QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, bytes, hipMemcpyHostToDevice,
                                h->stream));
        QH_CHECK(hipEventRecord(h->staged, h->stream));
        dev = static_cast<const NPDiagOp *>(h->optab.ptr);
        dev_ptab = static_cast<const unsigned char *>(h->optab.ptr) + dbytes;
    }
    NPDev d = dev_view(h);
    const dim3 grid(h->npt->nblk), blk(kBlock);
    size_t first_op = 0, ptab_at = 0;
    for (const Part &pt : parts) {
        if (pt.type == SINGLE) {
            np_launch_gate(h, gates[pt.first]);
            continue;
        }
        if (pt.type == PHASES) {
            NPPhaseRun run{};
            run.nops = (int)(pt.end - pt.first);
            run.nbytes = nbytes;
            for (int k = 0; k < run.nops; k++) {
                run.q[k] = gates[pt.first + k].q0;
                if (run.q[k] < L) run.lowmask |= 1u << k;
            }
            const uint32_t *lut = reinterpret_cast<const uint32_t *>(dev_ptab + ptab_at);
            const double *cs = reinterpret_cast<const double *>(dev_ptab + ptab_at +
                                                                (size_t)nbytes * 256 * 4);
            ptab_at += (size_t)nbytes * 256 * 4 + (size_t)run.nops * 16;
            if (h->prec == 8)
                hipLaunchKernelGGL(k_np_phases<double>, grid, blk, 0, h->stream,
                                   static_cast<cplx<double> *>(h->amps), d, lut, cs, run);
            else
                hipLaunchKernelGGL(k_np_phases<float>, grid, blk, 0, h->stream,
                                   static_cast<cplx<float> *>(h->amps), d, lut, cs, run);
            QH_CHECK(hipGetLastError());
            continue;
        }
        if (pt.type == DIAG) {
            const int nops = (int)(pt.end - pt.first);
            if (h->prec == 8)
                hipLaunchKernelGGL(k_np_diag<double>, grid, blk, 0, h->stream,
                                   static_cast<cplx<double> *>(h->amps), d, dev + first_op, nops);
            else
                hipLaunchKernelGGL(k_np_diag<float>, grid, blk, 0, h->stream,
                                   static_cast<cplx<float> *>(h->amps), d, dev + first_op, nops);
            first_op += (size_t)nops;
        } else if (pt.type == LOWBONDS) {
            NPLowBonds B{};
            B.m = (int)(pt.end - pt.first);
            for (int j = 0; j < B.m; j++) {
                const GateDesc &g = gates[pt.first + j];
                B.body[j] = body_of(g.kind);
                B.q0[j] = g.q0;
                B.q1[j] = g.q1;
                B.c[j] = g.c;
                B.s[j] = g.s;
            }
            np_run_blocks(h);
            const dim3 rgrid(h->npt->nrunblk);
            if (h->prec == 8)
                hipLaunchKernelGGL(k_np_lowbonds<double>, rgrid, blk, 0, h->stream,
                                   static_cast<cplx<double> *>(h->amps), d, h->npt->runblk, B);
            else
                hipLaunchKernelGGL(k_np_lowbonds<float>, rgrid, blk, 0, h->stream,
                                   static_cast<cplx<float> *>(h->amps), d, h->npt->runblk, B);
        } else if (pt.type == HIBONDS && pt.end - pt.first > (size_t)kBondMax) {
            NPHiBonds B{};
            B.m = (int)(pt.end - pt.first);
            QH_REQUIRE(B.m <= kHiBondMax, QSLHIP_EINVAL, "NP high-bond group too large");
            for (int j = 0; j < B.m; j++) {
                const GateDesc &g = gates[pt.first + j];
                B.body[j] = body_of(g.kind);
                B.q0[j] = g.q0;
                B.q1[j] = g.q1;
                B.c[j] = g.c;
                B.s[j] = g.s;
            }
            if (h->prec == 8)
                hipLaunchKernelGGL(k_np_hibonds_lds<double>, grid, blk, 0, h->stream,
                                   static_cast<cplx<double> *>(h->amps), d, B);
            else
                hipLaunchKernelGGL(k_np_hibonds_lds<float>, grid, blk, 0, h->stream,
                                   static_cast<cplx<float> *>(h->amps), d, B);
        } else {  // BONDS, HIBONDS
            NPBonds B{};
            B.m = (int)(pt.end - pt.first);
            for (int j = 0; j < B.m; j++) {
                const GateDesc &g = gates[pt.first + j];
                B.body[j] = body_of(g.kind);
                B.q0[j] = g.q0;
                B.q1[j] = g.q1;
                B.c[j] = g.c;
                B.s[j] = g.s;
            }
            if (pt.type == HIBONDS) {
                if (h->prec == 8)
                    hipLaunchKernelGGL(k_np_hibonds<double>, grid, blk, 0, h->stream,
                                       static_cast<cplx<double> *>(h->amps), d, B);
                else
                    hipLaunchKernelGGL(k_np_hibonds<float>, grid, blk, 0, h->stream,
                                       static_cast<cplx<float> *>(h->amps), d, B);
            } else if (h->prec == 8) {
                hipLaunchKernelGGL(k_np_bonds<double>, grid, blk, 0, h->stream,
                                   static_cast<cplx<double> *>(h->amps), d, B);
            } else {
                hipLaunchKernelGGL(k_np_bonds<float>, grid, blk, 0, h->stream,
                                   static_cast<cplx<float> *>(h->amps), d, B);
            }
        }
        QH_CHECK(hipGetLastError());
    }
}

double np_prob(qslhip_state *h, unsigned t, unsigned outcome)
{
    return h->prec == 8 ? prob_impl<double>(h, t, outcome) : prob_impl<float>(h, t, outcome);
}

void np_collapse(qslhip_state *h, unsigned t, unsigned outcome, double factor)
{
    NPDev d = dev_view(h);
    if (h->prec == 8)
        hipLaunchKernelGGL(k_np_collapse<double>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<cplx<double> *>(h->amps), d, t, outcome,
                           factor);
    else
        hipLaunchKernelGGL(k_np_collapse<float>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<cplx<float> *>(h->amps), d, t, outcome,
                           factor);
    QH_CHECK(hipGetLastError());
}

void np_scatter_full(qslhip_state *h, void *dev_full)
{
    NPDev d = dev_view(h);
    uint64_t dim = 1ull << h->n;
    unsigned zg = grid_for(dim, kBlock * 4);
    if (h->prec == 8) {
        hipLaunchKernelGGL(k_zero<double>, dim3(zg), dim3(kBlock), 0, h->stream,
                           static_cast<cplx<double> *>(dev_full), dim);
        hipLaunchKernelGGL(k_np_scatter<double>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<const cplx<double> *>(h->amps), d,
                           static_cast<cplx<double> *>(dev_full), 1);
    } else {
        hipLaunchKernelGGL(k_zero<float>, dim3(zg), dim3(kBlock), 0, h->stream,
                           static_cast<cplx<float> *>(dev_full), dim);
        hipLaunchKernelGGL(k_np_scatter<float>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<const cplx<float> *>(h->amps), d,
                           static_cast<cplx<float> *>(dev_full), 1);
    }
    QH_CHECK(hipGetLastError());
}

void np_gather_full(qslhip_state *h, const void *dev_full)
{
    NPDev d = dev_view(h);
    if (h->prec == 8)
        hipLaunchKernelGGL(k_np_scatter<double>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<const cplx<double> *>(h->amps), d,
                           const_cast<cplx<double> *>(static_cast<const cplx<double> *>(dev_full)),
                           0);
    else
        hipLaunchKernelGGL(k_np_scatter<float>, dim3(h->npt->nblk), dim3(kBlock), 0,
                           h->stream, static_cast<const cplx<float> *>(h->amps), d,
                           const_cast<cplx<float> *>(static_cast<const cplx<float> *>(dev_full)),
                           0);
    QH_CHECK(hipGetLastError());
}

}  // namespace qslhip

// CDF over the compact vector in rank order (= increasing basis index, as
// the reference's generateDist walk over lookup(n, k), np/measure.cpp:72-92)
// reuses the full-vector scan kernels of measure.hip.
namespace qslhip {
void build_cdf_any(qslhip_state *h);

void np_sample(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out,
               bool strict)
{
    build_cdf_any(h);
    if (nshots == 0) {
        QH_CHECK(hipStreamSynchronize(h->stream));
        return;
    }
    h->shots.ensure(nshots * 16);
    double *du = static_cast<double *>(h->shots.ptr);
    uint64_t *dout = reinterpret_cast<uint64_t *>(du + nshots);
    QH_CHECK(hipMemcpyAsync(du, u, nshots * sizeof(double), hipMemcpyHostToDevice,
                            h->stream));
    NPDev d = dev_view(h);
    hipLaunchKernelGGL(k_np_search, dim3(grid_for(nshots, kBlock)), dim3(kBlock), 0,
                       h->stream, static_cast<const double *>(h->cdf.ptr), h->namps, d,
                       (uint64_t)1 << h->npt->hib, du, nshots, (int)strict, dout);
    QH_CHECK(hipGetLastError());
    QH_CHECK(hipMemcpyAsync(out, dout, nshots * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
}
}  // namespace qslhip
