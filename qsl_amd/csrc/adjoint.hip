// adjoint.hip -- expectation values and adjoint (reverse-sweep) gradients.
//
// Observable H = sum_t c_t P_t of Pauli strings; P_t given by (xmask,
// zmask): qubit q carries X if only its x bit is set, Z if only its z bit,
// Y if both, so  P|j> = i^nY (-1)^popcount(j & z) |j ^ x>.
//
// Gradients of E = <psi|H|psi>, psi = G_L...G_1 |psi_0>, for every
// parametrised gate G_g = exp(-i theta P_g / 2) (rotations; phase and the
// controlled forms have P_g = -2|1><1| or a projector-weighted Pauli):
//   lambda = H psi_L;  for g = L..1:
//     dE/dtheta_g = Im <lambda_g| P_g |psi_g>;
//     psi_{g-1} = G_g^dag psi_g;  lambda_{g-1} = G_g^dag lambda_g.
// The reference has no gradients (no counterpart); results are checked
// against central finite differences and the parameter-shift rule on the
// oracle.
//
// Two paths:
//   * plain: one circuit on a handle, one kernel per gate (FP64 or FP32)
//     -- the "per-circuit GPU path";
//   * batch: thousands of circuits, FP32 states with FP64 energies and
//     gradients, one persistent kernel (a workgroup per circuit, largest
//     first from an atomic queue), captured once in a hipGraph and replayed
//     for each new set of angles.  States of <= 12 qubits live entirely in
//     LDS; larger ones stream through 2^12-amplitude LDS tile passes from a
//     per-workgroup HBM slot.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "tile_ops.h"

namespace qslhip {

// Generator of a parametrised gate.
enum GenKind { GEN_NONE = 0, GEN_X, GEN_Y, GEN_Z, GEN_P1, GEN_NPX, GEN_NPY };

__host__ __device__ inline int gen_of(int kind)
{
    switch (kind) {
    case QSLHIP_ROT_X: case QSLHIP_CROT_X: return GEN_X;
    case QSLHIP_ROT_Y: case QSLHIP_CROT_Y: return GEN_Y;
    case QSLHIP_ROT_Z: case QSLHIP_CROT_Z: return GEN_Z;
    case QSLHIP_PHASE: case QSLHIP_CPHASE: return GEN_P1;
    case QSLHIP_NP_ROT_X: return GEN_NPX;
    case QSLHIP_NP_ROT_Y: return GEN_NPY;
    default: return GEN_NONE;
    }
}

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ double im_conj_mul(const cplx<T> &l, const cplx<T> &p)
{
    // Im(conj(l) * p)
    return (double)l.re * (double)p.im - (double)l.im * (double)p.re;
}
template <typename T>
__device__ __forceinline__ double re_conj_mul(const cplx<T> &l, const cplx<T> &p)
{
    return (double)l.re * (double)p.re + (double)l.im * (double)p.im;
}

// Contribution of element i (value lambda_i) to Im<lambda|P|psi>, reading
// psi through `at(index)`.  tb/cb: target / control bit masks in the index
// space of `i`; for NP generators tb = bit of q0 (A side), cb = bit of q1.
template <typename T, typename F>
__device__ __forceinline__ double gen_term(int gen, uint64_t i, const cplx<T> &lam,
                                           uint64_t tb, uint64_t cb, bool ctrl_on,
                                           F at)
{
    switch (gen) {
    case GEN_X:
        if (!ctrl_on) return 0;
        return im_conj_mul(lam, at(i ^ tb));
    case GEN_Y:  // (Y psi)_i = -i psi_{i|b} if bit 0, +i psi_{i&~b} if bit 1
        if (!ctrl_on) return 0;
        return (i & tb) ? re_conj_mul(lam, at(i ^ tb)) : -re_conj_mul(lam, at(i ^ tb));
    case GEN_Z:
        if (!ctrl_on) return 0;
        return (i & tb) ? -im_conj_mul(lam, at(i)) : im_conj_mul(lam, at(i));
    case GEN_P1:
        if (!ctrl_on || !(i & tb)) return 0;
        return -2.0 * im_conj_mul(lam, at(i));
    case GEN_NPX:
    case GEN_NPY: {
        bool a = (i & tb) != 0, b = (i & cb) != 0;
        if (a == b) return 0;
        cplx<T> p = at(i ^ tb ^ cb);
        if (gen == GEN_NPX) return im_conj_mul(lam, p);
        return a ? -re_conj_mul(lam, p) : re_conj_mul(lam, p);
    }
    default:
        return 0;
    }
}

// ---- plain path kernels (full vectors in HBM) ---------------------------------

struct DevTerms {
    const qslhip_pauli *t;
    int n;
};

// lambda = H psi; part[block] = sum_i Re(conj(psi_i) lambda_i) (FP64).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_apply_h(const cplx<T> *__restrict__ psi,
                                                    cplx<T> *__restrict__ lam,
                                                    uint64_t namps, DevTerms terms,
                                                    double *__restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    double acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < namps; i += stride) {
        double lr = 0, li = 0;
        for (int t = 0; t < terms.n; t++) {
            const qslhip_pauli P = terms.t[t];
            uint64_t j = i ^ P.xmask;
            cplx<T> v = psi[j];
            int ny = __popcll(P.xmask & P.zmask) & 3;
            double sg = (__popcll(j & P.zmask) & 1) ? -P.coef : P.coef;
            double vr = sg * (double)v.re, vi = sg * (double)v.im;
            // multiply by i^ny
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
        cplx<T> o;
        o.re = (T)lr;
        o.im = (T)li;
        lam[i] = o;
        cplx<T> p = psi[i];
        acc += (double)p.re * lr + (double)p.im * li;
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

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grad_partial(const cplx<T> *__restrict__ psi,
                                                         const cplx<T> *__restrict__ lam,
                                                         uint64_t namps, int gen,
                                                         uint64_t tb, uint64_t cb,
                                                         uint64_t ctrl,
                                                         double *__restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    double acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    auto at = [&](uint64_t j) { return psi[j]; };
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < namps; i += stride)
        acc += gen_term<T>(gen, i, lam[i], tb, cb, (i & ctrl) == ctrl, at);
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < kBlock / 64; w++) r += sh[w];
        part[blockIdx.x] = r;
    }
}

// out[0] (+)= sum of part[0..n) in a fixed order.
__global__ __launch_bounds__(kBlock) void k_finish(const double *__restrict__ part, int n,
                                                   double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off >= 1; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

template <typename T>
void plain_impl(qslhip_state *h, const std::vector<GateDesc> &gates,
                const qslhip_pauli *terms, uint64_t nterms, double *energy, double *grads)
{
    const uint64_t namps = h->namps;
    const size_t sbytes = namps * sizeof(cplx<T>);
    // forward on the handle: it keeps psi_L
    for (const GateDesc &g : gates) launch_gate(h, g);
    DevBuf work, lam, scratch;
    work.ensure(sbytes);
    lam.ensure(sbytes);
    const size_t nslots = gates.size() + 1;
    scratch.ensure((kRedBlocks + nslots) * sizeof(double) + nterms * sizeof(qslhip_pauli) + 64);
    double *part = static_cast<double *>(scratch.ptr);
    double *slots = part + kRedBlocks;  // [0] energy, [1 + g] grads
    qslhip_pauli *dterms = reinterpret_cast<qslhip_pauli *>(slots + nslots);
    QH_CHECK(hipMemcpyAsync(dterms, terms, nterms * sizeof(qslhip_pauli),
                            hipMemcpyHostToDevice, h->stream));
    QH_CHECK(hipMemcpyAsync(work.ptr, h->amps, sbytes, hipMemcpyDeviceToDevice, h->stream));
    unsigned grid = grid_for(namps, kBlock * 4);
    if (grid > (unsigned)kRedBlocks) grid = kRedBlocks;
    cplx<T> *psi = static_cast<cplx<T> *>(work.ptr);
    cplx<T> *lm = static_cast<cplx<T> *>(lam.ptr);
    hipLaunchKernelGGL(k_apply_h<T>, dim3(grid), dim3(kBlock), 0, h->stream, psi, lm, namps,
                       DevTerms{dterms, (int)nterms}, part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid, slots);
    std::vector<char> is_param(gates.size(), 0);
    for (size_t gi = gates.size(); gi-- > 0;) {
        const GateDesc &g = gates[gi];
        int gen = gen_of(g.kind);
        if (gen != GEN_NONE) {
            is_param[gi] = 1;
            uint64_t tb, cb = 0, ctrl = 0;
            if (gen == GEN_NPX || gen == GEN_NPY) {
                tb = 1ull << g.q0;
                cb = 1ull << g.q1;
            } else if (is_gate2(g.kind)) {
                tb = 1ull << g.q1;
                ctrl = 1ull << g.q0;
            } else {
                tb = 1ull << g.q0;
            }
            hipLaunchKernelGGL(k_grad_partial<T>, dim3(grid), dim3(kBlock), 0, h->stream, psi,
                               lm, namps, gen, tb, cb, ctrl, part);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid,
                               slots + 1 + gi);
        }
        GateDesc inv = g;
        inv.s = -g.s;  // every non-parametrised kind here is an involution
        launch_gate(h, inv, work.ptr);
        launch_gate(h, inv, lam.ptr);
    }
    QH_CHECK(hipGetLastError());
    std::vector<double> host(nslots);
    QH_CHECK(hipMemcpyAsync(host.data(), slots, nslots * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    *energy = host[0];
    if (grads)
        for (size_t gi = 0; gi < gates.size(); gi++) grads[gi] = is_param[gi] ? host[1 + gi] : 0.0;
    work.release();
    lam.release();
    scratch.release();
}

template <typename T>
double expval_impl(qslhip_state *h, const qslhip_pauli *terms, uint64_t nterms)
{
    const uint64_t namps = h->namps;
    DevBuf lam, scratch;
    lam.ensure(namps * sizeof(cplx<T>));
    scratch.ensure((kRedBlocks + 8) * sizeof(double) + nterms * sizeof(qslhip_pauli));
    double *part = static_cast<double *>(scratch.ptr);
    double *res = part + kRedBlocks;
    qslhip_pauli *dterms = reinterpret_cast<qslhip_pauli *>(res + 8);
    QH_CHECK(hipMemcpyAsync(dterms, terms, nterms * sizeof(qslhip_pauli),
                            hipMemcpyHostToDevice, h->stream));
    unsigned grid = grid_for(namps, kBlock * 4);
    if (grid > (unsigned)kRedBlocks) grid = kRedBlocks;
    hipLaunchKernelGGL(k_apply_h<T>, dim3(grid), dim3(kBlock), 0, h->stream,
                       static_cast<const cplx<T> *>(h->amps), static_cast<cplx<T> *>(lam.ptr),
                       namps, DevTerms{dterms, (int)nterms}, part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid, res);
    QH_CHECK(hipGetLastError());
    double v = 0;
    QH_CHECK(hipMemcpyAsync(&v, res, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    lam.release();
    scratch.release();
    return v;
}

}  // namespace

void adjoint_plain(qslhip_state *h, const std::vector<GateDesc> &gates,
                   const qslhip_pauli *terms, uint64_t nterms, double *energy, double *grads)
{
    if (h->prec == 8) plain_impl<double>(h, gates, terms, nterms, energy, grads);
    else plain_impl<float>(h, gates, terms, nterms, energy, grads);
}

double expval(qslhip_state *h, const qslhip_pauli *terms, uint64_t nterms)
{
    return h->prec == 8 ? expval_impl<double>(h, terms, nterms)
                        : expval_impl<float>(h, terms, nterms);
}

// ============================ batch path ======================================

namespace {

constexpr int kBBlock = 1024;          // threads per circuit workgroup
constexpr int kBWaves = kBBlock / 64;
constexpr int kLdsQubits = 12;         // whole state in LDS up to 12 qubits
constexpr int kTileQubits = 12;        // tile passes of 2^12 amplitudes
constexpr int kMaxPassOps = 64;        // accumulator slots per pass
constexpr int kLowRun = 4;             // lowest tile qubits kept contiguous

struct BCirc {
    uint32_t n;
    uint32_t pass_first, npass;
    uint32_t op_first, nop;  // global op range (angles / grads index)
    uint32_t term_first, nterm;
    uint32_t lds;            // 1: whole state in LDS
};

struct BPass {
    uint64_t tile_mask;
    uint32_t top_first, top_count;  // into the pass-op table
    uint32_t crun, pad;
    uint32_t seg_first, seg_count;  // into the segment table
};

struct BOp {
    int32_t kind, pa, pb;
    uint32_t q0, q1;
    uint32_t gop;  // global op index
};

struct BDev {
    const BCirc *circ;
    const BPass *pass;
    const BOp *ops;
    const uint32_t *segs;
    const qslhip_pauli *terms;
    const double *angles;
    const uint32_t *order;
    uint32_t ncirc;
    uint32_t *queue;
    double *energy;
    double *grads;
    cplx<float> *slots;  // 2 * slot_amps per workgroup
    uint64_t slot_amps;
};

__device__ __forceinline__ TileOp make_tile_op(const BOp &b, double angle, bool inverse)
{
    TileOp t;
    t.kind = b.kind;
    t.pa = b.pa;
    t.pb = b.pb;
    t.q0 = b.q0;
    t.q1 = b.q1;
    t.pad = 0;
    float c = 0.f, s = 0.f;
    switch (b.kind) {
    case QSLHIP_PHASE: case QSLHIP_CPHASE:
        sincosf((float)angle, &s, &c);
        break;
    case QSLHIP_ROT_X: case QSLHIP_ROT_Y: case QSLHIP_ROT_Z:
    case QSLHIP_CROT_X: case QSLHIP_CROT_Y: case QSLHIP_CROT_Z:
    case QSLHIP_NP_ROT_X: case QSLHIP_NP_ROT_Y:
        sincosf((float)angle * 0.5f, &s, &c);
        break;
    default:
        break;
    }
    t.c = c;
    t.s = inverse ? -s : s;
    return t;
}

__device__ __forceinline__ double block_reduce(double v, double *sh)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBWaves; w++) r += sh[w];
    __syncthreads();
    return r;
}

// ---- segments: runs of ops applied in registers -----------------------------
// A pass's ops are cut into segments whose non-diagonal targets lie on at
// most two tile positions s0 < s1.  Thread t holds the four amplitudes
// base(t) | {0, 1<<s0, 1<<s1, both} of the tile and applies every op of the
// segment to them with the arithmetic of tile_op; diagonal factors and
// controls are read from each element's own index.  One barrier per segment
// instead of one per op.  Segment word: s0 | s1 << 8 | first op << 16 |
// op count << 24.

struct Block {
    cplx<float> v[4];
    uint32_t j[4];
};

// An op decoded once per pass (stage_ops): pair class, body and generator
// looked up, cos / sin in FP32, so the per-block loop only branches on them.
struct BlkOp {
    int32_t cls, body, gen, kind;
    int32_t pa, pb;
    uint32_t q0, q1;
    float c, s;
};

__device__ __forceinline__ void block_index(Block &b, uint32_t t, int s0, int s1)
{
    const uint32_t lo0 = t & ((1u << s0) - 1);
    const uint32_t x = ((t ^ lo0) << 1) | lo0;
    const uint32_t lo1 = x & ((1u << s1) - 1);
    const uint32_t base = ((x ^ lo1) << 1) | lo1;
    b.j[0] = base;
    b.j[1] = base | (1u << s0);
    b.j[2] = base | (1u << s1);
    b.j[3] = base | (1u << s0) | (1u << s1);
}

__device__ __forceinline__ bool jbit(uint32_t j, int p) { return (j >> p) & 1u; }

// Apply one op to a block (same element updates as tile_op).
__device__ __forceinline__ void block_op(Block &b, int s0, const BlkOp &op, uint64_t gbase)
{
    const int cls = op.cls;
    const int body = op.body;
    const float c = op.c, s = op.s;
    const int pa = op.pa, pb = op.pb;
    auto gbit = [&](uint32_t q) -> bool { return (gbase >> q) & 1ull; };
    cplx<float> *v = b.v;
    switch (cls) {
    case PC_PAIR1:
        if (pa >= 0) {
            if (pa == s0) {
                pair_update<float>(body, v[0], v[1], c, s);
                pair_update<float>(body, v[2], v[3], c, s);
            } else {
                pair_update<float>(body, v[0], v[2], c, s);
                pair_update<float>(body, v[1], v[3], c, s);
            }
        } else {
            const bool one = gbit(op.q0);
#pragma unroll
            for (int e = 0; e < 4; e++) rz_half<float>(one, v[e], c, s);
        }
        break;
    case PC_ONE:
        if (pa >= 0) {
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (jbit(b.j[e], pa)) one_update<float>(body, v[e], c, s);
        } else if (gbit(op.q0)) {
#pragma unroll
            for (int e = 0; e < 4; e++) one_update<float>(body, v[e], c, s);
        }
        break;
    case PC_CTRL:
        if (pa < 0 && !gbit(op.q0)) break;
        if (pb >= 0) {
            // pairs along pb; the control bit is the same in both members
            if (pb == s0) {
                if (pa < 0 || jbit(b.j[0], pa)) pair_update<float>(body, v[0], v[1], c, s);
                if (pa < 0 || jbit(b.j[2], pa)) pair_update<float>(body, v[2], v[3], c, s);
            } else {
                if (pa < 0 || jbit(b.j[0], pa)) pair_update<float>(body, v[0], v[2], c, s);
                if (pa < 0 || jbit(b.j[1], pa)) pair_update<float>(body, v[1], v[3], c, s);
            }
        } else {
            const bool one = gbit(op.q1);
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (pa < 0 || jbit(b.j[e], pa)) rz_half<float>(one, v[e], c, s);
        }
        break;
    case PC_DIAG2:
        if (pa < 0 && !gbit(op.q0)) break;
        if (pb < 0 && !gbit(op.q1)) break;
#pragma unroll
        for (int e = 0; e < 4; e++)
            if ((pa < 0 || jbit(b.j[e], pa)) && (pb < 0 || jbit(b.j[e], pb)))
                one_update<float>(body, v[e], c, s);
        break;
    case PC_SWAPLIKE:
    case PC_FSWAP:
        // element with the pa bit set (pb clear) first, as tile_op's j0
        if (pa == s0) pair_update<float>(body, v[1], v[2], c, s);
        else          pair_update<float>(body, v[2], v[1], c, s);
        if (cls == PC_FSWAP) one_update<float>(QSLHIP_PAULI_Z, v[3], c, s);
        break;
    default:
        break;
    }
}

// Sum of gen_term over a block's elements; the partner of element e is
// element e ^ LM (a compile-time register, no indexed access).  virt: a
// diagonal generator on a qubit outside the tile, evaluated at the virtual
// index vj with target bit 1.
template <int LM>
__device__ __forceinline__ double block_terms(const Block &p, const Block &l, int gen, uint64_t vj,
                                              uint64_t tb, uint64_t cb, uint64_t ctrl_tile,
                                              bool virt)
{
    double v = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint64_t j = p.j[e];
        const bool on = (j & ctrl_tile) == ctrl_tile;
        const cplx<float> partner = p.v[e ^ LM];
        v += gen_term<float>(gen, virt ? vj : j, l.v[e], tb, cb, on,
                             [&](uint64_t) { return partner; });
    }
    return v;
}

// Gradient partial of one op over a block (psi in p, lambda in l), summed
// per wave into acc[slot][wave].  Targets of non-diagonal generators are
// block positions; a diagonal generator on a qubit outside the tile takes
// its bit from the group's base index.
__device__ __forceinline__ void block_grad(const Block &p, const Block &l, int s0, int s1,
                                           uint64_t gbase, const BlkOp &op, double *acc, int slot)
{
    const int gen = op.gen;
    uint64_t tb, cb = 0, ctrl_tile = 0;
    bool ctrl_global_on = true;
    int tpos = -1;
    if (gen == GEN_NPX || gen == GEN_NPY) {
        tb = 1ull << op.pa;
        cb = 1ull << op.pb;
    } else if (is_gate2_dev(op.kind)) {
        tpos = op.pb;
        if (op.pa >= 0) ctrl_tile = 1ull << op.pa;
        else ctrl_global_on = (gbase >> op.q0) & 1ull;
        tb = (op.pb >= 0) ? (1ull << op.pb) : 0;
    } else {
        tpos = op.pa;
        tb = (op.pa >= 0) ? (1ull << op.pa) : 0;
    }
    double v = 0;
    if (ctrl_global_on) {
        const bool tglobal = (gen == GEN_Z || gen == GEN_P1) && tpos < 0;
        const uint64_t gq = is_gate2_dev(op.kind) ? op.q1 : op.q0;
        const uint64_t vj = tglobal ? ((gbase >> gq) & 1ull) : 0;
        // gen_term reads psi at the element's index with tb (and cb) flipped:
        // a block element whose local index differs by the uniform mask lm
        const uint64_t flip = (gen == GEN_X || gen == GEN_Y) ? tb
                            : (gen == GEN_NPX || gen == GEN_NPY) ? (tb | cb) : 0;
        const int lm = (int)((flip >> s0) & 1ull) | (int)(((flip >> s1) & 1ull) << 1);
        if (tglobal) v = block_terms<0>(p, l, gen, vj, 1ull, 0, ctrl_tile, true);
        else if (lm == 0) v = block_terms<0>(p, l, gen, 0, tb, cb, ctrl_tile, false);
        else if (lm == 1) v = block_terms<1>(p, l, gen, 0, tb, cb, ctrl_tile, false);
        else if (lm == 2) v = block_terms<2>(p, l, gen, 0, tb, cb, ctrl_tile, false);
        else v = block_terms<3>(p, l, gen, 0, tb, cb, ctrl_tile, false);
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) acc[slot * kBWaves + (threadIdx.x >> 6)] += v;
}

// Element j of a tile of nt amplitudes, zero past its end (blocks of a
// state of one qubit).  Selects on the value, not on the address.
__device__ __forceinline__ cplx<float> tile_at(const cplx<float> *tile, uint32_t j, uint32_t nt)
{
    cplx<float> x = tile[j < nt ? j : nt - 1];
    if (j >= nt) {
        x.re = 0.f;
        x.im = 0.f;
    }
    return x;
}

// Forward segments over a tile of nt amplitudes in LDS; ends with a barrier.
__device__ __forceinline__ void run_segments_fwd(cplx<float> *tile, uint32_t nt, uint64_t gbase,
                                                 const uint32_t *sseg, uint32_t nseg,
                                                 const BlkOp *sop)
{
    for (uint32_t sg = 0; sg < nseg; sg++) {
        const uint32_t w = sseg[sg];
        const int s0 = (int)(w & 255u), s1 = (int)((w >> 8) & 255u);
        const uint32_t first = (w >> 16) & 255u, count = w >> 24;
        Block b;
        block_index(b, threadIdx.x, s0, s1);
#pragma unroll
        for (int e = 0; e < 4; e++) b.v[e] = tile_at(tile, b.j[e], nt);
        for (uint32_t k = first; k < first + count; k++) block_op(b, s0, sop[k], gbase);
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (b.j[e] < nt) tile[b.j[e]] = b.v[e];
        __syncthreads();
    }
}

// Backward segments (last op first) over psi and lambda tiles: gradient
// partials, then the inverse ops (sop holds them).  Ends with a barrier.
__device__ __forceinline__ void run_segments_bwd(cplx<float> *psi, cplx<float> *lam, uint32_t nt,
                                                 uint64_t gbase, const uint32_t *sseg,
                                                 uint32_t nseg, const BlkOp *sop, double *acc)
{
    for (uint32_t sg = nseg; sg-- > 0;) {
        const uint32_t w = sseg[sg];
        const int s0 = (int)(w & 255u), s1 = (int)((w >> 8) & 255u);
        const uint32_t first = (w >> 16) & 255u, count = w >> 24;
        Block p, l;
        block_index(p, threadIdx.x, s0, s1);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            l.j[e] = p.j[e];
            p.v[e] = tile_at(psi, p.j[e], nt);
            l.v[e] = tile_at(lam, p.j[e], nt);
        }
        for (uint32_t k = first + count; k-- > first;) {
            if (sop[k].gen != GEN_NONE) block_grad(p, l, s0, s1, gbase, sop[k], acc, (int)k);
            block_op(p, s0, sop[k], gbase);
            block_op(l, s0, sop[k], gbase);
        }
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (p.j[e] < nt) {
                psi[p.j[e]] = p.v[e];
                lam[p.j[e]] = l.v[e];
            }
        __syncthreads();
    }
}

// lambda = H psi over the whole state (global slot or LDS), returns the
// workgroup's energy sum (valid in thread 0).
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

// One group of a tile pass over the HBM slot, forward: the first segment
// loads its blocks straight from the slot (or writes |0...0> on the first
// pass), the last one stores them straight back, and only the segments in
// between go through the LDS tile.  Group element j lives at slot index
// gbase | (j & lowm) | hoff[j >> crun].  Ends with a barrier.
__device__ __forceinline__ void group_fwd(cplx<float> *gpsi, cplx<float> *tile,
                                          const uint64_t *hoff, uint64_t gbase, uint32_t lowm,
                                          uint32_t crun, bool init, const uint32_t *sseg,
                                          uint32_t nseg, const BlkOp *sop)
{
    for (uint32_t sg = 0; sg < nseg; sg++) {
        const uint32_t w = sseg[sg];
        const int s0 = (int)(w & 255u), s1 = (int)((w >> 8) & 255u);
        const uint32_t first = (w >> 16) & 255u, count = w >> 24;
        Block b;
        block_index(b, threadIdx.x, s0, s1);
        if (sg == 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint64_t idx = gbase | (b.j[e] & lowm) | hoff[b.j[e] >> crun];
                if (init) {
                    b.v[e].re = (idx == 0) ? 1.f : 0.f;
                    b.v[e].im = 0.f;
                } else {
                    b.v[e] = gpsi[idx];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) b.v[e] = tile[b.j[e]];
        }
        for (uint32_t k = first; k < first + count; k++) block_op(b, s0, sop[k], gbase);
        if (sg + 1 == nseg) {
#pragma unroll
            for (int e = 0; e < 4; e++) gpsi[gbase | (b.j[e] & lowm) | hoff[b.j[e] >> crun]] = b.v[e];
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) tile[b.j[e]] = b.v[e];
        }
        __syncthreads();
    }
}

// The same group, backward (last segment first): psi and lambda from the
// slot into the first segment processed, back to the slot from the last
// one when `store` (the final backward pass keeps nothing).
__device__ __forceinline__ void group_bwd(cplx<float> *gpsi, cplx<float> *glam,
                                          cplx<float> *psi, cplx<float> *lam,
                                          const uint64_t *hoff, uint64_t gbase, uint32_t lowm,
                                          uint32_t crun, bool store, const uint32_t *sseg,
                                          uint32_t nseg, const BlkOp *sop, double *acc)
{
    for (uint32_t sg = nseg; sg-- > 0;) {
        const uint32_t w = sseg[sg];
        const int s0 = (int)(w & 255u), s1 = (int)((w >> 8) & 255u);
        const uint32_t first = (w >> 16) & 255u, count = w >> 24;
        Block p, l;
        block_index(p, threadIdx.x, s0, s1);
#pragma unroll
        for (int e = 0; e < 4; e++) l.j[e] = p.j[e];
        if (sg + 1 == nseg) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint64_t idx = gbase | (p.j[e] & lowm) | hoff[p.j[e] >> crun];
                p.v[e] = gpsi[idx];
                l.v[e] = glam[idx];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                p.v[e] = psi[p.j[e]];
                l.v[e] = lam[p.j[e]];
            }
        }
        for (uint32_t k = first + count; k-- > first;) {
            if (sop[k].gen != GEN_NONE) block_grad(p, l, s0, s1, gbase, sop[k], acc, (int)k);
            block_op(p, s0, sop[k], gbase);
            block_op(l, s0, sop[k], gbase);
        }
        if (sg == 0) {
            if (store) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint64_t idx = gbase | (p.j[e] & lowm) | hoff[p.j[e] >> crun];
                    gpsi[idx] = p.v[e];
                    glam[idx] = l.v[e];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                psi[p.j[e]] = p.v[e];
                lam[p.j[e]] = l.v[e];
            }
        }
        __syncthreads();
    }
}

// Stage one pass in LDS: its ops decoded (angles turned into cos/sin once
// per pass, not once per group; inverse for the backward sweep) and its
// segment words.  Ends with a barrier.
__device__ __forceinline__ void stage_ops(const BDev &d, const BPass &P, bool inverse,
                                          BlkOp *sop, uint32_t *sseg)
{
    for (uint32_t k = threadIdx.x; k < P.top_count; k += kBBlock) {
        const BOp o = d.ops[P.top_first + k];
        const TileOp t = make_tile_op(o, d.angles[o.gop], inverse);
        BlkOp b;
        b.cls = pair_class(t.kind);
        b.body = body_of(t.kind);
        b.gen = gen_of(t.kind);
        b.kind = t.kind;
        b.pa = t.pa;
        b.pb = t.pb;
        b.q0 = t.q0;
        b.q1 = t.q1;
        b.c = (float)t.c;  // computed in FP32 by make_tile_op: exact
        b.s = (float)t.s;
        sop[k] = b;
    }
    for (uint32_t k = threadIdx.x; k < P.seg_count; k += kBBlock) sseg[k] = d.segs[P.seg_first + k];
    __syncthreads();
}

// At most 64 VGPRs (8 waves per SIMD): two 1024-thread workgroups per CU,
// which the 80 KiB of LDS per workgroup allows.  The workgroups spend much
// of their time at barriers and on dependent LDS / scalar work, and a
// second one per CU runs 1.5x more circuits per second than the
// 82-VGPR build, with the few values it spills in cold paths.
__global__ __launch_bounds__(kBBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch(BDev d)
{
    // LDS: psi and lambda tiles of 2^12 FP32 amplitudes (2 x 32 KiB), the
    // offsets of the scattered tile qubits, and the gradient accumulators.
    __shared__ cplx<float> spsi[1 << kTileQubits];
    __shared__ cplx<float> slam[1 << kTileQubits];
    __shared__ uint64_t hoff[1 << (kTileQubits - kLowRun)];
    __shared__ double acc[kMaxPassOps * kBWaves];
    __shared__ double sh[kBWaves];
    __shared__ BlkOp sop[kMaxPassOps];
    __shared__ uint32_t sseg[kMaxPassOps];
    __shared__ uint32_t s_next;

    cplx<float> *gpsi = d.slots + (uint64_t)blockIdx.x * 2 * d.slot_amps;
    cplx<float> *glam = gpsi + d.slot_amps;

    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(d.queue, 1u);
        __syncthreads();
        const uint32_t qi = s_next;
        __syncthreads();
        if (qi >= d.ncirc) break;
        const BCirc c = d.circ[d.order[qi]];
        const uint64_t namps = 1ull << c.n;
        const uint64_t full = namps - 1;
        for (uint32_t g = threadIdx.x; g < c.nop; g += kBBlock) d.grads[c.op_first + g] = 0.0;

        if (c.lds) {
            // ---- whole state in LDS ---------------------------------------
            for (uint32_t j = threadIdx.x; j < namps; j += kBBlock) {
                cplx<float> z;
                z.re = (j == 0) ? 1.f : 0.f;
                z.im = 0.f;
                spsi[j] = z;
            }
            __syncthreads();
            for (uint32_t p = 0; p < c.npass; p++) {
                const BPass P = d.pass[c.pass_first + p];
                stage_ops(d, P, false, sop, sseg);
                run_segments_fwd(spsi, (uint32_t)namps, 0, sseg, P.seg_count, sop);
            }
            double e = apply_h(spsi, slam, namps, d.terms + c.term_first, (int)c.nterm, sh);
            if (threadIdx.x == 0) d.energy[d.order[qi]] = e;
            for (uint32_t p = c.npass; p-- > 0;) {
                const BPass P = d.pass[c.pass_first + p];
                for (uint32_t i = threadIdx.x; i < kMaxPassOps * kBWaves; i += kBBlock) acc[i] = 0;
                stage_ops(d, P, true, sop, sseg);
                run_segments_bwd(spsi, slam, (uint32_t)namps, 0, sseg, P.seg_count, sop, acc);
                for (uint32_t k = threadIdx.x; k < P.top_count; k += kBBlock) {
                    const BOp o = d.ops[P.top_first + k];
                    if (gen_of(o.kind) != GEN_NONE) {
                        double s = 0;
                        for (int w = 0; w < kBWaves; w++) s += acc[k * kBWaves + w];
                        d.grads[o.gop] += s;
                    }
                }
                __syncthreads();
            }
            continue;
        }

        // ---- tile passes over the workgroup's HBM slot -----------------------
        const int K = kTileQubits;
        static_assert(kBBlock * 4 == 1 << kTileQubits, "one 4-amplitude block per thread");
        const uint64_t ngroups = namps >> K;
        // qubits some tile so far held: from |0...0>, an amplitude whose
        // index has a 1 on any other qubit is still zero
        uint64_t support = 0;
        for (uint32_t p = 0; p < c.npass; p++) {
            const BPass P = d.pass[c.pass_first + p];
            const uint64_t hi_mask = P.tile_mask & ~((1ull << P.crun) - 1);
            const uint32_t lowm = (1u << P.crun) - 1;
            support |= P.tile_mask;
            for (int m = threadIdx.x; m < (1 << (K - P.crun)); m += kBBlock)
                hoff[m] = deposit((uint64_t)m, hi_mask);
            stage_ops(d, P, false, sop, sseg);
            for (uint64_t g = 0; g < ngroups; g++) {
                const uint64_t gbase = deposit(g, full & ~P.tile_mask);
                if (gbase & ~support) {
                    // A zero group stays zero.  Pass 0 stores its zeros (the
                    // slot still holds the previous circuit); a later pass
                    // stores nothing: each of its indices has a 1 outside
                    // every earlier pass's support too, so it lay in a zero
                    // group of pass 0 and no pass since has written it.
                    if (p == 0) {
                        const uint32_t t = threadIdx.x;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            const uint32_t j = t + (uint32_t)e * kBBlock;
                            cplx<float> z;
                            z.re = 0.f;
                            z.im = 0.f;
                            gpsi[gbase | (j & lowm) | hoff[j >> P.crun]] = z;
                        }
                    }
                    continue;
                }
                // the first pass starts from |0...0>
                group_fwd(gpsi, spsi, hoff, gbase, lowm, P.crun, p == 0, sseg, P.seg_count, sop);
            }
            // zero groups end without a barrier: every thread's stores (and
            // its reads of hoff) are done before the next pass restages
            // hoff and the ops, and before lambda = H psi reads the slot
            __syncthreads();
        }
        if (c.npass == 0) {  // no ops: the state stays |0...0>
            for (uint64_t i = threadIdx.x; i < namps; i += kBBlock) {
                cplx<float> z;
                z.re = (i == 0) ? 1.f : 0.f;
                z.im = 0.f;
                gpsi[i] = z;
            }
            __syncthreads();
        }
        double e = apply_h(gpsi, glam, namps, d.terms + c.term_first, (int)c.nterm, sh);
        if (threadIdx.x == 0) d.energy[d.order[qi]] = e;
        __syncthreads();
        for (uint32_t p = c.npass; p-- > 0;) {
            const BPass P = d.pass[c.pass_first + p];
            const uint64_t hi_mask = P.tile_mask & ~((1ull << P.crun) - 1);
            const uint32_t lowm = (1u << P.crun) - 1;
            for (int m = threadIdx.x; m < (1 << (K - P.crun)); m += kBBlock)
                hoff[m] = deposit((uint64_t)m, hi_mask);
            for (uint32_t i = threadIdx.x; i < kMaxPassOps * kBWaves; i += kBBlock) acc[i] = 0;
            stage_ops(d, P, true, sop, sseg);
            for (uint64_t g = 0; g < ngroups; g++) {
                const uint64_t gbase = deposit(g, full & ~P.tile_mask);
                group_bwd(gpsi, glam, spsi, slam, hoff, gbase, lowm, P.crun, p > 0, sseg,
                          P.seg_count, sop, acc);
            }
            for (uint32_t k = threadIdx.x; k < P.top_count; k += kBBlock) {
                const BOp o = d.ops[P.top_first + k];
                if (gen_of(o.kind) != GEN_NONE) {
                    double s = 0;
                    for (int w = 0; w < kBWaves; w++) s += acc[k * kBWaves + w];
                    d.grads[o.gop] += s;
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
    DevBuf circ, pass, ops, segs, terms, angles, order, queue, energy, grads, slots;
    PinnedBuf pin_angles, pin_out;
    unsigned grid = 0;
    BDev dev{};
};

// Cut a pass's ops into segments (see run_segments_fwd): a new segment
// starts when the non-diagonal targets would span more than two positions;
// a segment with fewer is padded with the lowest free positions.
static void add_segments(std::vector<uint32_t> &segs, const BOp *ops, uint32_t count)
{
    auto need = [](const BOp &o) -> uint32_t {
        switch (pair_class(o.kind)) {
        case PC_PAIR1: return o.pa >= 0 ? 1u << o.pa : 0u;
        case PC_CTRL: return o.pb >= 0 ? 1u << o.pb : 0u;
        case PC_SWAPLIKE:
        case PC_FSWAP:
            QH_REQUIRE(o.pa >= 0 && o.pb >= 0, QSLHIP_EINVAL, "swap-like op outside its tile");
            return (1u << o.pa) | (1u << o.pb);
        default: return 0u;
        }
    };
    uint32_t first = 0, req = 0;
    auto flush = [&](uint32_t end) {
        if (end == first) return;
        uint32_t m = req;
        for (unsigned p = 0; __builtin_popcount(m) < 2; p++) m |= 1u << p;
        const uint32_t s0 = (uint32_t)__builtin_ctz(m), s1 = 31u - (uint32_t)__builtin_clz(m);
        segs.push_back(s0 | (s1 << 8) | (first << 16) | ((end - first) << 24));
        first = end;
    };
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t nd = need(ops[i]);
        if (__builtin_popcount(req | nd) > 2) {
            flush(i);
            req = 0;
        }
        req |= nd;
    }
    flush(count);
}

static void batch_plan(Batch &b, const uint32_t *nq, const uint64_t *op_off,
                       const qslhip_op *ops, const uint64_t *term_off,
                       const qslhip_pauli *terms)
{
    std::vector<BCirc> circ(b.ncirc);
    std::vector<BPass> passes;
    std::vector<BOp> bops;
    std::vector<uint32_t> segs;
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
        }
        // ops first..first+count-1 of the applied order; ord maps that order
        // to the circuit's (null: program order)
        auto add_pass = [&](uint64_t mask, unsigned crun, size_t first, size_t count,
                            const std::vector<TileOp> *tops, const uint32_t *ord) {
            BPass P;
            P.tile_mask = mask;
            P.crun = crun;
            P.pad = 0;
            P.top_first = (uint32_t)bops.size();
            P.top_count = (uint32_t)count;
            for (size_t i = 0; i < count; i++) {
                BOp o;
                const size_t gi = ord ? ord[first + i] : first + i;
                const GateDesc &g = gates[gi];
                o.kind = g.kind;
                o.q0 = g.q0;
                o.q1 = g.q1;
                o.gop = C.op_first + (uint32_t)gi;
                if (tops) {
                    QH_REQUIRE((*tops)[i].kind == g.kind, QSLHIP_EINVAL,
                               "batch plan: tile op out of step with its gate");
                    o.pa = (*tops)[i].pa;
                    o.pb = (*tops)[i].pb;
                } else {  // whole state in LDS: tile index == basis index
                    o.pa = (int)g.q0;
                    o.pb = is_gate2(g.kind) ? (int)g.q1 : -1;
                }
                bops.push_back(o);
            }
            P.seg_first = (uint32_t)segs.size();
            add_segments(segs, bops.data() + P.top_first, P.top_count);
            P.seg_count = (uint32_t)segs.size() - P.seg_first;
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
                    std::vector<TileOp> slice(to.begin() + P.op_first + f,
                                              to.begin() + P.op_first + f + cnt);
                    add_pass(P.tile_mask, crun, P.op_first + f, cnt, &slice, ord.data());
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
    up(b.ops, bops.data(), bops.size() * sizeof(BOp));
    up(b.segs, segs.data(), segs.size() * sizeof(uint32_t));
    up(b.terms, terms, term_off[b.ncirc] * sizeof(qslhip_pauli));
    up(b.order, order.data(), order.size() * sizeof(uint32_t));
    b.angles.ensure(b.nops * sizeof(double) + 8);
    b.queue.ensure(64);
    b.energy.ensure(b.ncirc * sizeof(double) + 8);
    b.grads.ensure(b.nops * sizeof(double) + 8);
    b.slots.ensure((size_t)b.grid * 2 * slot_amps * sizeof(cplx<float>));
    b.pin_angles.ensure(b.nops * sizeof(double) + 8);
    b.pin_out.ensure((b.ncirc + b.nops) * sizeof(double) + 8);

    BDev &d = b.dev;
    d.circ = static_cast<const BCirc *>(b.circ.ptr);
    d.pass = static_cast<const BPass *>(b.pass.ptr);
    d.ops = static_cast<const BOp *>(b.ops.ptr);
    d.segs = static_cast<const uint32_t *>(b.segs.ptr);
    d.terms = static_cast<const qslhip_pauli *>(b.terms.ptr);
    d.angles = static_cast<const double *>(b.angles.ptr);
    d.order = static_cast<const uint32_t *>(b.order.ptr);
    d.ncirc = b.ncirc;
    d.queue = static_cast<uint32_t *>(b.queue.ptr);
    d.energy = static_cast<double *>(b.energy.ptr);
    d.grads = static_cast<double *>(b.grads.ptr);
    d.slots = static_cast<cplx<float> *>(b.slots.ptr);
    d.slot_amps = slot_amps;

    // Capture once: angles in, queue reset (a kernel, not a memset node),
    // the persistent kernel, results out.
    QH_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    QH_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    QH_CHECK(hipMemcpyAsync(b.angles.ptr, b.pin_angles.ptr, b.nops * sizeof(double),
                            hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, b.stream, d.queue);
    hipLaunchKernelGGL(k_batch, dim3(b.grid), dim3(kBBlock), 0, b.stream, d);
    double *po = static_cast<double *>(b.pin_out.ptr);
    QH_CHECK(hipMemcpyAsync(po, b.energy.ptr, b.ncirc * sizeof(double),
                            hipMemcpyDeviceToHost, b.stream));
    QH_CHECK(hipMemcpyAsync(po + b.ncirc, b.grads.ptr, b.nops * sizeof(double),
                            hipMemcpyDeviceToHost, b.stream));
    QH_CHECK(hipStreamEndCapture(b.stream, &b.graph));
    QH_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

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
        for (DevBuf *d : {&b.circ, &b.pass, &b.ops, &b.segs, &b.terms, &b.angles, &b.order,
                          &b.queue, &b.energy, &b.grads, &b.slots})
            d->release();
        b.pin_angles.release();
        b.pin_out.release();
        if (b.stream) (void)hipStreamDestroy(b.stream);
        delete bt;
    });
}

}  // extern "C"
