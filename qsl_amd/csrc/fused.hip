// fused.hip -- LDS tile passes: a run of gates applied in one HBM round trip.
//
// A pass owns a set of 2^K "tile" qubits (the lowest `crun` qubits, kept
// contiguous for coalescing, plus up to 8 others).  Workgroup g takes the
// 2^K amplitudes that share the values of every non-tile qubit (the group
// base), loads them into LDS with contiguous 16-byte lanes, applies every op
// of the pass in program order, and writes them back.  Each op updates the
// same amplitudes with the same arithmetic as the one-gate kernels
// (gate_math.h), so a pass is bit-identical to running its gates one by one:
//   - non-diagonal targets are tile qubits (the planner guarantees it);
//   - controls and diagonal factors on non-tile qubits are constant over the
//     group, so the whole op is either applied or skipped per workgroup.
// HBM traffic per pass is one read and one write of the state, however many
// gates it holds; LDS carries the per-gate traffic.
#include <cstring>

#include "gate_math.h"

namespace qslhip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxHiBits = 8;

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

template <typename T>
__device__ __forceinline__ void tile_op(cplx<T> *tile, int K, uint64_t gbase,
                                        const TileOp &op)
{
    const int tid = threadIdx.x;
    const int cls = pair_class(op.kind);
    const int body = body_of(op.kind);
    const T c = (T)op.c, s = (T)op.s;
    const int pa = op.pa, pb = op.pb;
    const uint32_t nt = 1u << K;
    // value of a qubit that is outside the tile (uniform over the group)
    auto gbit = [&](uint32_t q) -> bool { return (gbase >> q) & 1ull; };

    switch (cls) {
    case PC_PAIR1:
        if (pa >= 0) {
            for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                uint32_t j0 = (uint32_t)ins0(f, pa), j1 = j0 | (1u << pa);
                cplx<T> a = tile[j0], b = tile[j1];
                pair_update<T>(body, a, b, c, s);
                tile[j0] = a;
                tile[j1] = b;
            }
        } else {  // rotateZ on a non-tile qubit: one half-formula for all
            bool one = gbit(op.q0);
            for (uint32_t j = tid; j < nt; j += kBlock) rz_half<T>(one, tile[j], c, s);
        }
        break;
    case PC_ONE:
        if (pa >= 0) {
            for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                uint32_t j = (uint32_t)ins0(f, pa) | (1u << pa);
                one_update<T>(body, tile[j], c, s);
            }
        } else if (gbit(op.q0)) {
            for (uint32_t j = tid; j < nt; j += kBlock) one_update<T>(body, tile[j], c, s);
        }
        break;
    case PC_CTRL:
        if (pa < 0 && !gbit(op.q0)) break;  // control off for this group
        if (pb >= 0) {
            if (pa >= 0) {
                int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
                for (uint32_t f = tid; f < nt / 4; f += kBlock) {
                    uint32_t base = (uint32_t)ins0(ins0(f, lo), hi);
                    uint32_t j0 = base | (1u << pa), j1 = j0 | (1u << pb);
                    cplx<T> a = tile[j0], b = tile[j1];
                    pair_update<T>(body, a, b, c, s);
                    tile[j0] = a;
                    tile[j1] = b;
                }
            } else {
                for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                    uint32_t j0 = (uint32_t)ins0(f, pb), j1 = j0 | (1u << pb);
                    cplx<T> a = tile[j0], b = tile[j1];
                    pair_update<T>(body, a, b, c, s);
                    tile[j0] = a;
                    tile[j1] = b;
                }
            }
        } else {  // controlRotateZ with its target outside the tile
            bool one = gbit(op.q1);
            if (pa >= 0) {
                for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                    uint32_t j = (uint32_t)ins0(f, pa) | (1u << pa);
                    rz_half<T>(one, tile[j], c, s);
                }
            } else {
                for (uint32_t j = tid; j < nt; j += kBlock) rz_half<T>(one, tile[j], c, s);
            }
        }
        break;
    case PC_DIAG2: {
        if (pa < 0 && !gbit(op.q0)) break;
        if (pb < 0 && !gbit(op.q1)) break;
        if (pa >= 0 && pb >= 0) {
            int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
            for (uint32_t f = tid; f < nt / 4; f += kBlock) {
                uint32_t j = (uint32_t)ins0(ins0(f, lo), hi) | (1u << pa) | (1u << pb);
                one_update<T>(body, tile[j], c, s);
            }
        } else if (pa >= 0 || pb >= 0) {
            int p = pa >= 0 ? pa : pb;
            for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                uint32_t j = (uint32_t)ins0(f, p) | (1u << p);
                one_update<T>(body, tile[j], c, s);
            }
        } else {
            for (uint32_t j = tid; j < nt; j += kBlock) one_update<T>(body, tile[j], c, s);
        }
        break;
    }
    case PC_SWAPLIKE:
    case PC_FSWAP: {
        int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
        for (uint32_t f = tid; f < nt / 4; f += kBlock) {
            uint32_t base = (uint32_t)ins0(ins0(f, lo), hi);
            uint32_t j0 = base | (1u << pa), j1 = base | (1u << pb);
            cplx<T> a = tile[j0], b = tile[j1];
            pair_update<T>(body, a, b, c, s);
            tile[j0] = a;
            tile[j1] = b;
            if (cls == PC_FSWAP) {
                uint32_t j = base | (1u << pa) | (1u << pb);
                one_update<T>(QSLHIP_PAULI_Z, tile[j], c, s);
            }
        }
        break;
    }
    default:
        break;
    }
}

// One workgroup per group of 2^K amplitudes.
//   tile_mask: the K tile qubits; crun: length of its low contiguous run
//   (qubits 0..crun-1); the other K-crun <= 8 tile qubits are addressed
//   through an LDS table of their 2^(K-crun) offsets.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tile(cplx<T> *__restrict__ st,
                                                 unsigned n, int K, int crun,
                                                 uint64_t tile_mask,
                                                 const TileOp *__restrict__ ops,
                                                 int nops)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T> *tile = reinterpret_cast<cplx<T> *>(smem);
    uint64_t *hoff = reinterpret_cast<uint64_t *>(smem + (sizeof(cplx<T>) << K));

    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const uint64_t gbase = deposit((uint64_t)blockIdx.x, full & ~tile_mask);
    const uint64_t hi_mask = tile_mask & ~((1ull << crun) - 1);
    const int hib = K - crun;
    for (int m = threadIdx.x; m < (1 << hib); m += kBlock) hoff[m] = deposit(m, hi_mask);
    __syncthreads();

    const uint32_t nt = 1u << K;
    const uint32_t lowm = (1u << crun) - 1;
    for (uint32_t e = threadIdx.x; e < nt; e += kBlock)
        tile[e] = st[gbase | (e & lowm) | hoff[e >> crun]];
    __syncthreads();

    for (int i = 0; i < nops; i++) {
        TileOp op = ops[i];
        tile_op<T>(tile, K, gbase, op);
        __syncthreads();
    }

    for (uint32_t e = threadIdx.x; e < nt; e += kBlock)
        st[gbase | (e & lowm) | hoff[e >> crun]] = tile[e];
}

template <typename T>
void run(qslhip_state *h, const std::vector<TilePass> &passes, const TileOp *dops)
{
    for (const TilePass &p : passes) {
        int crun = 0;
        while (crun < 64 && ((p.tile_mask >> crun) & 1ull)) crun++;
        if (crun > p.logt) crun = p.logt;
        QH_REQUIRE(p.logt - crun <= kMaxHiBits, QSLHIP_EINVAL,
                   "tile pass with too many scattered tile qubits");
        size_t shmem = (sizeof(cplx<T>) << p.logt) +
                       sizeof(uint64_t) * ((size_t)1 << (p.logt - crun));
        QH_REQUIRE(shmem <= 64 * 1024, QSLHIP_EINVAL, "tile pass needs more than 64 KiB of LDS");
        uint64_t groups = 1ull << (h->n - p.logt);
        hipLaunchKernelGGL(k_tile<T>, dim3((unsigned)groups), dim3(kBlock), shmem,
                           h->stream, static_cast<cplx<T> *>(h->amps), h->n,
                           p.logt, crun, p.tile_mask, dops + p.op_first,
                           (int)p.op_count);
        QH_CHECK(hipGetLastError());
    }
}

}  // namespace

void launch_tile_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                        const std::vector<TileOp> &ops)
{
    if (passes.empty() || ops.empty()) return;
    size_t bytes = ops.size() * sizeof(TileOp);
    // The pinned staging buffer may still be read by the previous copy.
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    h->pin.ensure(bytes);
    h->optab.ensure(bytes);
    std::memcpy(h->pin.ptr, ops.data(), bytes);
    QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, bytes, hipMemcpyHostToDevice,
                            h->stream));
    QH_CHECK(hipEventRecord(h->staged, h->stream));
    const TileOp *dops = static_cast<const TileOp *>(h->optab.ptr);
    if (h->prec == 8) run<double>(h, passes, dops);
    else run<float>(h, passes, dops);
}

}  // namespace qslhip
