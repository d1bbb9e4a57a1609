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

#include "tile_ops.h"

namespace qslhip {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxHiBits = 8;

// Group <-> LDS tile copies, eight elements per thread at a time: all
// eight loads are issued before the first wait (a loop of single
// load-then-store keeps one load in flight per thread).
template <bool S, typename T>
__device__ __forceinline__ void tile_load(cplx<T> *tile, const cplx<T> *st, uint32_t nt,
                                          uint64_t gbase, uint32_t lowm, int crun,
                                          const uint64_t *hoff)
{
    for (uint32_t e0 = threadIdx.x; e0 < nt; e0 += 8 * kBlock) {
        cplx<T> v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            // past the tile (K < 11): reload the first element, not stored
            const uint32_t e = e0 + k * kBlock < nt ? e0 + k * kBlock : e0;
            v[k] = amp_ld<S>(st + (gbase | (e & lowm) | hoff[e >> crun]));
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t e = e0 + k * kBlock;
            if (e < nt) tile[e] = v[k];
        }
    }
}

template <bool S, typename T>
__device__ __forceinline__ void tile_store(const cplx<T> *tile, cplx<T> *st, uint32_t nt,
                                           uint64_t gbase, uint32_t lowm, int crun,
                                           const uint64_t *hoff)
{
    for (uint32_t e0 = threadIdx.x; e0 < nt; e0 += 8 * kBlock) {
        cplx<T> v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t e = e0 + k * kBlock < nt ? e0 + k * kBlock : e0;
            v[k] = tile[e];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t e = e0 + k * kBlock;
            if (e < nt) amp_st<S>(st + (gbase | (e & lowm) | hoff[e >> crun]), v[k]);
        }
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
    tile_load<false>(tile, st, nt, gbase, lowm, crun, hoff);
    __syncthreads();

    for (int i = 0; i < nops; i++) {
        TileOp op = ops[i];
        tile_op<T>(tile, K, gbase, op);
        __syncthreads();
    }

    tile_store<false>(tile, st, nt, gbase, lowm, crun, hoff);
}

// The same pass with one op passed by value (the per-gate path for low
// targets): no op table to stage.
template <typename T, bool S>
__global__ __launch_bounds__(kBlock) void k_tile1(cplx<T> *__restrict__ st, unsigned n,
                                                  int K, int crun, uint64_t tile_mask,
                                                  TileOp op)
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
    tile_load<S>(tile, st, nt, gbase, lowm, crun, hoff);
    __syncthreads();
    tile_op<T>(tile, K, gbase, op);
    __syncthreads();
    tile_store<S>(tile, st, nt, gbase, lowm, crun, hoff);
}

template <typename T>
void run(qslhip_state *h, const std::vector<TilePass> &passes, const TileOp *dops,
         void *amps, const TileOp *single)
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
        if (single) {
            // whole contiguous runs of at least 128 bytes: streaming accesses
            // for states far larger than the caches
            const bool stream = (int)h->n >= tuning().stream_min_qubits &&
                                (sizeof(cplx<T>) << crun) >= 128;
            if (stream)
                hipLaunchKernelGGL((k_tile1<T, true>), dim3((unsigned)groups), dim3(kBlock),
                                   shmem, h->stream, static_cast<cplx<T> *>(amps), h->n,
                                   p.logt, crun, p.tile_mask, *single);
            else
                hipLaunchKernelGGL((k_tile1<T, false>), dim3((unsigned)groups), dim3(kBlock),
                                   shmem, h->stream, static_cast<cplx<T> *>(amps), h->n,
                                   p.logt, crun, p.tile_mask, *single);
            QH_CHECK(hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL(k_tile<T>, dim3((unsigned)groups), dim3(kBlock), shmem,
                           h->stream, static_cast<cplx<T> *>(amps), h->n,
                           p.logt, crun, p.tile_mask, dops + p.op_first,
                           (int)p.op_count);
        QH_CHECK(hipGetLastError());
    }
}

}  // namespace

bool launch_tile_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                        const std::vector<TileOp> &ops, void *amps, bool exact, uint64_t basis,
                        double *runs)
{
    if (!amps) amps = h->amps;
    QH_REQUIRE(basis == qslhip_state::kNoBasis || amps == h->amps, QSLHIP_EINVAL,
               "a pending basis state belongs to the handle's own vector");
    const Tuning &tu = tuning();
    const int threads = h->prec == 8 ? tu.rthreads_f64 : tu.rthreads_f32;
    bool reg = tu.reg_tile != 0 && h->n >= (unsigned)tu.reg_min_qubits && !ops.empty() &&
               !(passes.size() == 1 && ops.size() == 1);
    for (const TilePass &p : passes) reg = reg && reg_tile_ok(p.logt, threads, h->prec);
    if (reg) {
        // run sums need qubits 0..2 on the lanes of the last pass's layout,
        // and eight groups for the sampler to regroup them (k_subs_from_runs)
        if (runs && ((passes.back().tile_mask & 7) != 7 || h->n < (unsigned)passes.back().logt + 3))
            runs = nullptr;
        return launch_reg_passes(h, passes, ops, amps, threads, tu.rgroups, exact, basis, runs) &&
               runs != nullptr;
    }
    // the other kernels read the state: write the basis state first
    if (basis != qslhip_state::kNoBasis) launch_fill_basis(h, basis);
    if (passes.empty() || ops.empty()) return false;
    if (passes.size() == 1 && ops.size() == 1) {
        if (h->prec == 8) run<double>(h, passes, nullptr, amps, &ops[0]);
        else run<float>(h, passes, nullptr, amps, &ops[0]);
        return false;
    }
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
    if (h->prec == 8) run<double>(h, passes, dops, amps, nullptr);
    else run<float>(h, passes, dops, amps, nullptr);
    return false;
}

}  // namespace qslhip
