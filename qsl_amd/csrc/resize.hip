// resize.hip -- the resizeable backend's moves: a qubit added
// (addQubit, src/qubits/resize/resize.cpp:107-147) and a qubit removed after
// collapse (collapseOut, src/qubits/resize/measure.cpp:62-94), on the
// device, in place where the allocation allows.
//
// With k = 2^targ, a block is k consecutive amplitudes.
//   add:     block m -> block 2m, block 2m+1 <- 0     (data moves up)
//   remove:  block 2m+outcome -> block m, times factor (data moves down)
// Both are pure streaming moves (read 2^n amplitudes, write as many or
// twice as many), so they are HBM-bound.  In place, one launch cannot move
// every block at once: a block's destination is another block's source.
// The blocks are therefore moved in ranges that double, in an order where
// no launch writes a block that a later launch still reads:
//   add:     m in [H/2, H), then [H/4, H/2), ..., [1, 2), then [0, 1)
//            (H = 2^(n-targ) source blocks; the range [h, 2h) writes blocks
//            [2h, 4h), which no later launch reads)
//   remove:  m in [0, 1), then [1, 2), [2, 4), ..., [M/2, M)
//            (M = 2^(n-1-targ) destination blocks; the range [a, 2a) reads
//            blocks [2a, 4a] and writes [a, 2a), which no later launch reads)
// Within a launch the read and written index ranges are disjoint, except
// block 0, which one thread reads and writes at the same index.  Launches
// are ordered on the handle's stream.
#include "gate_math.h"

namespace qslhip {

namespace {

constexpr int kBlock = 256;
constexpr unsigned kMaxGrid = 65536;

template <typename T>
__device__ __forceinline__ cplx<T> scaled(cplx<T> a, double factor)
{
    // the factor is FP64, as in k_collapse; for FP32 states the exact
    // double product is rounded once, which equals the float product of a
    // float factor
    if constexpr (sizeof(T) == 8) {
        a.re *= factor;
        a.im *= factor;
    } else {
        a.re = (T)((double)a.re * factor);
        a.im = (T)((double)a.im * factor);
    }
    return a;
}

// Source amplitudes [first, first + count): block m = i >> targ goes to
// block 2m, and block 2m + 1 is zeroed.
// Both kernels take kUnroll amplitudes per thread per step, all loads
// issued before the stores (a load-then-store loop keeps one load in flight
// per thread).  Within a launch the read and written ranges are disjoint
// (or the same index), so the order inside a step does not matter.
constexpr int kUnroll = 4;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_expand(const cplx<T> *src, cplx<T> *dst,
                                                   uint64_t first, uint64_t count, unsigned targ)
{
    const uint64_t lomask = (1ull << targ) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t f0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f0 < count;
         f0 += kUnroll * stride) {
        cplx<T> a[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint64_t f = f0 + u * stride;
            a[u] = src[first + (f < count ? f : f0)];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint64_t f = f0 + u * stride;
            if (f >= count) continue;
            const uint64_t i = first + f;
            const uint64_t to = ((i & ~lomask) << 1) | (i & lomask);
            cplx<T> z;
            z.re = 0;
            z.im = 0;
            dst[to] = a[u];
            dst[to | (lomask + 1)] = z;
        }
    }
}

// Destination amplitudes [first, first + count): block m = j >> targ takes
// block 2m + outcome of the source, times factor.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_compact(const cplx<T> *src, cplx<T> *dst,
                                                    uint64_t first, uint64_t count, unsigned targ,
                                                    unsigned outcome, double factor)
{
    const uint64_t lomask = (1ull << targ) - 1;
    const uint64_t ob = (uint64_t)outcome << targ;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t f0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f0 < count;
         f0 += kUnroll * stride) {
        cplx<T> a[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint64_t f = f0 + u * stride;
            const uint64_t j = first + (f < count ? f : f0);
            a[u] = src[((j & ~lomask) << 1) | ob | (j & lomask)];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint64_t f = f0 + u * stride;
            if (f < count) dst[first + f] = scaled(a[u], factor);
        }
    }
}

unsigned grid_of(uint64_t count)
{
    uint64_t g = (count + kBlock - 1) / kBlock;
    if (g > kMaxGrid) g = kMaxGrid;
    return (unsigned)(g ? g : 1);
}

template <typename T>
void expand(qslhip_state *h, const void *src_v, void *dst_v, unsigned targ)
{
    const cplx<T> *src = static_cast<const cplx<T> *>(src_v);
    cplx<T> *dst = static_cast<cplx<T> *>(dst_v);
    const uint64_t k = 1ull << targ, H = h->namps >> targ;
    auto launch = [&](uint64_t m0, uint64_t m1) {
        const uint64_t first = m0 * k, count = (m1 - m0) * k;
        hipLaunchKernelGGL(k_expand<T>, dim3(grid_of(count)), dim3(kBlock), 0, h->stream, src,
                           dst, first, count, targ);
    };
    if (src_v != dst_v) {
        launch(0, H);  // out of place: one launch
    } else {
        for (uint64_t m0 = H / 2; m0 >= 1; m0 /= 2) launch(m0, 2 * m0);
        launch(0, 1);
    }
    QH_CHECK(hipGetLastError());
}

template <typename T>
void compact(qslhip_state *h, unsigned targ, unsigned outcome, double factor)
{
    cplx<T> *st = static_cast<cplx<T> *>(h->amps);
    const uint64_t k = 1ull << targ, M = (h->namps / 2) >> targ;
    auto launch = [&](uint64_t m0, uint64_t m1) {
        const uint64_t first = m0 * k, count = (m1 - m0) * k;
        hipLaunchKernelGGL(k_compact<T>, dim3(grid_of(count)), dim3(kBlock), 0, h->stream, st,
                           st, first, count, targ, outcome, factor);
    };
    launch(0, 1);
    for (uint64_t m0 = 1; m0 < M; m0 *= 2) launch(m0, 2 * m0);
    QH_CHECK(hipGetLastError());
}

}  // namespace

void resize_add_qubit(qslhip_state *h, unsigned targ)
{
    const uint64_t bytes_new = 2 * h->namps * h->amp_bytes();
    void *dst = h->amps;
    if (2 * h->namps > h->cap) {
        // the allocation doubles, as the reference's state.resize(2 * dim)
        // (resize.cpp:112-117); the move then runs out of place
        hipError_t e = hipMalloc(&dst, bytes_new);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(QSLHIP_ENOMEM, "hipMalloc of " + std::to_string(bytes_new) +
                                           " bytes for the larger state failed");
        }
    }
    if (h->prec == 8) expand<double>(h, h->amps, dst, targ);
    else expand<float>(h, h->amps, dst, targ);
    if (dst != h->amps) {
        QH_CHECK(hipStreamSynchronize(h->stream));
        QH_CHECK(hipFree(h->amps));
        h->amps = dst;
        h->cap = 2 * h->namps;
    }
    h->namps *= 2;
    h->n += 1;
}

void resize_collapse_out(qslhip_state *h, unsigned targ, unsigned outcome, double factor)
{
    if (h->prec == 8) compact<double>(h, targ, outcome, factor);
    else compact<float>(h, targ, outcome, factor);
    h->namps /= 2;
    h->n -= 1;
}

}  // namespace qslhip
