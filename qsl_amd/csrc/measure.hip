// measure.hip -- reductions, collapse and sampling for full-vector handles.
//
// prob/norm: each lane accumulates |a|^2 in FP64 (the per-amplitude square
// is formed in Fp first, as the reference's `Fp prob = re*re + im*im` in
// src/qubits/default/measure.cpp:66-67), then a 64-lane xor-shuffle tree,
// then one partial per workgroup; a single workgroup adds the partials in a
// fixed order, so results are deterministic run to run.
//
// Sampling (measure.cpp:60-78, 164-188): a three-kernel reduce-then-scan
// builds the inclusive FP64 CDF over all 2^n amplitudes (zero entries leave
// the running sum unchanged, so "first index with cdf >= u" never lands on
// a zero-probability outcome, exactly like generateDist's skip), then one
// lane per shot binary-searches it.  A uniform above the total mass gives
// 0, as in drawSample.
#include <cstring>

#include "gate_math.h"

namespace qslhip {
namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048;       // partials of a reduction
constexpr int kScanItems = 16;         // amplitudes per lane in the scan
constexpr int kScanChunk = kBlock * kScanItems;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    return r;
}

template <typename T>
__device__ __forceinline__ double sq(const cplx<T> &a)
{
    T p = a.re * a.re + a.im * a.im;  // in Fp, like the reference
    return (double)p;
}

// Sum of |a|^2 over amplitudes with bit t == outcome (t < 0: all).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prob_partial(const cplx<T> *__restrict__ st,
                                                         uint64_t nitems, int t,
                                                         unsigned outcome,
                                                         double *__restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    double acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t ob = (t >= 0) ? ((uint64_t)outcome << t) : 0;
    for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < nitems; f += stride) {
        uint64_t i = (t >= 0) ? (ins0(f, (unsigned)t) | ob) : f;
        acc += sq(st[i]);
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *__restrict__ part,
                                                         int n, double *__restrict__ out)
{
    __shared__ double sh[kBlock / 64];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += part[i];
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = r;
}

// collapse (measure.cpp:31-57): kept half times factor, the other zeroed.
// The factor is FP64; a float state is scaled in FP64 and rounded once.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_collapse(cplx<T> *__restrict__ st,
                                                     uint64_t namps, unsigned t,
                                                     unsigned outcome, double factor)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < namps; i += stride) {
        cplx<T> a = st[i];
        if (((i >> t) & 1ull) == outcome) {
            if constexpr (sizeof(T) == 8) {
                a.re *= factor;
                a.im *= factor;
            } else {
                a.re = (T)((double)a.re * factor);
                a.im = (T)((double)a.im * factor);
            }
        } else {
            a.re = 0;
            a.im = 0;
        }
        st[i] = a;
    }
}

// Scan pass 1: per-chunk sums.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const cplx<T> *__restrict__ st,
                                                       uint64_t namps,
                                                       double *__restrict__ sums)
{
    __shared__ double sh[kBlock / 64];
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    double acc = 0;
#pragma unroll 4
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = c0 + (uint64_t)j * kBlock + threadIdx.x;
        if (i < namps) acc += sq(st[i]);
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// Scan pass 2: exclusive scan of the chunk sums in one workgroup (each lane
// takes a contiguous run, then a block scan of the run totals).
__global__ __launch_bounds__(1024) void k_scan_sums(double *__restrict__ sums, uint64_t n)
{
    __shared__ double sh[1024];
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = threadIdx.x * per;
    const uint64_t e = (b + per < n) ? b + per : n;
    double tot = 0;
    for (uint64_t i = b; i < e; i++) tot += sums[i];
    sh[threadIdx.x] = tot;
    __syncthreads();
    // Hillis-Steele inclusive scan of 1024 totals.
    for (int off = 1; off < 1024; off <<= 1) {
        double v = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    double run = (threadIdx.x > 0) ? sh[threadIdx.x - 1] : 0.0;
    for (uint64_t i = b; i < e; i++) {
        double v = sums[i];
        sums[i] = run;
        run += v;
    }
}

// Scan pass 3: inclusive CDF inside each chunk, offset by the chunk prefix.
// Lane l owns amplitudes [l*kScanItems, (l+1)*kScanItems) of the chunk so
// that its serial sum follows index order; loads go through LDS to stay
// coalesced.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_chunk_scan(const cplx<T> *__restrict__ st,
                                                       uint64_t namps,
                                                       const double *__restrict__ pre,
                                                       double *__restrict__ cdf)
{
    __shared__ double p[kScanChunk];
    __shared__ double lane_tot[kBlock];
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    for (int j = 0; j < kScanItems; j++) {
        int e = j * kBlock + threadIdx.x;
        uint64_t i = c0 + e;
        p[e] = (i < namps) ? sq(st[i]) : 0.0;
    }
    __syncthreads();
    double tot = 0;
    const int b = threadIdx.x * kScanItems;
    for (int j = 0; j < kScanItems; j++) tot += p[b + j];
    lane_tot[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        double v = (threadIdx.x >= (unsigned)off) ? lane_tot[threadIdx.x - off] : 0.0;
        __syncthreads();
        lane_tot[threadIdx.x] += v;
        __syncthreads();
    }
    double run = pre[blockIdx.x] + ((threadIdx.x > 0) ? lane_tot[threadIdx.x - 1] : 0.0);
    for (int j = 0; j < kScanItems; j++) {
        run += p[b + j];
        p[b + j] = run;
    }
    __syncthreads();
    for (int j = 0; j < kScanItems; j++) {
        int e = j * kBlock + threadIdx.x;
        uint64_t i = c0 + e;
        if (i < namps) cdf[i] = p[e];
    }
}

// First index i with cdf[i] >= u (u > 0), or with cdf[i] > 0 (u <= 0: the
// reference would index out of bounds; this picks the first outcome of
// nonzero probability); namps when u exceeds the total -> outcome 0.
// With `strict`, the first index with cdf[i] > u instead: the merge rule of
// sampleAll2 (measure.cpp:237-267, `rand < sum`).
__device__ __forceinline__ uint64_t cdf_search(const double *__restrict__ cdf,
                                               uint64_t namps, double u, bool strict)
{
    uint64_t lo = 0, hi = namps;
    const bool zero = strict || !(u > 0.0);
    const double key = strict ? u : 0.0;
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        double v = cdf[mid];
        bool ge = zero ? (v > key) : (v >= u);
        if (ge) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_search(const double *__restrict__ cdf,
                                                   uint64_t namps,
                                                   const double *__restrict__ u,
                                                   uint64_t nshots, int strict,
                                                   uint64_t *__restrict__ out)
{
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nshots) return;
    uint64_t r = cdf_search(cdf, namps, u[i], strict != 0);
    // past the total mass: drawSample returns 0; sampleAll2 counts nothing
    out[i] = (r >= namps) ? (strict ? ~0ull : 0ull) : r;
}

template <typename T>
double prob_impl(qslhip_state *h, int t, unsigned outcome)
{
    uint64_t nitems = (t >= 0) ? (h->namps >> 1) : h->namps;
    unsigned grid = grid_for(nitems, kBlock * 8);
    if (grid > (unsigned)kRedBlocks) grid = kRedBlocks;
    h->red.ensure((kRedBlocks + 8) * sizeof(double));
    double *part = static_cast<double *>(h->red.ptr);
    double *res = part + kRedBlocks;
    hipLaunchKernelGGL(k_prob_partial<T>, dim3(grid), dim3(kBlock), 0, h->stream,
                       static_cast<const cplx<T> *>(h->amps), nitems, t, outcome, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, h->stream, part,
                       (int)grid, res);
    QH_CHECK(hipGetLastError());
    h->pin.ensure(64);
    // the staging buffer may still be the source of an op-table copy
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    QH_CHECK(hipMemcpyAsync(h->pin.ptr, res, sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    double v;
    std::memcpy(&v, h->pin.ptr, sizeof(double));
    return v;
}

template <typename T>
void build_cdf(qslhip_state *h)
{
    const uint64_t namps = h->namps;
    const uint64_t nchunks = (namps + kScanChunk - 1) / kScanChunk;
    h->cdf.ensure(namps * sizeof(double) + (nchunks + 1) * sizeof(double));
    double *cdf = static_cast<double *>(h->cdf.ptr);
    double *sums = cdf + namps;
    const cplx<T> *st = static_cast<const cplx<T> *>(h->amps);
    hipLaunchKernelGGL(k_chunk_sums<T>, dim3((unsigned)nchunks), dim3(kBlock), 0,
                       h->stream, st, namps, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, sums, nchunks);
    hipLaunchKernelGGL(k_chunk_scan<T>, dim3((unsigned)nchunks), dim3(kBlock), 0,
                       h->stream, st, namps, sums, cdf);
    QH_CHECK(hipGetLastError());
}

}  // namespace

double device_prob(qslhip_state *h, unsigned t, unsigned outcome)
{
    return h->prec == 8 ? prob_impl<double>(h, (int)t, outcome)
                        : prob_impl<float>(h, (int)t, outcome);
}

double device_norm2(qslhip_state *h)
{
    return h->prec == 8 ? prob_impl<double>(h, -1, 0) : prob_impl<float>(h, -1, 0);
}

void launch_collapse(qslhip_state *h, unsigned t, unsigned outcome, double factor)
{
    unsigned grid = grid_for(h->namps, kBlock * 4);
    if (h->prec == 8)
        hipLaunchKernelGGL(k_collapse<double>, dim3(grid), dim3(kBlock), 0, h->stream,
                           static_cast<cplx<double> *>(h->amps), h->namps, t, outcome,
                           factor);
    else
        hipLaunchKernelGGL(k_collapse<float>, dim3(grid), dim3(kBlock), 0, h->stream,
                           static_cast<cplx<float> *>(h->amps), h->namps, t, outcome,
                           factor);
    QH_CHECK(hipGetLastError());
}

// out[i] = basis index sampled by uniform u[i] (host arrays).
void device_sample(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out,
                   bool strict)
{
    if (h->prec == 8) build_cdf<double>(h);
    else build_cdf<float>(h);
    if (nshots == 0) {
        QH_CHECK(hipStreamSynchronize(h->stream));
        return;
    }
    h->shots.ensure(nshots * 16);
    double *du = static_cast<double *>(h->shots.ptr);
    uint64_t *dout = reinterpret_cast<uint64_t *>(du + nshots);
    QH_CHECK(hipMemcpyAsync(du, u, nshots * sizeof(double), hipMemcpyHostToDevice,
                            h->stream));
    hipLaunchKernelGGL(k_search, dim3(grid_for(nshots, kBlock)), dim3(kBlock), 0,
                       h->stream, static_cast<const double *>(h->cdf.ptr), h->namps, du,
                       nshots, (int)strict, dout);
    QH_CHECK(hipGetLastError());
    QH_CHECK(hipMemcpyAsync(out, dout, nshots * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
}

}  // namespace qslhip

namespace qslhip {
// Prefix sums over the handle's device amplitudes, whatever their layout
// (used by the compact NP backend, whose rank order is index order).
void build_cdf_any(qslhip_state *h)
{
    if (h->prec == 8) build_cdf<double>(h);
    else build_cdf<float>(h);
}
}  // namespace qslhip
