// measure.hip -- reductions, collapse and sampling for full-vector handles.
//
// prob/norm: each lane accumulates |a|^2 in FP64 (the per-amplitude square
// is formed in Fp first, as the reference's `Fp prob = re*re + im*im` in
// src/qubits/default/measure.cpp:66-67), then a 64-lane xor-shuffle tree,
// then one partial per workgroup; a single workgroup adds the partials in a
// fixed order, so results are deterministic run to run.
//
// Sampling (measure.cpp:60-78, 164-188): the outcome of uniform u is the
// first index whose running FP64 sum of |a|^2 reaches u (zero entries
// leave the running sum unchanged, so it never lands on a zero-probability
// outcome, exactly like generateDist's skip).  The running sums are not
// stored: one pass over the state writes sub-chunk and chunk sums, a
// single workgroup scans the chunk sums, and one wavefront per shot
// resolves chunk, sub-chunk and amplitude with 64-lane searches (k_sample).
// A uniform above the total mass gives 0, as in drawSample.  The compact
// NP layout keeps a stored CDF (build_cdf_any).
#include <cstring>
#include <mutex>
#include <string>

#include "gate_math.h"

namespace qslhip {

int g_run_sums_used = 0;

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048;       // partials of a reduction
constexpr int kScanItems = 16;         // amplitudes per lane in the scan
constexpr int kScanChunk = kBlock * kScanItems;

__device__ __forceinline__ double wave_sum(double v) { return wave_sum_xor(v); }

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
    return amp_prob(a);  // in Fp, like the reference
}

// Sum of |a|^2 over amplitudes with bit t == outcome (t < 0: all).
template <typename T, bool S>
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
        acc += sq(amp_ld<S>(st + i));
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
template <typename T, bool S>
__global__ __launch_bounds__(kBlock) void k_collapse(cplx<T> *__restrict__ st,
                                                     uint64_t namps, unsigned t,
                                                     unsigned outcome, double factor)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < namps; i += stride) {
        cplx<T> a = amp_ld<S>(st + i);
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
        amp_st<S>(st + i, a);
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
    if (threadIdx.x == 1023) sums[n] = run;  // the total: the last run's end
}

// Longer scans, in three steps: each workgroup scans 1024 entries in place
// (lane runs of 4, then a block scan of the run totals) and leaves its
// total; k_scan_sums scans the totals; k_scan_add offsets every entry.
constexpr int kScanPer = 4;
constexpr int kScanBlockN = kBlock * kScanPer;

__global__ __launch_bounds__(kBlock) void k_scan_local(double *__restrict__ sums, uint64_t n,
                                                       double *__restrict__ tot)
{
    __shared__ double sh[kBlock];
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlockN + threadIdx.x * kScanPer;
    double v[kScanPer], run = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        v[j] = b0 + j < n ? sums[b0 + j] : 0.0;
        run += v[j];
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const double o = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += o;
        __syncthreads();
    }
    double ex = threadIdx.x ? sh[threadIdx.x - 1] : 0.0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        if (b0 + j < n) sums[b0 + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == kBlock - 1) tot[blockIdx.x] = sh[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_add(double *__restrict__ sums, uint64_t n,
                                                     const double *__restrict__ tot, uint64_t nb)
{
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlockN + threadIdx.x * kScanPer;
    const double off = tot[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanPer; j++)
        if (b0 + j < n) sums[b0 + j] += off;
    if (blockIdx.x == 0 && threadIdx.x == 0) sums[n] = tot[nb];  // the total
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

// ---- sampling without a stored CDF ---------------------------------------
// A shot's outcome is the first index whose running sum of |a|^2 reaches
// the uniform, resolved in three levels that each hold 64 entries: chunks
// of kScanChunk amplitudes (exclusive prefix `pre`, one FP64 per chunk),
// their 64 sub-chunks of 64 amplitudes (sums `subs`), and the amplitudes
// of one sub-chunk, re-read from the state.  Only the state is read in
// full (once, by k_sub_sums); a shot reads 64 prefix values per chunk
// level and 1 KiB of amplitudes.

constexpr int kSub = 64;                          // amplitudes per sub-chunk
constexpr int kSubsPerChunk = kScanChunk / kSub;  // 64
constexpr int kWaves = kBlock / 64;
constexpr int kSubsPerWave = kSubsPerChunk / kWaves;  // 16

// Sub-chunk sums (one wave reduction each) and chunk sums (their wave
// totals added in sub order, then the four waves in order).
template <typename T, bool S>
__global__ __launch_bounds__(kBlock) void k_sub_sums(const cplx<T> *__restrict__ st,
                                                     uint64_t namps, double *__restrict__ subs,
                                                     double *__restrict__ sums)
{
    __shared__ double sh[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t a0 = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)w * kSubsPerWave * kSub;
    double p[kSubsPerWave];
#pragma unroll
    for (int j = 0; j < kSubsPerWave; j++) {
        const uint64_t i = a0 + (uint64_t)j * kSub + lane;
        p[j] = i < namps ? sq(amp_ld<S>(st + i)) : 0.0;
    }
    double tot = 0, mine = 0;
#pragma unroll
    for (int j = 0; j < kSubsPerWave; j++) {
        // 8-amplitude runs first, then the runs: the association a gate
        // pass's run sums (reg_run_sums) and k_subs_from_runs reproduce
        const double sj = lane_sum_up<6>(p[j]);
        tot += sj;
        if (lane == j) mine = sj;
    }
    if (lane < kSubsPerWave)
        subs[(uint64_t)blockIdx.x * kSubsPerChunk + w * kSubsPerWave + lane] = mine;
    if (lane == 0) sh[w] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        double c = 0;
        for (int k = 0; k < kWaves; k++) c += sh[k];
        sums[blockIdx.x] = c;
    }
}

// Where the run sums of a gate pass sit (reg_run_sums): runs[g * 2^lb + l]
// holds the run of tile index l << 3 in group g, lb = K - 3.
struct RunLayout {
    uint64_t tile, outside;  // the pass's tile qubits, the others
    uint32_t other;          // block bits that index a block's sub-chunks
    uint32_t qbit[3];        // block bit of qubit 3, 4, 5
    int lb;
};

// Sub-chunk sums from a gate pass's run sums.  Workgroup b takes groups
// 8b..8b+7 into LDS in order: block bits 0..lb-1 are a run's tile bits
// (tile qubits above 2, ascending), bits lb..lb+2 the group's three lowest
// bits (the three lowest outside qubits).  Qubits 3..5 are among them, so
// the eight runs of a sub-chunk are all in the block, combined by the tree
// lane_sum_up<6> builds over them in k_sub_sums; the other lb bits
// enumerate the block's 2^lb sub-chunks.
__global__ __launch_bounds__(kBlock) void k_subs_from_runs(const double *__restrict__ runs,
                                                           RunLayout L,
                                                           double *__restrict__ subs)
{
    extern __shared__ double blk[];
    const uint32_t nb = 8u << L.lb;
    const double *src = runs + (uint64_t)blockIdx.x * nb;
    // eight loads in flight per thread before the first LDS write
    constexpr uint32_t kU = 8;
    uint32_t i = threadIdx.x;
    for (; i + (kU - 1) * kBlock < nb; i += kU * kBlock) {
        double t[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) t[u] = src[i + u * kBlock];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) blk[i + u * kBlock] = t[u];
    }
    for (; i < nb; i += kBlock) blk[i] = src[i];
    __syncthreads();
    const uint32_t lmask = (1u << L.lb) - 1;
    for (uint32_t u = threadIdx.x; u <= lmask; u += kBlock) {
        const uint32_t p0 = (uint32_t)deposit(u, L.other);
        double r[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            uint32_t p = p0;
#pragma unroll
            for (int b = 0; b < 3; b++)
                if ((k >> b) & 1) p |= 1u << L.qbit[b];
            r[k] = blk[p];
        }
        const double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        const uint64_t g = (uint64_t)blockIdx.x * 8 + (p0 >> L.lb);
        const uint64_t a = deposit(g, L.outside) | deposit((uint64_t)(p0 & lmask) << 3, L.tile);
        subs[a >> 6] = s;
    }
}

// Chunk sums from sub-chunk sums: the same order of additions as
// k_sub_sums (a wave's 16 in order, then the four waves).
__global__ __launch_bounds__(kBlock) void k_chunk_from_subs(const double *__restrict__ subs,
                                                            uint64_t nchunks,
                                                            double *__restrict__ sums)
{
    const uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nchunks) return;
    const double *sc = subs + c * kSubsPerChunk;
    double tot = 0;
    for (int w = 0; w < kWaves; w++) {
        double tw = 0;
        for (int j = 0; j < kSubsPerWave; j++) tw += sc[w * kSubsPerWave + j];
        tot += tw;
    }
    sums[c] = tot;
}

__device__ __forceinline__ double wave_scan_incl(double v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Whether running sum v has reached the shot: v >= u, or v > key for the
// strict rule of sampleAll2 (measure.cpp:237-267, `rand < sum`) and for
// u <= 0 (key 0: the first outcome of nonzero probability, where the
// reference would index out of bounds).
struct ShotRule {
    double u, key;
    bool gt;
    __device__ bool operator()(double v) const { return gt ? (v > key) : (v >= u); }
};

// First lane whose flag is set, or -1.
__device__ __forceinline__ int first_lane(bool f)
{
    const uint64_t b = __ballot(f);
    return b ? __ffsll((unsigned long long)b) - 1 : -1;
}
__device__ __forceinline__ int last_lane(bool f)
{
    const uint64_t b = __ballot(f);
    return b ? 63 - __clzll((long long)b) : -1;
}

// Search of one shot.  Level 1: the first chunk c whose inclusive end
// pre[c + 1] reaches the shot (64-way search, three rounds for 2^18
// chunks).  Levels 2 and 3: running sums from pre[c] over the sub-chunk
// sums, then over the chosen sub-chunk's |a|^2; the outcome is the first
// entry of nonzero probability whose running sum reaches the shot (zero
// entries leave the reference's running sum unchanged, and generateDist
// skips them), else the last entry of nonzero probability (the chunk's
// total reached the shot, its partial sums fell short by rounding).

// One round of level 1 on [lo, hi]: hi is a chunk known to hit, or nchunks
// (none).  Lane l probes the end of the l-th of 64 equal steps.
struct Level1 {
    uint64_t lo, hi, step, c;
    __device__ void probe(int lane)
    {
        const uint64_t span = hi - lo;
        step = span <= 64 ? 1 : (span + 63) / 64;
        c = lo + (uint64_t)(lane + 1) * step - 1;
        if (c > hi - 1) c = hi - 1;
    }
    __device__ void narrow(int f)
    {
        if (f < 0) {
            lo = hi;  // the last probe is hi - 1
        } else {
            const uint64_t cf = lo + (uint64_t)(f + 1) * step - 1;
            if (f > 0) lo += (uint64_t)f * step;  // past the probe of lane f - 1
            hi = cf < hi - 1 ? cf : hi - 1;
        }
    }
};

// Levels 2 and 3 from chunk c on; a chunk whose partial sums all stay zero
// (its prefix moved only by rounding) hands the shot to the next chunk.
// Past the total mass: drawSample returns 0; sampleAll2 counts nothing.
template <typename T>
__device__ uint64_t resolve_from(const cplx<T> *__restrict__ st, uint64_t namps,
                                 const double *__restrict__ pre, uint64_t nchunks,
                                 const double *__restrict__ subs, const ShotRule &hit,
                                 int strict, uint64_t c, int lane)
{
    for (; c < nchunks; c++) {
        const double base = pre[c];
        const double sv = subs[c * kSubsPerChunk + lane];
        const double run = base + wave_scan_incl(sv, lane);
        int j = first_lane(sv > 0.0 && hit(run));
        if (j < 0) j = last_lane(sv > 0.0);
        if (j < 0) continue;
        const double sbase = j ? __shfl(run, j - 1, 64) : base;
        const uint64_t i = (c * kSubsPerChunk + (uint64_t)j) * kSub + (uint64_t)lane;
        const double p = i < namps ? sq(st[i]) : 0.0;
        const double run2 = sbase + wave_scan_incl(p, lane);
        int e = first_lane(p > 0.0 && hit(run2));
        if (e < 0) e = last_lane(p > 0.0);
        if (e < 0) continue;
        return (c * kSubsPerChunk + (uint64_t)j) * kSub + (uint64_t)e;
    }
    return strict ? ~0ull : 0ull;
}

// First chunk attempt of levels 2 and 3, split in phases so that two
// shots' loads are in flight together: loads, sub-chunk scan, amplitude
// load, amplitude scan.  e stays -1 when the chunk held nothing (the shot
// then continues with resolve_from at c + 1).
struct FirstChunk {
    uint64_t c;
    bool live;
    double base, sv, run, p;
    int j, e;
    uint64_t i;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sample(const cplx<T> *__restrict__ st,
                                                   uint64_t namps,
                                                   const double *__restrict__ pre,
                                                   uint64_t nchunks,
                                                   const double *__restrict__ subs,
                                                   const double *__restrict__ u,
                                                   uint64_t nshots, int strict,
                                                   uint64_t *__restrict__ out)
{
    // two shots per wavefront, searched in lockstep: each shot's arithmetic
    // is the one-shot search above, the second shot's memory latency hides
    // behind the first's
    const int lane = threadIdx.x & 63;
    const uint64_t shot = ((uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6)) * 2;
    if (shot >= nshots) return;  // the whole wave
    const bool two = shot + 1 < nshots;
    ShotRule hit[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double x = (s == 0 || two) ? u[shot + s] : 0.0;
        hit[s] = ShotRule{x, strict ? x : 0.0, strict != 0 || !(x > 0.0)};
    }
    Level1 L[2];
    L[0].lo = 0;
    L[0].hi = nchunks;
    L[1].lo = 0;
    L[1].hi = two ? nchunks : 0;
    while (L[0].lo < L[0].hi || L[1].lo < L[1].hi) {
        double pv[2] = {0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 2; s++)
            if (L[s].lo < L[s].hi) {
                L[s].probe(lane);
                pv[s] = pre[L[s].c + 1];
            }
#pragma unroll
        for (int s = 0; s < 2; s++)
            if (L[s].lo < L[s].hi) L[s].narrow(first_lane(hit[s](pv[s])));
    }
    FirstChunk F[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        F[s].c = L[s].lo;
        F[s].live = (s == 0 || two) && F[s].c < nchunks;
        F[s].j = -1;
        F[s].e = -1;
        F[s].base = F[s].sv = F[s].run = F[s].p = 0.0;
        F[s].i = 0;
        if (F[s].live) {
            F[s].base = pre[F[s].c];
            F[s].sv = subs[F[s].c * kSubsPerChunk + lane];
        }
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
        if (F[s].live) {
            F[s].run = F[s].base + wave_scan_incl(F[s].sv, lane);
            int j = first_lane(F[s].sv > 0.0 && hit[s](F[s].run));
            if (j < 0) j = last_lane(F[s].sv > 0.0);
            F[s].j = j;
            if (j >= 0) {
                F[s].base = j ? __shfl(F[s].run, j - 1, 64) : F[s].base;
                F[s].i = (F[s].c * kSubsPerChunk + (uint64_t)j) * kSub + (uint64_t)lane;
                F[s].p = F[s].i < namps ? sq(st[F[s].i]) : 0.0;
            }
        }
#pragma unroll
    for (int s = 0; s < 2; s++)
        if (F[s].j >= 0) {
            const double run2 = F[s].base + wave_scan_incl(F[s].p, lane);
            int e = first_lane(F[s].p > 0.0 && hit[s](run2));
            if (e < 0) e = last_lane(F[s].p > 0.0);
            F[s].e = e;
        }
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s == 1 && !two) break;
        uint64_t o;
        if (F[s].e >= 0)
            o = (F[s].c * kSubsPerChunk + (uint64_t)F[s].j) * kSub + (uint64_t)F[s].e;
        else
            o = resolve_from<T>(st, namps, pre, nchunks, subs, hit[s], strict,
                                F[s].live ? F[s].c + 1 : nchunks, lane);
        if (lane == 0) out[shot + s] = o;
    }
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
    if ((int)h->n >= tuning().stream_min_qubits)
        hipLaunchKernelGGL((k_prob_partial<T, true>), dim3(grid), dim3(kBlock), 0, h->stream,
                           static_cast<const cplx<T> *>(h->amps), nitems, t, outcome, part);
    else
        hipLaunchKernelGGL((k_prob_partial<T, false>), dim3(grid), dim3(kBlock), 0, h->stream,
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

// Exclusive scan of sums[0..n) in place, with the total in sums[n].
void scan_exclusive(qslhip_state *h, double *sums, uint64_t n)
{
    if (n <= 4096) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, sums, n);
        return;
    }
    const uint64_t nb = (n + kScanBlockN - 1) / kScanBlockN;
    h->red.ensure((nb + 1) * sizeof(double));  // block totals; free between reductions
    double *tot = static_cast<double *>(h->red.ptr);
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kBlock), 0, h->stream, sums, n,
                       tot);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, tot, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kBlock), 0, h->stream, sums, n,
                       static_cast<const double *>(tot), nb);
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
    scan_exclusive(h, sums, nchunks);
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

template <typename T, bool S>
void collapse_impl(qslhip_state *h, unsigned t, unsigned outcome, double factor)
{
    unsigned grid = grid_for(h->namps, kBlock * 4);
    hipLaunchKernelGGL((k_collapse<T, S>), dim3(grid), dim3(kBlock), 0, h->stream,
                       static_cast<cplx<T> *>(h->amps), h->namps, t, outcome, factor);
}

void launch_collapse(qslhip_state *h, unsigned t, unsigned outcome, double factor)
{
    const bool stream = (int)h->n >= tuning().stream_min_qubits;
    if (h->prec == 8) {
        if (stream) collapse_impl<double, true>(h, t, outcome, factor);
        else collapse_impl<double, false>(h, t, outcome, factor);
    } else {
        if (stream) collapse_impl<float, true>(h, t, outcome, factor);
        else collapse_impl<float, false>(h, t, outcome, factor);
    }
    QH_CHECK(hipGetLastError());
}

namespace {

// [chunk prefix: nchunks + 1][sub-chunk sums: 64 per chunk][8-amplitude
// run sums: 8 per sub-chunk]; returns the sub-chunk sums
double *run_sums_buffer(qslhip_state *h)
{
    const uint64_t nchunks = (h->namps + kScanChunk - 1) / kScanChunk;
    const uint64_t nsubs = nchunks * kSubsPerChunk;
    h->cdf.ensure((nchunks + 1 + nsubs + nsubs * 8) * sizeof(double));
    return static_cast<double *>(h->cdf.ptr) + nchunks + 1;
}

RunLayout run_layout(const qslhip_state *h)
{
    const int K = h->sums_logt;
    QH_REQUIRE((h->sums_tile & 7) == 7 && __builtin_popcountll(h->sums_tile) == K &&
                   h->n >= (unsigned)K + 3 && K >= 6 && K <= 13,
               QSLHIP_EINVAL, "run sums from an unexpected tile");
    RunLayout L{};
    L.tile = h->sums_tile;
    L.outside = ((h->n >= 64) ? ~0ull : ((1ull << h->n) - 1)) & ~L.tile;
    L.lb = K - 3;
    int bit[6] = {-1, -1, -1, -1, -1, -1};  // block bit of qubits 0..5
    int lpos = 0;
    for (unsigned q = 3; q < h->n && lpos < L.lb; q++)
        if ((L.tile >> q) & 1) {
            if (q < 6) bit[q] = lpos;
            lpos++;
        }
    int opos = 0;
    for (unsigned q = 3; q < h->n && opos < 3; q++)
        if ((L.outside >> q) & 1) {
            if (q < 6) bit[q] = L.lb + opos;
            opos++;
        }
    uint32_t all = (1u << (L.lb + 3)) - 1, m345 = 0;
    for (int b = 0; b < 3; b++) {
        QH_REQUIRE(bit[3 + b] >= 0, QSLHIP_EINVAL, "qubit 3..5 outside the run block");
        L.qbit[b] = (uint32_t)bit[3 + b];
        m345 |= 1u << bit[3 + b];
    }
    L.other = all & ~m345;
    return L;
}

// out[i] = basis index sampled by uniform u[i] (host arrays).
template <typename T>
void sample_impl(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out, bool strict)
{
    const uint64_t namps = h->namps;
    const uint64_t nchunks = (namps + kScanChunk - 1) / kScanChunk;
    double *subs = run_sums_buffer(h);
    double *pre = subs - (nchunks + 1);
    const cplx<T> *st = static_cast<const cplx<T> *>(h->amps);
    if (h->sums_valid) {  // the last gate pass wrote the run sums
        g_run_sums_used++;
        const uint64_t nsubs = nchunks * kSubsPerChunk;
        const RunLayout L = run_layout(h);
        const uint64_t nblocks = 1ull << (h->n - h->sums_logt - 3);
        hipLaunchKernelGGL(k_subs_from_runs, dim3((unsigned)nblocks), dim3(kBlock),
                           (size_t)64 << L.lb, h->stream, subs + nsubs, L, subs);
        hipLaunchKernelGGL(k_chunk_from_subs, dim3(grid_for(nchunks, kBlock)), dim3(kBlock), 0,
                           h->stream, subs, nchunks, pre);
    } else {
        if ((int)h->n >= tuning().stream_min_qubits)
            hipLaunchKernelGGL((k_sub_sums<T, true>), dim3((unsigned)nchunks), dim3(kBlock), 0,
                               h->stream, st, namps, subs, pre);
        else
            hipLaunchKernelGGL((k_sub_sums<T, false>), dim3((unsigned)nchunks), dim3(kBlock), 0,
                               h->stream, st, namps, subs, pre);
    }
    scan_exclusive(h, pre, nchunks);
    QH_CHECK(hipGetLastError());
    if (nshots == 0) {
        QH_CHECK(hipStreamSynchronize(h->stream));
        return;
    }
    h->shots.ensure(nshots * 16);
    double *du = static_cast<double *>(h->shots.ptr);
    uint64_t *dout = reinterpret_cast<uint64_t *>(du + nshots);
    if (!h->copy_stream) {
        QH_CHECK(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
        for (hipEvent_t &e : h->xev) QH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // The uniforms go up on the copy stream while the sums kernels run (they
    // are queued first: a copy from pageable memory may hold the host), and
    // each quarter of the outcomes comes down while the next is computed.
    QH_CHECK(hipMemcpyAsync(du, u, nshots * sizeof(double), hipMemcpyHostToDevice,
                            h->copy_stream));
    QH_CHECK(hipEventRecord(h->xev[0], h->copy_stream));
    QH_CHECK(hipStreamWaitEvent(h->stream, h->xev[0], 0));
    const uint64_t parts = nshots >= (1u << 18) ? 4 : 1;
    const uint64_t per = (nshots + parts - 1) / parts;
    for (uint64_t c = 0; c < parts; c++) {
        const uint64_t c0 = c * per, cn = c0 + per < nshots ? per : nshots - c0;
        hipLaunchKernelGGL(k_sample<T>, dim3(grid_for(cn, 2 * kWaves)), dim3(kBlock), 0, h->stream,
                           st, namps, pre, nchunks, subs, du + c0, cn, (int)strict, dout + c0);
        QH_CHECK(hipGetLastError());
        QH_CHECK(hipEventRecord(h->xev[1 + c], h->stream));
    }
    for (uint64_t c = 0; c < parts; c++) {
        const uint64_t c0 = c * per, cn = c0 + per < nshots ? per : nshots - c0;
        QH_CHECK(hipStreamWaitEvent(h->copy_stream, h->xev[1 + c], 0));
        QH_CHECK(hipMemcpyAsync(out + c0, dout + c0, cn * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, h->copy_stream));
    }
    QH_CHECK(hipStreamSynchronize(h->copy_stream));
}

}  // namespace

// ---- sampling without allocation (qslhip_sample_all_noalloc) -------------------
//
// The same kernels as sample_impl (sub-chunk sums from the state, one
// chunk scan, the two-shot wavefront search), with every buffer in one
// file-scope device array instead of the handle's lazily allocated ones:
// [chunk prefix: kNoAllocChunks + 1][sub-chunk sums][uniforms][outcomes].
// Host code never passes `scratch` itself (its name is a host-side
// stand-in); it takes each device's address once with hipGetSymbolAddress
// and passes that.  Every call on one device shares the array and overwrites
// it; nothing in it must start at zero (each kernel writes what the next
// reads).  Calls that overlap on one device race.
namespace {

constexpr uint64_t kNoAllocChunks = kNoAllocMaxAmps / kScanChunk;  // 4096: one-block scan
constexpr uint64_t kNoAllocDoubles =
    (kNoAllocChunks + 1) + kNoAllocChunks * kSubsPerChunk + 2 * kNoAllocMaxShots;
static_assert(kNoAllocMaxAmps % kScanChunk == 0 && kNoAllocChunks <= 4096,
              "the chunk scan of the static sampler runs in one workgroup");

__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};  // per-device address of `scratch`
std::mutex g_scratch_mu;

// The device address of `scratch` on device `dev` (taken once per device,
// with that device current).
double *scratch_on(int dev)
{
    QH_REQUIRE(dev >= 0 && dev < kMaxDevices, QSLHIP_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        int cur = 0;
        QH_CHECK(hipGetDevice(&cur));
        if (cur != dev) QH_CHECK(hipSetDevice(dev));
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        QH_CHECK(e);
        QH_REQUIRE(ptr != nullptr, QSLHIP_EDEVICE, "no device address for the sampling scratch");
        g_scratch_dev[dev] = static_cast<double *>(ptr);
    }
    return g_scratch_dev[dev];
}

// FP32 handles compare float uniforms, as the reference's Random<float>.
__global__ void k_round_to_float(double *u, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (double)(float)u[i];
}

template <typename T>
void sample_noalloc_impl(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out)
{
    const uint64_t namps = h->namps;
    const uint64_t nchunks = (namps + kScanChunk - 1) / kScanChunk;
    double *base = scratch_on(h->device);
    double *pre = base;
    double *subs = pre + kNoAllocChunks + 1;
    double *du = subs + kNoAllocChunks * kSubsPerChunk;
    uint64_t *dout = reinterpret_cast<uint64_t *>(du + kNoAllocMaxShots);
    const cplx<T> *st = static_cast<const cplx<T> *>(h->amps);
    hipLaunchKernelGGL((k_sub_sums<T, false>), dim3((unsigned)nchunks), dim3(kBlock), 0,
                       h->stream, st, namps, subs, pre);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, pre, nchunks);
    QH_CHECK(hipMemcpyAsync(du, u, nshots * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (sizeof(T) == 4)
        hipLaunchKernelGGL(k_round_to_float, dim3(grid_for(nshots, kBlock)), dim3(kBlock), 0,
                           h->stream, du, nshots);
    hipLaunchKernelGGL(k_sample<T>, dim3(grid_for(nshots, 2 * kWaves)), dim3(kBlock), 0,
                       h->stream, st, namps, pre, nchunks, subs, du, nshots, 0, dout);
    QH_CHECK(hipGetLastError());
    QH_CHECK(hipMemcpyAsync(out, dout, nshots * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
}

}  // namespace

void device_sample_noalloc(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out)
{
    QH_REQUIRE(!h->np, QSLHIP_EKIND, "the allocation-free sampler takes full-vector handles");
    QH_REQUIRE(h->namps <= kNoAllocMaxAmps, QSLHIP_EINVAL,
               "the allocation-free sampler takes states of at most 2^24 amplitudes "
               "(24 qubits); this one has " + std::to_string(h->namps));
    QH_REQUIRE(nshots <= kNoAllocMaxShots, QSLHIP_EINVAL,
               "the allocation-free sampler takes at most 2^18 shots per call; got " +
                   std::to_string(nshots));
    if (nshots == 0) return;
    if (h->prec == 8) sample_noalloc_impl<double>(h, u, nshots, out);
    else sample_noalloc_impl<float>(h, u, nshots, out);
}

double *sample_run_sums(qslhip_state *h)
{
    // whole chunks only: every sub-chunk sum k_chunk_from_subs reads is written
    if (h->np || h->namps % kScanChunk) return nullptr;
    const uint64_t nchunks = h->namps / kScanChunk, nsubs = nchunks * kSubsPerChunk;
    // only a handle that has sampled before holds the buffer (namps bytes of
    // run sums); gate passes never allocate it
    if (h->cdf.bytes < (nchunks + 1 + nsubs + nsubs * 8) * sizeof(double)) return nullptr;
    return run_sums_buffer(h) + nsubs;
}

void device_sample(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out,
                   bool strict)
{
    if (h->prec == 8) sample_impl<double>(h, u, nshots, out, strict);
    else sample_impl<float>(h, u, nshots, out, strict);
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
