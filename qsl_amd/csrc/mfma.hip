// mfma.hip -- dense F-qubit unitaries (F = 4, 5) on FP32 matrix cores.
//
// A dense op acts on F tile positions of a 2^12-amplitude LDS tile: the
// tile splits into 2^(12-F) column vectors of 2^F amplitudes, and the pass
// computes Y = U X for all of them as one real GEMM with the interleaved
// embedding
//     [Re y; Im y] rows 2r, 2r+1  =  [[Ur, -Ui], [Ui, Ur]] (rows 2r/2r+1,
//                                      cols 2s/2s+1)  x  [Re x; Im x],
// i.e. M = 2^(F+1) rows, K = 2^(F+1), N = 32 columns per wave, issued as
// v_mfma_f32_32x32x2f32 (A 32x2, B 2x32, C/D 32x32 FP32 per wave): one
// k-step consumes one complex input amplitude (lane/32 selects re or im).
// Operand layouts (CDNA3/CDNA4 ISA, 32x32x2 f32, one block):
//   A: lane l holds A[l % 32][l / 32];  B: lane l holds B[l / 32][l % 32];
//   D: accumulator j of lane l holds D[8*(j/4) + 4*(l/32) + j%4][l % 32].
// Per amplitude a dense op costs 8 * 2^F real flops (256 for F = 5) against
// 16 bytes of HBM traffic for the whole pass, so a pass holding two or more
// 5-qubit unitaries is matrix-core bound.
#include <cstring>

#include "dense.h"
#include "tile_ops.h"

namespace qslhip {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t deposit32(uint32_t v, uint32_t mask)
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

template <int F>
__device__ __forceinline__ void dense_apply(cplx<float> *tile, int K, const DenseOp &op,
                                            const float2 *__restrict__ mats)
{
    constexpr int V = 1 << F;         // amplitudes per column vector
    constexpr int MT = (2 * V) / 32;  // 32-row tiles of the real embedding
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int nblk = ((1 << K) >> F) / 32;  // 32-column blocks
    uint32_t epos[F];
    uint32_t tmask = 0;
#pragma unroll
    for (int j = 0; j < F; j++) {
        epos[j] = 1u << op.pos[j];
        tmask |= epos[j];
    }
    const uint32_t cmask = ((1u << K) - 1) & ~tmask;
    const float2 *U = mats + op.mat;
    float *tf = reinterpret_cast<float *>(tile);
    const int hi = lane >> 5, lo = lane & 31;
    for (int cb = wave; cb < nblk; cb += nwaves) {
        const uint32_t colbase = deposit32((uint32_t)(cb * 32 + lo), cmask);
        floatx16 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[mt][j] = 0.f;
        for (int ks = 0; ks < V; ks++) {
            uint32_t eo = 0;
#pragma unroll
            for (int j = 0; j < F; j++)
                if ((ks >> j) & 1) eo |= epos[j];
            const float b = tf[2 * (colbase | eo) + hi];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const int row = mt * 32 + lo;
                const int r = row >> 1, rp = row & 1;
                const float2 u = U[r * V + ks];
                const float a = (rp == hi) ? u.x : (rp == 0 ? -u.y : u.y);
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[mt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int R = mt * 32 + 8 * (j >> 2) + 4 * hi + (j & 3);
                const int r = R >> 1, comp = R & 1;
                uint32_t eo = 0;
#pragma unroll
                for (int q = 0; q < F; q++)
                    if ((r >> q) & 1) eo |= epos[q];
                tf[2 * (colbase | eo) + comp] = acc[mt][j];
            }
        }
    }
}

template <int F>
__global__ __launch_bounds__(kBlock) void k_dense(cplx<float> *__restrict__ st, unsigned n,
                                                  int K, int crun, uint64_t tile_mask,
                                                  const DenseOp *__restrict__ ops, int nops,
                                                  const float2 *__restrict__ mats)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<float> *tile = reinterpret_cast<cplx<float> *>(smem);
    uint64_t *hoff = reinterpret_cast<uint64_t *>(smem + (sizeof(cplx<float>) << K));
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
        const DenseOp op = ops[i];
        if (op.kind == 0) tile_op<float>(tile, K, gbase, op.top);
        else dense_apply<F>(tile, K, op, mats);
        __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < nt; e += kBlock)
        st[gbase | (e & lowm) | hoff[e >> crun]] = tile[e];
}

}  // namespace

void launch_dense_passes(qslhip_state *h, int F, const std::vector<DensePass> &passes,
                         const std::vector<DenseOp> &ops, const std::vector<float> &mats)
{
    QH_REQUIRE(h->prec == 4, QSLHIP_EINVAL, "dense MFMA passes run on FP32 states");
    QH_REQUIRE(F == 4 || F == 5, QSLHIP_EINVAL, "fusion width must be 4 or 5");
    if (passes.empty()) return;
    const int K = kDenseTile;
    QH_REQUIRE((int)h->n >= K, QSLHIP_EINVAL, "dense passes need at least 12 qubits");
    size_t ob = ops.size() * sizeof(DenseOp);
    size_t ob_al = (ob + 255) & ~(size_t)255;
    size_t mb = mats.size() * sizeof(float);
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    h->pin.ensure(ob_al + mb + 16);
    h->optab.ensure(ob_al + mb + 16);
    std::memcpy(h->pin.ptr, ops.data(), ob);
    if (mb) std::memcpy(static_cast<char *>(h->pin.ptr) + ob_al, mats.data(), mb);
    QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, ob_al + mb, hipMemcpyHostToDevice,
                            h->stream));
    QH_CHECK(hipEventRecord(h->staged, h->stream));
    const DenseOp *dops = static_cast<const DenseOp *>(h->optab.ptr);
    const float2 *dm = reinterpret_cast<const float2 *>(static_cast<char *>(h->optab.ptr) + ob_al);
    for (const DensePass &p : passes) {
        int crun = 0;
        while (crun < 64 && ((p.tile_mask >> crun) & 1ull)) crun++;
        if (crun > K) crun = K;
        QH_REQUIRE(K - crun <= 8, QSLHIP_EINVAL, "dense pass with too many scattered qubits");
        size_t shmem = (sizeof(cplx<float>) << K) + sizeof(uint64_t) * ((size_t)1 << (K - crun));
        uint64_t groups = 1ull << (h->n - K);
        if (F == 5)
            hipLaunchKernelGGL(k_dense<5>, dim3((unsigned)groups), dim3(kBlock), shmem, h->stream,
                               static_cast<cplx<float> *>(h->amps), h->n, K, crun, p.tile_mask,
                               dops + p.op_first, (int)p.op_count, dm);
        else
            hipLaunchKernelGGL(k_dense<4>, dim3((unsigned)groups), dim3(kBlock), shmem, h->stream,
                               static_cast<cplx<float> *>(h->amps), h->n, K, crun, p.tile_mask,
                               dops + p.op_first, (int)p.op_count, dm);
        QH_CHECK(hipGetLastError());
    }
}

}  // namespace qslhip
