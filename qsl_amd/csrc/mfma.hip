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
// 5-qubit unitaries is matrix-core bound.  F = 4 is the default: half the
// flops per unitary, and layered circuits need only a few percent more
// 4-qubit unitaries than 5-qubit ones.
//
// The A operand of every (m-tile, k-step) is laid out on the host as the 64
// floats the 64 lanes need, so a wave fetches it with one coalesced 256-byte
// load (L2-resident: 16 KiB per 5-qubit unitary).  The LDS tile is stored
// XOR-swizzled, element e at e ^ ((e >> 5) & 31): column vectors whose
// stride is a multiple of 32 elements (targets on the low tile qubits) then
// fall into distinct banks instead of one.
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

struct Swizzle32 {
    __device__ __forceinline__ uint32_t operator()(uint32_t e) const
    {
        return e ^ ((e >> 5) & 31u);
    }
};

// Y = U X for the 32-column blocks of a tile.  Per op, every lane loads its
// A values for all (m-tile, k-step) pairs once (2^(F+1) registers: 64 at
// F = 5) and the swizzled offsets of the 2^F k-step elements are computed
// once as uniform values; the swizzle is XOR-linear, so an element at
// column c (per lane) and k-step offset e (uniform) sits at sw(c) ^ sw(e).
// Each block then runs the MT independent accumulator chains back to back,
// one LDS read of B per k-step (scheduled ahead by the compiler).
template <int F>
__device__ __forceinline__ void dense_apply(cplx<float> *tile, int K, const DenseOp &op,
                                            const float *__restrict__ frags)
{
    const Swizzle32 sw;
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
    uint32_t swe[V];
#pragma unroll
    for (int ks = 0; ks < V; ks++) {
        uint32_t eo = 0;
#pragma unroll
        for (int j = 0; j < F; j++)
            if ((ks >> j) & 1) eo |= epos[j];
        swe[ks] = sw(eo);
    }
    float a[MT * V];
    const float *A = frags + op.frag + lane;
#pragma unroll
    for (int i = 0; i < MT * V; i++) a[i] = A[i * 64];
    const uint32_t cmask = ((1u << K) - 1) & ~tmask;
    float *tf = reinterpret_cast<float *>(tile);
    const int hi = lane >> 5, lo = lane & 31;
    for (int cb = wave; cb < nblk; cb += nwaves) {
        const uint32_t csw = sw(deposit32((uint32_t)(cb * 32 + lo), cmask));
        floatx16 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[mt][j] = 0.f;
#pragma unroll
        for (int ks = 0; ks < V; ks++) {
            const float b = tf[2 * (csw ^ swe[ks]) + hi];
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[mt * V + ks], b, acc[mt], 0, 0,
                                                               0);
        }
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                // output row R of the embedding: element r = R / 2, part R % 2
                const int R0 = mt * 32 + 8 * (j >> 2) + (j & 3);  // hi = 0
                const uint32_t e0 = swe[R0 >> 1], e1 = swe[(R0 + 4) >> 1];
                // blend without a select (a select here picks between two
                // addresses of swe and keeps the table out of registers)
                const uint32_t e = e0 ^ ((e0 ^ e1) & (0u - (uint32_t)hi));
                tf[2 * (csw ^ e) + (R0 & 1)] = acc[mt][j];
            }
        }
    }
}

template <int F>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_dense(cplx<float> *__restrict__ st, unsigned n,
                                                  int K, int crun, uint64_t tile_mask,
                                                  const DenseOp *__restrict__ ops, int nops,
                                                  const float *__restrict__ frags)
{
    const Swizzle32 sw;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<float> *tile = reinterpret_cast<cplx<float> *>(smem);
    uint64_t *hoff = reinterpret_cast<uint64_t *>(smem + (sizeof(cplx<float>) << K));
    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const uint64_t gbase = deposit((uint64_t)blockIdx.x, full & ~tile_mask);
    const uint64_t hi_mask = tile_mask & ~((1ull << crun) - 1);
    const int hib = K - crun;
    for (int m = threadIdx.x; m < (1 << hib); m += kBlock) hoff[m] = deposit(m, hi_mask);
    __syncthreads();
    // the tile is 2^kDenseTile amplitudes (K == kDenseTile): every thread
    // issues its NE loads before the first wait, then fills LDS
    // (in two halves of NE / 2: fewer live registers)
    constexpr int NE = (1 << kDenseTile) / kBlock, NH = NE / 2;
    const uint32_t lowm = (1u << crun) - 1;
    for (int h = 0; h < 2; h++) {
        cplx<float> v[NH];
#pragma unroll
        for (int k = 0; k < NH; k++) {
            const uint32_t e = threadIdx.x + (h * NH + k) * kBlock;
            v[k] = st[gbase | (e & lowm) | hoff[e >> crun]];
        }
#pragma unroll
        for (int k = 0; k < NH; k++) tile[sw(threadIdx.x + (h * NH + k) * kBlock)] = v[k];
    }
    __syncthreads();
    for (int i = 0; i < nops; i++) {
        const DenseOp op = ops[i];
        if (op.kind == 0) tile_op<float>(tile, K, gbase, op.top, sw);
        else dense_apply<F>(tile, K, op, frags);
        __syncthreads();
    }
    // the store addresses are recomputed here, not kept in registers from
    // the loads across the op loop (where they would spill): the empty asm
    // statements make the thread index and crun opaque values, so nothing
    // derived from them before the loop is reused
    uint32_t tid = threadIdx.x, crun_s = (uint32_t)crun;
    asm volatile("" : "+v"(tid));
    asm volatile("" : "+s"(crun_s));
    const uint32_t lowm_s = (1u << crun_s) - 1;
    for (int h = 0; h < 2; h++) {
        cplx<float> v[NH];
#pragma unroll
        for (int k = 0; k < NH; k++) v[k] = tile[sw(tid + (h * NH + k) * kBlock)];
#pragma unroll
        for (int k = 0; k < NH; k++) {
            const uint32_t e = tid + (h * NH + k) * kBlock;
            st[gbase | (e & lowm_s) | hoff[e >> crun_s]] = v[k];
        }
    }
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
    // A-operand fragments: for dense op d, m-tile mt, k-step ks, lane l the
    // float A[mt*32 + l%32][2*ks + l/32] of the interleaved real embedding
    // of U (rows 2r / 2r+1 = Re / Im of y_r, columns 2s / 2s+1 = Re / Im of
    // x_s): (0,0) Re U, (0,1) -Im U, (1,0) Im U, (1,1) Re U.
    std::vector<DenseOp> hops(ops);
    std::vector<float> frags;
    const int V = 1 << F, MT = 2 * V / 32;
    for (DenseOp &o : hops) {
        if (o.kind != 1) continue;
        o.frag = (uint32_t)frags.size();
        const float *U = mats.data() + 2 * (size_t)o.mat;
        for (int mt = 0; mt < MT; mt++)
            for (int ks = 0; ks < V; ks++)
                for (int l = 0; l < 64; l++) {
                    const int row = mt * 32 + (l & 31), kp = l >> 5;
                    const int r = row >> 1, rp = row & 1;
                    const float ur = U[2 * ((size_t)r * V + ks)];
                    const float ui = U[2 * ((size_t)r * V + ks) + 1];
                    frags.push_back(rp == kp ? ur : (rp == 0 ? -ui : ui));
                }
    }
    size_t ob = hops.size() * sizeof(DenseOp);
    size_t ob_al = (ob + 255) & ~(size_t)255;
    size_t mb = frags.size() * sizeof(float);
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    h->pin.ensure(ob_al + mb + 16);
    h->optab.ensure(ob_al + mb + 16);
    std::memcpy(h->pin.ptr, hops.data(), ob);
    if (mb) std::memcpy(static_cast<char *>(h->pin.ptr) + ob_al, frags.data(), mb);
    QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, ob_al + mb, hipMemcpyHostToDevice,
                            h->stream));
    QH_CHECK(hipEventRecord(h->staged, h->stream));
    const DenseOp *dops = static_cast<const DenseOp *>(h->optab.ptr);
    const float *dm = reinterpret_cast<const float *>(static_cast<char *>(h->optab.ptr) + ob_al);
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
