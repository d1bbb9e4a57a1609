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
// This file holds the plain path: one circuit on a handle, one kernel per
// gate (FP64 or FP32) -- the per-circuit GPU path.  The batched path
// (thousands of circuits in one persistent kernel) is batch.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "adjoint.h"
#include "tile_ops.h"

namespace qslhip {

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

}  // namespace qslhip
