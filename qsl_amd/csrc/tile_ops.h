// tile_ops.h -- ops applied to a group of 2^K amplitudes held in LDS.
//
// Shared by the LDS tile passes (fused.hip) and the batched variational
// kernels (adjoint.hip).  Every op updates the same amplitudes with the same
// arithmetic as the one-gate kernels (gate_math.h):
//   - non-diagonal targets are tile qubits (the planner guarantees it);
//   - controls and diagonal factors on qubits outside the tile are constant
//     over the group (read from the group's base index gbase), so the op is
//     either applied or skipped for the whole group.
// Threads of the workgroup stride over the tile; the caller puts a
// barrier between ops.
#pragma once

#include "gate_math.h"

namespace qslhip {

// Identity element order of an LDS tile (see mfma.hip for a swizzled one).
struct NoSwizzle {
    __device__ __forceinline__ uint32_t operator()(uint32_t e) const { return e; }
};

template <typename T, typename SW = NoSwizzle>
__device__ __forceinline__ void tile_op(cplx<T> *tile, int K, uint64_t gbase,
                                        const TileOp &op, SW sw = SW())
{
    const int tid = threadIdx.x;
    const uint32_t kBlock = blockDim.x;
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
                cplx<T> a = tile[sw(j0)], b = tile[sw(j1)];
                pair_update<T>(body, a, b, c, s);
                tile[sw(j0)] = a;
                tile[sw(j1)] = b;
            }
        } else {  // rotateZ on a non-tile qubit: one half-formula for all
            bool one = gbit(op.q0);
            for (uint32_t j = tid; j < nt; j += kBlock) rz_half<T>(one, tile[sw(j)], c, s);
        }
        break;
    case PC_ONE:
        if (pa >= 0) {
            for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                uint32_t j = (uint32_t)ins0(f, pa) | (1u << pa);
                one_update<T>(body, tile[sw(j)], c, s);
            }
        } else if (gbit(op.q0)) {
            for (uint32_t j = tid; j < nt; j += kBlock) one_update<T>(body, tile[sw(j)], c, s);
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
                    cplx<T> a = tile[sw(j0)], b = tile[sw(j1)];
                    pair_update<T>(body, a, b, c, s);
                    tile[sw(j0)] = a;
                    tile[sw(j1)] = b;
                }
            } else {
                for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                    uint32_t j0 = (uint32_t)ins0(f, pb), j1 = j0 | (1u << pb);
                    cplx<T> a = tile[sw(j0)], b = tile[sw(j1)];
                    pair_update<T>(body, a, b, c, s);
                    tile[sw(j0)] = a;
                    tile[sw(j1)] = b;
                }
            }
        } else {  // controlRotateZ with its target outside the tile
            bool one = gbit(op.q1);
            if (pa >= 0) {
                for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                    uint32_t j = (uint32_t)ins0(f, pa) | (1u << pa);
                    rz_half<T>(one, tile[sw(j)], c, s);
                }
            } else {
                for (uint32_t j = tid; j < nt; j += kBlock) rz_half<T>(one, tile[sw(j)], c, s);
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
                one_update<T>(body, tile[sw(j)], c, s);
            }
        } else if (pa >= 0 || pb >= 0) {
            int p = pa >= 0 ? pa : pb;
            for (uint32_t f = tid; f < nt / 2; f += kBlock) {
                uint32_t j = (uint32_t)ins0(f, p) | (1u << p);
                one_update<T>(body, tile[sw(j)], c, s);
            }
        } else {
            for (uint32_t j = tid; j < nt; j += kBlock) one_update<T>(body, tile[sw(j)], c, s);
        }
        break;
    }
    case PC_SWAPLIKE:
    case PC_FSWAP: {
        int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
        for (uint32_t f = tid; f < nt / 4; f += kBlock) {
            uint32_t base = (uint32_t)ins0(ins0(f, lo), hi);
            uint32_t j0 = base | (1u << pa), j1 = base | (1u << pb);
            cplx<T> a = tile[sw(j0)], b = tile[sw(j1)];
            pair_update<T>(body, a, b, c, s);
            tile[sw(j0)] = a;
            tile[sw(j1)] = b;
            if (cls == PC_FSWAP) {
                uint32_t j = base | (1u << pa) | (1u << pb);
                one_update<T>(QSLHIP_PAULI_Z, tile[sw(j)], c, s);
            }
        }
        break;
    }
    default:
        break;
    }
}

}  // namespace qslhip
