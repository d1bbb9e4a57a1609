// dense.h -- fused dense-unitary passes (host planner dense.cpp, kernels
// mfma.hip).
#pragma once

#include "internal.h"

namespace qslhip {

constexpr int kDenseTile = 12;    // 2^12 FP32 amplitudes per tile (32 KiB)
constexpr int kDenseLowRun = 4;   // lowest qubits always in the tile

// One op of a dense pass: an ordinary tile op (kind 0) or a dense
// 2^nq x 2^nq complex unitary on tile positions pos[0..nq) (kind 1), whose
// row-major FP32 (re, im) entries start at complex offset `mat`.
struct DenseOp {
    int32_t kind;
    int32_t nq;
    int32_t pos[5];
    uint32_t mat;
    uint32_t gidx;  // first gate of the op in the input list (introspection)
    uint32_t frag;  // float offset of its MFMA A-operand fragments (device)
    TileOp top;
};

struct DensePass {
    uint64_t tile_mask;
    uint32_t op_first, op_count;
};

void plan_dense_passes(unsigned n, const std::vector<GateDesc> &gates, int F,
                       std::vector<DensePass> &passes, std::vector<DenseOp> &ops,
                       std::vector<float> &mats);
void launch_dense_passes(qslhip_state *h, int F, const std::vector<DensePass> &passes,
                         const std::vector<DenseOp> &ops, const std::vector<float> &mats);

}  // namespace qslhip
