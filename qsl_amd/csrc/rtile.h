// rtile.h -- host side of the register-tile passes: the decoded op format,
// the planner (regtile.hip) and the launchers (regtile.hip, rtile_jit.cpp).
#pragma once

#include <string>
#include <vector>

#include "internal.h"
#include "rtile_core.h"

namespace qslhip {

enum RegCode : int32_t { RC_REMAP = 0, RC_PAIR = 1, RC_SWAP = 2, RC_ONE = 3, RC_RZ = 4 };

// One decoded op of a register-tile pass.  The first 64 bytes hold what
// every gate op reads.
struct RegOpHead {
    int32_t code;    // RegCode
    int32_t body;    // gate_math body (RC_PAIR, RC_SWAP, RC_ONE)
    int32_t ia, ib;  // RC_PAIR: register bit of the target; RC_SWAP: of q0 and q1
    uint32_t kmask;  // registers taking part (RC_PAIR: the bit-ia-clear one of each pair)
    uint32_t kneg;   // RC_RZ: registers whose target bit is 1
    uint32_t tmask;  // thread predicate (jt & tmask) == tmask
    uint32_t tneg;   // RC_RZ: thread-bit target, as a tile-position mask
    uint32_t flags;  // RF_FULL: kmask holds every register (pair op: every pair); RF_FSWAP
    uint32_t rho;    // RC_REMAP: register positions of the new layout
    uint64_t gmask;  // group predicate (gbase & gmask) == gmask
    double c, s;
};
static_assert(sizeof(RegOpHead) == 64, "one 64-byte scalar load");

enum : uint32_t { RF_FULL = 1, RF_FSWAP = 2 };

struct RegOp {
    RegOpHead h;
    uint64_t gneg;            // RC_RZ: target qubit outside the tile, as a mask
    uint32_t koff[kRegMaxR];  // RC_REMAP: swizzled LDS byte offset of register k
};

// Register positions of a pass's load and store layouts (0: the top r
// tile positions).
struct RegLayout {
    uint32_t in = 0, out = 0;
};

void plan_reg_ops(const TileOp *ops, uint32_t count, int K, int tb, int A,
                  std::vector<RegOp> &out, RegLayout *lay = nullptr);
uint32_t swz_bytes(uint32_t e, int A);
// where one pass's tile lives (load and store layouts, thread and register parts)
RegTileArgs make_tile_args(unsigned n, const TilePass &p, int tb, int A,
                           const RegLayout &lay = RegLayout());
// grid of a pass: gpw groups per workgroup when groups are plentiful
uint64_t reg_grid(uint64_t groups, int gpw);

// Per-launch settings of a pass beyond its ops (defaults: load the state,
// no run sums, every group).
struct PassExtras {
    uint64_t basis = ~0ull;  // RegTileArgs::basis
    double *runs = nullptr;  // RegTileArgs::runs
    uint64_t gvar = ~0ull;   // ~0: every group; else RegTileArgs::gvar / gfix
    uint64_t gfix = 0;
    uint64_t zmask = 0;      // RegTileArgs::zmask / zval
    uint64_t zval = 0;
    RegLayout lay;           // the pass's load and store layouts (plan_reg_ops)
    uint32_t gorder = 0;     // RegTileArgs::gorder
};
// the tile arguments of a pass with its extras applied, and its group count
RegTileArgs pass_args(unsigned n, const TilePass &p, int tb, int A, const PassExtras &x,
                      uint64_t *groups);

// Passes compiled at run time (rtile_jit.cpp): a pass structure seen before
// (or every pass, tuning "jit" = 2) runs as straight-line code generated
// from its decoded ops.  Returns false when the pass should take the
// ahead-of-time kernel instead (not repeated yet, or the compiler is not
// available); then nothing was launched.
bool jit_launch_pass(qslhip_state *h, void *amps, const TilePass &p, int threads, int gpw,
                     bool exact, const RegOp *ops, uint32_t nops, const double *dev_coef,
                     const PassExtras &x);
// the gate coefficients (c, s) of a pass's ops in the order the generated
// code reads them
void jit_coefficients(const RegOp *ops, uint32_t nops, std::vector<double> &coef);
// introspection: the generated source of a pass, and a compile check
std::string jit_pass_source(unsigned n, int prec, const TilePass &p, int threads, bool exact,
                            const RegOp *ops, uint32_t nops, const RegLayout &lay);
bool jit_check_source(const std::string &src, const std::string &arch, std::string &log);

}  // namespace qslhip
