// dense.cpp -- host side of the MFMA path: fuse gate runs into dense
// unitaries on up to F tile qubits (F = 4 or 5), in FP64, and lay out the
// passes that apply them (mfma.hip).
//
// Packing reuses the LDS tile planner (every pass owns 2^12 tile qubits).
// Inside a pass, consecutive gates are folded into an open block while the
// union of their tile qubits stays within F; a gate that fits neither is
// either (a) disjoint from the block, so it commutes with it and is emitted
// as an ordinary tile op ahead of it, or (b) overlapping, which closes the
// block.  Gates with a control or diagonal factor on a non-tile qubit stay
// ordinary tile ops.  A block of a single gate is emitted as that gate.
// The block matrix is the ordered product of the constituent 2x2 / 4x4
// matrices of docs/overview.rst (the reference's gate conventions),
// computed in FP64 and rounded once to FP32.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "gate_math.h"
#include "dense.h"

namespace qslhip {

using cd = std::complex<double>;

// The 2x2 matrix of a one-qubit body (row = output bit, col = input bit).
static void mat2(int body, double c, double s, cd m[2][2])
{
    const cd I(0, 1);
    const double r2 = 1.0 / std::sqrt(2.0);
    switch (body) {
    case QSLHIP_PAULI_X: m[0][0] = 0; m[0][1] = 1; m[1][0] = 1; m[1][1] = 0; break;
    case QSLHIP_PAULI_Y: m[0][0] = 0; m[0][1] = -I; m[1][0] = I; m[1][1] = 0; break;
    case QSLHIP_PAULI_Z: m[0][0] = 1; m[0][1] = 0; m[1][0] = 0; m[1][1] = -1; break;
    case QSLHIP_HADAMARD: m[0][0] = r2; m[0][1] = r2; m[1][0] = r2; m[1][1] = -r2; break;
    case QSLHIP_PHASE: m[0][0] = 1; m[0][1] = 0; m[1][0] = 0; m[1][1] = cd(c, s); break;
    case QSLHIP_ROT_X: m[0][0] = c; m[0][1] = -I * s; m[1][0] = -I * s; m[1][1] = c; break;
    case QSLHIP_ROT_Y: m[0][0] = c; m[0][1] = -s; m[1][0] = s; m[1][1] = c; break;
    case QSLHIP_ROT_Z: m[0][0] = cd(c, -s); m[0][1] = 0; m[1][0] = 0; m[1][1] = cd(c, s); break;
    default: m[0][0] = 1; m[0][1] = 0; m[1][0] = 0; m[1][1] = 1; break;
    }
}

// Apply gate g (qubits given as positions inside the block, pos0/pos1) to
// the columns of the dim x dim block matrix M (M <- G M).
static void left_apply(std::vector<cd> &M, int dim, const GateDesc &g, int p0, int p1)
{
    const int cls = pair_class(g.kind);
    const int body = body_of(g.kind);
    cd m[2][2];
    mat2(body, g.c, g.s, m);
    for (int col = 0; col < dim; col++) {
        auto at = [&](int row) -> cd & { return M[(size_t)row * dim + col]; };
        if (cls == PC_PAIR1 || cls == PC_ONE) {
            const int b = 1 << p0;
            for (int r = 0; r < dim; r++) {
                if (r & b) continue;
                cd a0 = at(r), a1 = at(r | b);
                at(r) = m[0][0] * a0 + m[0][1] * a1;
                at(r | b) = m[1][0] * a0 + m[1][1] * a1;
            }
        } else if (cls == PC_CTRL) {  // p0 control, p1 target
            const int cb = 1 << p0, tb = 1 << p1;
            for (int r = 0; r < dim; r++) {
                if (!(r & cb) || (r & tb)) continue;
                cd a0 = at(r), a1 = at(r | tb);
                at(r) = m[0][0] * a0 + m[0][1] * a1;
                at(r | tb) = m[1][0] * a0 + m[1][1] * a1;
            }
        } else if (cls == PC_DIAG2) {
            const int mb = (1 << p0) | (1 << p1);
            for (int r = 0; r < dim; r++)
                if ((r & mb) == mb) at(r) = m[1][1] * at(r);
        } else {  // swap-like on |10>,|01> (a = q0 set), fswap negates |11>
            const int ab = 1 << p0, bb = 1 << p1;
            for (int r = 0; r < dim; r++) {
                if (!(r & ab) || (r & bb)) continue;
                int r2 = (r ^ ab) | bb;
                cd a0 = at(r), a1 = at(r2);
                at(r) = m[0][0] * a0 + m[0][1] * a1;
                at(r2) = m[1][0] * a0 + m[1][1] * a1;
            }
            if (cls == PC_FSWAP)
                for (int r = 0; r < dim; r++)
                    if ((r & ab) && (r & bb)) at(r) = -at(r);
        }
    }
}

void plan_dense_passes(unsigned n, const std::vector<GateDesc> &gates, int F,
                       std::vector<DensePass> &passes, std::vector<DenseOp> &ops,
                       std::vector<float> &mats)
{
    const int K = (int)std::min<unsigned>(kDenseTile, n);
    std::vector<TilePass> tp;
    std::vector<TileOp> to;
    // The tile passes: in program order, or commutation-aware (tuning
    // dense_reorder 1: gates on disjoint qubits, or both diagonal, taken in
    // either order, the reordered mode's planner).  order maps a tile op to
    // its gate.
    std::vector<uint32_t> order;
    if (tuning().dense_reorder) {
        plan_tile_passes_reorder(n, gates, K, kDenseLowRun, tp, to, order);
    } else {
        plan_tile_passes(n, 4, gates, K, kDenseLowRun, tp, to);
        order.resize(to.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (uint32_t)i;
    }
    QH_REQUIRE(order.size() == to.size(), QSLHIP_EINVAL, "dense plan: op order mismatch");
    for (const TilePass &P : tp) {
        DensePass D;
        D.tile_mask = P.tile_mask;
        D.op_first = (uint32_t)ops.size();
        // Open blocks: disjoint sets of tile positions, each folding gates
        // that touch only its positions.  Blocks on disjoint qubits commute,
        // so their emission order among themselves is free; a gate must
        // follow every block it shares a position with.
        struct Block {
            std::vector<int> S;               // tile positions
            std::vector<size_t> members;      // gate indices, program order
            std::vector<TileOp> tops;
        };
        std::vector<Block> open;
        auto emit_tile = [&](size_t gi, const TileOp &t) {
            DenseOp o{};
            o.kind = 0;
            o.gidx = (uint32_t)gi;
            o.top = t;
            ops.push_back(o);
        };
        auto has = [](const Block &b, int p) {
            return std::find(b.S.begin(), b.S.end(), p) != b.S.end();
        };
        auto emit_block = [&](Block &b) {
            if (b.members.size() == 1) {
                emit_tile(b.members[0], b.tops[0]);
                return;
            }
            // pad the block to F positions with unused tile positions
            std::vector<int> Sp = b.S;
            for (int p = 0; p < K && (int)Sp.size() < F; p++)
                if (std::find(Sp.begin(), Sp.end(), p) == Sp.end()) Sp.push_back(p);
            const int dim = 1 << F;
            std::vector<cd> M((size_t)dim * dim, 0.0);
            for (int i = 0; i < dim; i++) M[(size_t)i * dim + i] = 1.0;
            for (size_t k = 0; k < b.members.size(); k++) {
                const TileOp &t = b.tops[k];
                int p0 = (int)(std::find(Sp.begin(), Sp.end(), t.pa) - Sp.begin());
                int p1 = t.pb >= 0 ? (int)(std::find(Sp.begin(), Sp.end(), t.pb) - Sp.begin())
                                   : -1;
                left_apply(M, dim, gates[b.members[k]], p0, p1);
            }
            DenseOp o{};
            o.kind = 1;
            o.gidx = (uint32_t)b.members[0];
            o.nq = F;
            for (int j = 0; j < F; j++) o.pos[j] = Sp[j];
            o.mat = (uint32_t)(mats.size() / 2);
            for (int r = 0; r < dim; r++)
                for (int c = 0; c < dim; c++) {
                    mats.push_back((float)M[(size_t)r * dim + c].real());
                    mats.push_back((float)M[(size_t)r * dim + c].imag());
                }
            ops.push_back(o);
        };
        // emit (in opening order) and close every open block touching q
        auto flush_touching = [&](const std::vector<int> &q) {
            std::vector<Block> keep;
            for (Block &b : open) {
                bool t = false;
                for (int p : q) t |= has(b, p);
                if (t) emit_block(b);
                else keep.push_back(std::move(b));
            }
            open.swap(keep);
        };
        for (uint32_t k = 0; k < P.op_count; k++) {
            const size_t gi = order[P.op_first + k];
            const TileOp &t = to[P.op_first + k];
            const bool local = t.pa >= 0 && (!is_gate2(t.kind) || t.pb >= 0);
            std::vector<int> q;
            if (t.pa >= 0) q.push_back(t.pa);
            if (is_gate2(t.kind) && t.pb >= 0) q.push_back(t.pb);
            if (!local) {
                // a factor on a non-tile qubit stays an ordinary tile op
                flush_touching(q);
                emit_tile(gi, t);
                continue;
            }
            std::vector<size_t> touching;
            for (size_t i = 0; i < open.size(); i++) {
                bool tt = false;
                for (int p : q) tt |= has(open[i], p);
                if (tt) touching.push_back(i);
            }
            auto join = [&](Block &b) {
                for (int p : q)
                    if (!has(b, p)) b.S.push_back(p);
                b.members.push_back(gi);
                b.tops.push_back(t);
            };
            if (touching.size() == 1) {
                Block &b = open[touching[0]];
                int add = 0;
                for (int p : q) add += !has(b, p);
                if ((int)b.S.size() + add <= F) {
                    join(b);
                    continue;
                }
            }
            if (touching.empty()) {
                // any open block with room (disjoint, so they commute)
                bool placed = false;
                for (Block &b : open)
                    if ((int)(b.S.size() + q.size()) <= F) {
                        join(b);
                        placed = true;
                        break;
                    }
                if (placed) continue;
            } else {
                flush_touching(q);
            }
            Block nb;
            join(nb);
            open.push_back(std::move(nb));
        }
        for (Block &b : open) emit_block(b);
        open.clear();
        D.op_count = (uint32_t)(ops.size() - D.op_first);
        passes.push_back(D);
    }
}

}  // namespace qslhip
