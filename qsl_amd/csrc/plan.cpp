// plan.cpp -- host side of the gate path: gate descriptors with the
// reference's Fp coefficients, and the planner that packs a gate list into
// LDS tile passes.
#include <cmath>
#include <cstring>
#include <set>

#include "gate_math.h"

namespace qslhip {

bool is_gate1(int kind) { return kind >= 0 && kind <= 7; }
bool is_gate2(int kind) { return kind >= 16 && kind <= 28; }

// Gates that keep the Hamming weight (the NP API, include/qsl/concepts.hpp:67-86).
bool is_np_kind(int kind)
{
    switch (kind) {
    case QSLHIP_PAULI_Z: case QSLHIP_PHASE: case QSLHIP_ROT_Z:
    case QSLHIP_CZ: case QSLHIP_CROT_Z: case QSLHIP_CPHASE:
    case QSLHIP_SWAP: case QSLHIP_FSWAP: case QSLHIP_NP_ROT_X:
    case QSLHIP_NP_ROT_Y: case QSLHIP_NP_HADAMARD:
        return true;
    default:
        return false;
    }
}

// cos and sin of one argument the way the reference's g++ -O2 build gets
// them: GCC folds the adjacent std::cos(x), std::sin(x) calls of every gate
// into one glibc sincos(x) call, whose cosine can differ from cos(x) by one
// ulp.  Calling sincos here keeps the coefficients bit-identical.
static inline void fp_sincos(double x, double &c, double &s) { ::sincos(x, &s, &c); }
static inline void fp_sincos(float x, float &c, float &s) { ::sincosf(x, &s, &c); }

// Coefficients exactly as the reference computes them, in Fp:
//   rotations:  Fp cos = std::cos(angle/2), sin = std::sin(angle/2)
//               (src/qubits/default/gates.cpp:124-125)
//   phase:      qsl::complex<Fp> phase{std::cos(angle), std::sin(angle)}
//               (gates.cpp:106, 440)
template <typename Fp>
static void coeffs(int kind, Fp angle, double &c, double &s)
{
    Fp fc, fs;
    switch (kind) {
    case QSLHIP_PHASE: case QSLHIP_CPHASE:
        fp_sincos(angle, fc, fs);
        c = (double)fc;
        s = (double)fs;
        break;
    case QSLHIP_ROT_X: case QSLHIP_ROT_Y: case QSLHIP_ROT_Z:
    case QSLHIP_CROT_X: case QSLHIP_CROT_Y: case QSLHIP_CROT_Z:
    case QSLHIP_NP_ROT_X: case QSLHIP_NP_ROT_Y:
        fp_sincos((Fp)(angle / 2), fc, fs);
        c = (double)fc;
        s = (double)fs;
        break;
    default:
        c = 0;
        s = 0;
        break;
    }
}

GateDesc make_desc(int prec, int kind, unsigned q0, unsigned q1, double angle)
{
    GateDesc g{kind, q0, q1, 0.0, 0.0};
    if (prec == 8) coeffs<double>(kind, angle, g.c, g.s);
    else coeffs<float>(kind, (float)angle, g.c, g.s);
    return g;
}

// Qubits that must sit inside the tile for a gate: the qubits a
// non-diagonal body moves amplitudes across.  Diagonal gates (pauliZ,
// phase, rotateZ, controlZ, controlPhase, controlRotateZ) and controls act
// per amplitude and can read a non-tile qubit from the group base.
static uint64_t needs_local(const GateDesc &g)
{
    switch (g.kind) {
    case QSLHIP_PAULI_X: case QSLHIP_PAULI_Y: case QSLHIP_HADAMARD:
    case QSLHIP_ROT_X: case QSLHIP_ROT_Y:
        return 1ull << g.q0;
    case QSLHIP_CNOT: case QSLHIP_CY: case QSLHIP_CROT_X: case QSLHIP_CROT_Y:
    case QSLHIP_CHADAMARD:
        return 1ull << g.q1;
    case QSLHIP_SWAP: case QSLHIP_FSWAP: case QSLHIP_NP_ROT_X:
    case QSLHIP_NP_ROT_Y: case QSLHIP_NP_HADAMARD:
        return (1ull << g.q0) | (1ull << g.q1);
    default:
        return 0;
    }
}

static int tile_pos(uint64_t mask, unsigned q)
{
    if (!((mask >> q) & 1ull)) return -1;
    return __builtin_popcountll(mask & ((1ull << q) - 1));
}

// The TileOp of gate g in a pass whose tile qubits are `mask`.
static TileOp tile_op_of(const GateDesc &g, uint64_t mask)
{
    TileOp op;
    op.kind = g.kind;
    op.q0 = g.q0;
    op.q1 = is_gate2(g.kind) ? g.q1 : 0;
    op.pa = tile_pos(mask, g.q0);
    op.pb = is_gate2(g.kind) ? tile_pos(mask, g.q1) : -1;
    op.pad = 0;
    op.c = g.c;
    op.s = g.s;
    return op;
}

// Greedy packing in program order: a gate joins the open pass while the
// union of tile qubits stays within K; otherwise the pass is closed,
// topped up to K qubits with the lowest free ones (longer contiguous runs
// coalesce better), and a new one opens.  Gate order inside a pass is kept,
// which is all the bit-exactness argument in fused.hip needs.
void plan_tile_passes(unsigned n, int prec, const std::vector<GateDesc> &gates,
                      int logt, int low_bits, std::vector<TilePass> &passes,
                      std::vector<TileOp> &ops)
{
    (void)prec;
    int K = logt < (int)n ? logt : (int)n;
    int c = low_bits < K ? low_bits : K;
    if (K - c > 8) c = K - 8;  // at most 8 scattered tile qubits
    const uint64_t base = (c >= 64) ? ~0ull : ((1ull << c) - 1);

    uint64_t mask = base;
    size_t first = 0;  // index into gates of the open pass
    auto close = [&](size_t end) {
        if (end == first) return;
        uint64_t m = mask;
        for (unsigned q = 0; q < n && __builtin_popcountll(m) < K; q++) m |= 1ull << q;
        TilePass p;
        p.logt = K;
        p.tile_mask = m;
        p.op_first = (uint32_t)ops.size();
        for (size_t i = first; i < end; i++) ops.push_back(tile_op_of(gates[i], m));
        p.op_count = (uint32_t)(ops.size() - p.op_first);
        passes.push_back(p);
    };
    for (size_t i = 0; i < gates.size(); i++) {
        uint64_t need = needs_local(gates[i]);
        if ((need & ~mask) == 0) continue;
        if (__builtin_popcountll(mask | need) <= K) {
            mask |= need;
            continue;
        }
        close(i);
        first = i;
        mask = base | need;
    }
    close(gates.size());
}

}  // namespace qslhip

namespace qslhip {

// Commutation-aware packing (QSLHIP_MODE_REORDER).  Two gates commute, and
// may run in either order, when they act on disjoint qubits or are both
// diagonal in the computational basis (pauliZ, phase, rotateZ, controlZ,
// controlPhase, controlRotateZ).  Every other pair keeps program order: a
// gate depends on each earlier gate sharing a qubit, back to the last
// non-diagonal gate on that qubit.  A pass takes, in index order, every
// ready gate whose non-diagonal targets are already tile qubits; when none
// is left it grows the tile by a ready gate that still fits (the one adding
// the fewest qubits no earlier tile held, the earliest of those), and it
// closes when no ready gate fits.  The result equals the program-order
// product in exact arithmetic; in floating point the amplitudes differ by
// rounding only.  `order` receives the gate index of every TileOp.
void plan_tile_passes_reorder(unsigned n, const std::vector<GateDesc> &gates, int logt,
                              int low_bits, std::vector<TilePass> &passes,
                              std::vector<TileOp> &ops, std::vector<uint32_t> &order)
{
    const int K = logt < (int)n ? logt : (int)n;
    int c = low_bits < K ? low_bits : K;
    if (K - c > 8) c = K - 8;
    const uint64_t base = (c >= 64) ? ~0ull : ((1ull << c) - 1);
    const size_t m = gates.size();
    std::vector<uint64_t> need(m);
    std::vector<std::vector<uint32_t>> succ(m);
    std::vector<uint32_t> npred(m, 0), stamp(m, 0);
    std::vector<std::vector<std::pair<uint32_t, bool>>> last(n);
    for (size_t i = 0; i < m; i++) {
        const GateDesc &g = gates[i];
        need[i] = needs_local(g);
        const bool diag = need[i] == 0;
        const unsigned qs[2] = {g.q0, g.q1};
        const int nq = is_gate2(g.kind) ? 2 : 1;
        for (int k = 0; k < nq; k++) {
            auto &lq = last[qs[k]];
            for (const auto &e : lq) {
                if (diag && e.second) continue;  // diagonal gates commute
                if (stamp[e.first] == i + 1) continue;
                stamp[e.first] = (uint32_t)(i + 1);
                succ[e.first].push_back((uint32_t)i);
                npred[i]++;
            }
            if (diag) lq.emplace_back((uint32_t)i, true);
            else lq.assign(1, {(uint32_t)i, false});
        }
    }
    std::set<uint32_t> ready;
    for (size_t i = 0; i < m; i++)
        if (!npred[i]) ready.insert((uint32_t)i);
    size_t remaining = m;
    std::vector<std::vector<uint32_t>> takes;
    std::vector<uint64_t> masks;
    // Qubits some earlier tile held.  A circuit applied to a basis state is
    // zero outside the groups that agree with it on every other qubit, and
    // its passes skip those groups: growing the tile with the qubits the
    // earlier tiles already held keeps the later passes sparse longer.
    uint64_t support = base;
    while (remaining) {
        uint64_t mask = base;
        std::vector<uint32_t> take;
        for (;;) {
            for (auto it = ready.begin(); it != ready.end();) {
                const uint32_t i = *it;
                if (need[i] & ~mask) {
                    ++it;
                    continue;
                }
                it = ready.erase(it);
                take.push_back(i);
                remaining--;
                for (uint32_t s : succ[i])
                    if (--npred[s] == 0) ready.insert(s);  // s > i: met later in this sweep
            }
            // grow by the ready gate that fits and brings the fewest qubits
            // no earlier tile held, the earliest of those
            int best = -1, best_new = 65;
            for (uint32_t i : ready) {
                if (__builtin_popcountll(mask | need[i]) > K) continue;
                const int nw = __builtin_popcountll(need[i] & ~mask & ~support);
                if (nw < best_new) {
                    best_new = nw;
                    best = (int)i;
                    if (nw == 0) break;
                }
            }
            if (best < 0) break;
            mask |= need[best];
        }
        QH_REQUIRE(!take.empty(), QSLHIP_EINVAL, "tile too narrow for a gate's targets");
        // top up with the lowest qubits an earlier tile held, then the lowest
        uint64_t mk = mask;
        for (unsigned q = 0; q < n && __builtin_popcountll(mk) < K; q++)
            if ((support >> q) & 1ull) mk |= 1ull << q;
        for (unsigned q = 0; q < n && __builtin_popcountll(mk) < K; q++) mk |= 1ull << q;
        support |= mk;
        takes.push_back(std::move(take));
        masks.push_back(mk);
    }
    // Balance: the sweep puts every ready diagonal gate into the earliest
    // pass, which makes the first pass the longest.  A diagonal gate needs
    // no tile qubit, so it may run in any later pass up to the first one
    // holding one of its successors; walking the gates backwards, each moves
    // to the lightest such pass while that one is lighter by two or more.
    // It runs first there: its predecessors are all in earlier passes, and
    // every op of that pass that it does not commute with is a successor.
    const size_t np = takes.size();
    std::vector<uint32_t> pass_of(m), count(np);
    for (size_t p = 0; p < np; p++) {
        count[p] = (uint32_t)takes[p].size();
        for (uint32_t i : takes[p]) pass_of[i] = (uint32_t)p;
    }
    std::vector<char> moved(m, 0);
    for (size_t i = m; i-- > 0;) {
        if (need[i]) continue;
        const uint32_t p = pass_of[i];
        uint32_t lim = (uint32_t)np - 1;
        for (uint32_t s : succ[i]) lim = pass_of[s] < lim ? pass_of[s] : lim;
        uint32_t best = p;
        for (uint32_t q = p + 1; q <= lim && q < np; q++)
            if (count[q] < count[best]) best = q;
        if (best != p && count[best] + 1 < count[p]) {
            pass_of[i] = best;
            count[p]--;
            count[best]++;
            moved[i] = 1;
        }
    }
    std::vector<std::vector<uint32_t>> front(np);  // moved-in gates, program order
    for (size_t i = 0; i < m; i++)
        if (moved[i]) front[pass_of[i]].push_back((uint32_t)i);
    for (size_t p = 0; p < np; p++) {
        TilePass tp;
        tp.logt = K;
        tp.tile_mask = masks[p];
        tp.op_first = (uint32_t)ops.size();
        auto emit = [&](uint32_t i) {
            ops.push_back(tile_op_of(gates[i], masks[p]));
            order.push_back(i);
        };
        for (uint32_t i : front[p]) emit(i);
        for (uint32_t i : takes[p])
            if (pass_of[i] == p) emit(i);
        tp.op_count = (uint32_t)(ops.size() - tp.op_first);
        passes.push_back(tp);
    }
}

}  // namespace qslhip
