// plan.cpp -- host side of the gate path: gate descriptors with the
// reference's Fp coefficients, and the planner that packs a gate list into
// LDS tile passes.
#include <cmath>
#include <cstring>

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
        for (size_t i = first; i < end; i++) {
            const GateDesc &g = gates[i];
            TileOp op;
            op.kind = g.kind;
            op.q0 = g.q0;
            op.q1 = is_gate2(g.kind) ? g.q1 : 0;
            op.pa = tile_pos(m, g.q0);
            op.pb = is_gate2(g.kind) ? tile_pos(m, g.q1) : -1;
            op.pad = 0;
            op.c = g.c;
            op.s = g.s;
            ops.push_back(op);
        }
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
