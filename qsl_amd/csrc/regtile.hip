// regtile.hip -- LDS tile passes with the amplitudes held in registers.
//
// A workgroup of NT = 2^tb threads owns one group of 2^K amplitudes, R =
// 2^r = 2^(K-tb) per thread.  A "layout" rho (r of the K tile positions)
// says which tile bits index a thread's registers; the other tb index the
// thread:  tile index of register k of thread t = dep(t, ~rho) | dep(k, rho).
//
// A non-diagonal gate needs its target bit(s) among the register bits; the
// host planner (plan_reg_ops) inserts a remap (registers -> LDS ->
// registers in a new layout covering the next targets) only when they are
// not.  Controls and diagonal factors are predicates on single bits, each
// of which is a register bit (uniform per register: decided on the host,
// a scalar branch on the device), a thread bit (one lane test per op) or a
// qubit outside the tile (uniform per workgroup).  The planner decodes every
// gate into that form, so the device loop does the arithmetic and little
// else.  Remaps address LDS through an XOR-linear swizzle: the offset of
// (thread part ^ register part) is the XOR of the two offsets, the register
// parts come precomputed with the op, and one XOR per element remains.
//
// EX selects the arithmetic: the reference's exact per-gate formulas
// (bit-identical), or for QSLHIP_MODE_REORDER the tolerance-bound forms with
// fused multiply-adds (gate_math.h pair_update_fast).
//
// Loads and stores keep the tile's four lowest positions on lanes
// (coalesced); the registers sit on the positions the pass's first and last
// gates need, or on the top r positions (plan_reg_ops).  A workgroup may
// take several groups and prefetch the next while computing one (k_rtile),
// in either group order (group_of).  The per-amplitude arithmetic is
// gate_math.h's in program order: results are bit-identical to the one-gate
// kernels.
#include <cstring>

#include "rtile.h"

namespace qslhip {

namespace {

// Workgroup b takes groups b, b + gridDim.x, ... (one group when the grid
// covers them all; persistent workgroups when it is the device's resident
// count) and loads its next group into a second register set while it
// computes and stores the current one, so HBM traffic overlaps the gate
// arithmetic.
template <typename T, int NT, int R, bool EX>
__global__ __launch_bounds__(NT) void k_rtile(cplx<T> *__restrict__ st, unsigned n,
                                              RegTileArgs ta, const RegOp *__restrict__ ops,
                                              int nops, uint64_t ngroups)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int A = (int)sizeof(cplx<T>);
    constexpr int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    const uint32_t t = threadIdx.x;
    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const uint64_t outside = full & ~ta.tile_mask;
    const uint64_t tthread = deposit((uint64_t)t, ta.tlow);
    const uint64_t tthread_out = deposit((uint64_t)t, ta.tlow_out);
    uint64_t kpart[R], kpart_out[R];
    uint32_t kt_out[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        uint64_t kp = 0, ko = 0;
#pragma unroll
        for (int i = 0; i < r; i++)
            if ((k >> i) & 1) {
                kp |= ta.kq[i];
                ko |= ta.kq_out[i];
            }
        kpart[k] = kp;
        kpart_out[k] = ko;
        kt_out[k] = reg_tile_bits((uint32_t)k, ta.rho_out);
    }
    // this thread's tile-index bits (zeros at register positions) in the
    // load and the store layout
    const uint32_t jt_in = thread_tile_bits(t, ta.rho_in);
    const uint32_t jt_out = thread_tile_bits(t, ta.rho_out);
    uint64_t g = blockIdx.x;
    if (g >= ngroups) return;
    uint64_t gbase_n = deposit(group_of(g, ngroups, ta.gorder), ta.gvar) | ta.gfix;
    cplx<T> v[R], vn[R];
    reg_load<T, R>(vn, st, gbase_n | tthread, kpart, ta);
    wait_vmcnt<0>();
    for (;;) {
        const uint64_t gbase = gbase_n;
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = vn[k];
        const uint64_t gn = g + gridDim.x;
        const bool more = gn < ngroups;
        if (more) {
            gbase_n = deposit(group_of(gn, ngroups, ta.gorder), ta.gvar) | ta.gfix;
            reg_load<T, R>(vn, st, gbase_n | tthread, kpart, ta);
            // in flight: the previous group's R stores and these R loads;
            // everything older (this group's loads) is complete
            wait_vmcnt<2 * R>();
        } else {
            wait_vmcnt<R>();  // only the previous group's stores may remain
        }
        uint32_t jt = jt_in;
        uint32_t base = lds_swz<A>(jt) * A;
        uint32_t koff[R];
#pragma unroll
        for (int k = 0; k < R; k++) koff[k] = ta.koff0[k];
        // from a basis state, a group without the basis index is zero and
        // stays zero (gates never mix groups): store it as it is.  Program
        // order (EX) runs the ops anyway: they may turn +0 into -0 there
        const int gops = (!EX && ta.basis != ~0ull && (ta.basis & outside) != gbase) ? 0 : nops;
        for (int i = 0; i < gops; i++) {
            const RegOpHead &o = ops[i].h;
            const int code = o.code;
            if (code == RC_REMAP) {
#pragma unroll
                for (int k = 0; k < R; k++)
                    *reinterpret_cast<cplx<T> *>(smem + (base ^ koff[k])) = v[k];
                lds_barrier();
                jt = t;
                for (uint32_t rho = o.rho; rho; rho &= rho - 1) {  // insert zeros at rho
                    const uint32_t lowm = (rho & (0u - rho)) - 1;
                    jt = (jt & lowm) | ((jt & ~lowm) << 1);
                }
                base = lds_swz<A>(jt) * A;
#pragma unroll
                for (int k = 0; k < R; k++) koff[k] = ops[i].koff[k];
#pragma unroll
                for (int k = 0; k < R; k++)
                    v[k] = *reinterpret_cast<const cplx<T> *>(smem + (base ^ koff[k]));
                lds_barrier();
                continue;
            }
            if ((gbase & o.gmask) != o.gmask) continue;  // a control outside the tile is 0
            const T c = (T)o.c, s = (T)o.s;
            const uint32_t tmask = o.tmask;
            const bool sel = tmask != 0;
            const bool tp = (jt & tmask) == tmask;
            const uint32_t kmask = o.kmask;
            const bool full = (o.flags & RF_FULL) != 0;
            switch (code) {
            case RC_PAIR:
                // a control is a thread bit (sel), a register bit (partial
                // kmask) or outside the tile (already tested): never two
                if (sel) dispatch_body<T, R, true, true, EX>(v, o.body, o.ia, kmask, tp, c, s);
                else if (full) dispatch_body<T, R, false, true, EX>(v, o.body, o.ia, kmask, tp, c, s);
                else dispatch_body<T, R, false, false, EX>(v, o.body, o.ia, kmask, tp, c, s);
                break;
            case RC_SWAP:
                switch (o.body) {
                case QSLHIP_PAULI_X:
                    if (o.flags & RF_FSWAP)
                        dispatch_swap<T, R, QSLHIP_PAULI_X, true, EX>(v, o.ia, o.ib, c, s);
                    else dispatch_swap<T, R, QSLHIP_PAULI_X, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                case QSLHIP_ROT_X:
                    dispatch_swap<T, R, QSLHIP_ROT_X, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                case QSLHIP_ROT_Y:
                    dispatch_swap<T, R, QSLHIP_ROT_Y, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                default:
                    dispatch_swap<T, R, QSLHIP_HADAMARD, false, EX>(v, o.ia, o.ib, c, s);
                    break;
                }
                break;
            case RC_ONE:
                if (o.body == QSLHIP_PAULI_Z) {
                    if (sel) {
                        if (full) reg_one<T, R, QSLHIP_PAULI_Z, true, true, EX>(v, kmask, tp, c, s);
                        else reg_one<T, R, QSLHIP_PAULI_Z, true, false, EX>(v, kmask, tp, c, s);
                    } else {
                        if (full) reg_one<T, R, QSLHIP_PAULI_Z, false, true, EX>(v, kmask, tp, c, s);
                        else reg_one<T, R, QSLHIP_PAULI_Z, false, false, EX>(v, kmask, tp, c, s);
                    }
                } else {
                    if (sel) {
                        if (full) reg_one<T, R, QSLHIP_PHASE, true, true, EX>(v, kmask, tp, c, s);
                        else reg_one<T, R, QSLHIP_PHASE, true, false, EX>(v, kmask, tp, c, s);
                    } else {
                        if (full) reg_one<T, R, QSLHIP_PHASE, false, true, EX>(v, kmask, tp, c, s);
                        else reg_one<T, R, QSLHIP_PHASE, false, false, EX>(v, kmask, tp, c, s);
                    }
                }
                break;
            default: {  // RC_RZ
                const bool gneg = (gbase & ops[i].gneg) != 0;
                const uint32_t kneg = o.kneg;
                if (o.tneg) {  // target on a thread bit: one sign per lane
                    const T sl = flip_sign(s, (jt & o.tneg) != 0);
                    if (sel) {
                        if (full) reg_rz<T, R, true, true, true, EX>(v, kmask, kneg, gneg, tp, c, sl);
                        else reg_rz<T, R, true, false, true, EX>(v, kmask, kneg, gneg, tp, c, sl);
                    } else {
                        if (full) reg_rz<T, R, false, true, true, EX>(v, kmask, kneg, gneg, tp, c, sl);
                        else reg_rz<T, R, false, false, true, EX>(v, kmask, kneg, gneg, tp, c, sl);
                    }
                } else {  // a register or group target: one sign per register
                    if (sel) {
                        if (full) reg_rz<T, R, true, true, false, EX>(v, kmask, kneg, gneg, tp, c, s);
                        else reg_rz<T, R, true, false, false, EX>(v, kmask, kneg, gneg, tp, c, s);
                    } else {
                        if (full) reg_rz<T, R, false, true, false, EX>(v, kmask, kneg, gneg, tp, c, s);
                        else reg_rz<T, R, false, false, false, EX>(v, kmask, kneg, gneg, tp, c, s);
                    }
                }
                break;
            }
            }
        }
        if (ta.runs)
            reg_run_sums<T, R, NT>(v, ta.runs, group_of(g, ngroups, ta.gorder), jt_out, kt_out);
        reg_store<T, R>(v, st, gbase | tthread_out, kpart_out, ta);
        if (!more) break;
        g = gn;
    }
}

// Tile positions a non-diagonal body moves amplitudes across.
void targets_of(const TileOp &t, int out[2], int &cnt)
{
    cnt = 0;
    const int cls = pair_class(t.kind), body = body_of(t.kind);
    if (cls == PC_PAIR1 && body != QSLHIP_ROT_Z) out[cnt++] = t.pa;
    else if (cls == PC_CTRL && body != QSLHIP_ROT_Z) out[cnt++] = t.pb;
    else if (cls == PC_SWAPLIKE || cls == PC_FSWAP) {
        out[cnt++] = t.pa;
        out[cnt++] = t.pb;
    }
}

}  // namespace

static uint32_t dep_host(uint32_t v, uint32_t mask)
{
    uint32_t out = 0;
    for (int b = 0; mask; b++, mask &= mask - 1)
        if ((v >> b) & 1u) out |= mask & (0u - mask);
    return out;
}

uint32_t swz_bytes(uint32_t e, int A)
{
    return A == 16 ? lds_swz<16>(e) * 16 : lds_swz<8>(e) * 8;
}


// Host planner: decode the TileOps of one pass (tile width K, 2^tb
// threads, A bytes per amplitude) into register ops.  A remap goes before a
// gate whose targets are not register bits, to the gate's targets plus the
// next gates' targets in order (the set that postpones the following remap
// longest), topped up from the current layout.  Without `lay` the pass
// loads and stores with registers on the top r positions.  With it (and
// tuning free_layout), the load layout is the one the first gate would
// remap to and the store layout the one the last gate leaves, each as long
// as it keeps the tile's four lowest positions on lanes (the coalesced
// runs, and the sampling runs' lanes): up to two remaps fewer.
void plan_reg_ops(const TileOp *ops, uint32_t count, int K, int tb, int A,
                  std::vector<RegOp> &out, RegLayout *lay)
{
    const int r = K - tb, R = 1 << r;
    const uint32_t all_k = (R >= 32) ? ~0u : ((1u << R) - 1);
    const uint32_t rho0 = ((1u << K) - 1) & ~((1u << tb) - 1);
    const uint32_t lanes = 0xFu;  // positions that stay on lanes at load and store
    const bool free_layout = lay != nullptr && tuning().free_layout != 0;
    // the layout a remap before op i picks when the current one is cur
    auto choose = [&](uint32_t i, uint32_t cur) {
        int tg[2], nt;
        targets_of(ops[i], tg, nt);
        uint32_t rho = 0;
        for (int k = 0; k < nt; k++) rho |= 1u << tg[k];
        for (uint32_t j = i + 1; j < count && __builtin_popcount(rho) < r; j++) {
            int tj[2], nj;
            targets_of(ops[j], tj, nj);
            for (int k = 0; k < nj && __builtin_popcount(rho) < r; k++) rho |= 1u << tj[k];
        }
        for (int p = K - 1; p >= 0 && __builtin_popcount(rho) < r; p--)
            if ((cur >> p) & 1u) rho |= 1u << p;
        for (int p = K - 1; p >= 0 && __builtin_popcount(rho) < r; p--) rho |= 1u << p;
        return rho;
    };
    uint32_t rho_in = rho0;
    if (free_layout && count > 0) {
        int tg[2], nt;
        targets_of(ops[0], tg, nt);
        uint32_t need = 0;
        for (int k = 0; k < nt; k++) need |= 1u << tg[k];
        if (need & ~rho0) {
            const uint32_t rho = choose(0, rho0);
            if (!(rho & lanes)) rho_in = rho;
        }
    }
    uint32_t cur = rho_in;
    auto slot = [&](int pos) { return __builtin_popcount(cur & ((1u << pos) - 1)); };
    auto kset = [&](int pos) {  // registers whose bit at tile position pos is 1
        const int sl = slot(pos);
        uint32_t m = 0;
        for (int k = 0; k < R; k++)
            if ((k >> sl) & 1) m |= 1u << k;
        return m;
    };
    auto kset_pos_slot = [&](int sl) {  // registers whose register bit sl is 1
        uint32_t m = 0;
        for (int k = 0; k < R; k++)
            if ((k >> sl) & 1) m |= 1u << k;
        return m;
    };
    auto remap = [&](uint32_t rho) {
        RegOp m{};
        m.h.code = RC_REMAP;
        m.h.rho = rho;
        for (int k = 0; k < R; k++) m.koff[k] = swz_bytes(dep_host((uint32_t)k, rho), A);
        out.push_back(m);
        cur = rho;
    };
    auto pred = [&](RegOp &op, int pos, uint32_t q) {  // require qubit q (tile pos) == 1
        if (pos < 0) op.h.gmask |= 1ull << q;
        else if ((cur >> pos) & 1u) op.h.kmask &= kset(pos);
        else op.h.tmask |= 1u << pos;
    };
    auto neg = [&](RegOp &op, int pos, uint32_t q) {  // rotateZ half follows qubit q
        if (pos < 0) op.gneg = 1ull << q;
        else if ((cur >> pos) & 1u) op.h.kneg = kset(pos);
        else op.h.tneg = 1u << pos;
    };
    for (uint32_t i = 0; i < count; i++) {
        const TileOp &t = ops[i];
        int tg[2], nt;
        targets_of(t, tg, nt);
        uint32_t need = 0;
        for (int k = 0; k < nt; k++) need |= 1u << tg[k];
        if ((need & ~cur) != 0) remap(choose(i, cur));
        RegOp op{};
        RegOpHead &o = op.h;
        o.c = t.c;
        o.s = t.s;
        o.kmask = all_k;
        const int body = body_of(t.kind);
        switch (t.kind) {
        case QSLHIP_PAULI_X: case QSLHIP_PAULI_Y: case QSLHIP_HADAMARD:
        case QSLHIP_ROT_X: case QSLHIP_ROT_Y:
            o.code = RC_PAIR;
            o.body = body;
            o.ia = slot(t.pa);
            o.kmask = all_k & ~kset(t.pa);
            break;
        case QSLHIP_ROT_Z:
            o.code = RC_RZ;
            neg(op, t.pa, t.q0);
            break;
        case QSLHIP_PAULI_Z: case QSLHIP_PHASE:
            o.code = RC_ONE;
            o.body = body;
            pred(op, t.pa, t.q0);
            break;
        case QSLHIP_CNOT: case QSLHIP_CY: case QSLHIP_CROT_X: case QSLHIP_CROT_Y:
        case QSLHIP_CHADAMARD:
            o.code = RC_PAIR;
            o.body = body;
            o.ia = slot(t.pb);
            o.kmask = all_k & ~kset(t.pb);
            pred(op, t.pa, t.q0);
            break;
        case QSLHIP_CROT_Z:
            o.code = RC_RZ;
            pred(op, t.pa, t.q0);
            neg(op, t.pb, t.q1);
            break;
        case QSLHIP_CZ: case QSLHIP_CPHASE:
            o.code = RC_ONE;
            o.body = body;
            pred(op, t.pa, t.q0);
            pred(op, t.pb, t.q1);
            break;
        default:  // swap-like
            o.code = RC_SWAP;
            o.body = body;
            o.ia = slot(t.pa);
            o.ib = slot(t.pb);
            if (t.kind == QSLHIP_FSWAP) o.flags |= RF_FSWAP;
            break;
        }
        const bool full = o.code == RC_PAIR ? o.kmask == (all_k & ~kset_pos_slot(o.ia))
                                            : o.kmask == all_k;
        if (full) o.flags |= RF_FULL;
        // the kernel's pair dispatch relies on it: one control qubit is a
        // thread bit, a register bit or outside the tile, never two of these
        QH_REQUIRE(!(o.code == RC_PAIR && o.tmask && !full), QSLHIP_EINVAL,
                   "pair op with both a thread and a register predicate");
        out.push_back(op);
    }
    uint32_t rho_out = rho0;
    if (free_layout && !(cur & lanes)) rho_out = cur;
    if (cur != rho_out) remap(rho_out);
    if (lay) {
        lay->in = rho_in;
        lay->out = rho_out;
    }
}

RegTileArgs make_tile_args(unsigned n, const TilePass &p, int tb, int A, const RegLayout &lay)
{
    const int K = p.logt, R = 1 << (K - tb);
    const uint32_t rho0 = ((1u << K) - 1) & ~((1u << tb) - 1);
    RegTileArgs ta{};
    ta.rho_in = lay.in ? lay.in : rho0;
    ta.rho_out = lay.out ? lay.out : rho0;
    QH_REQUIRE(__builtin_popcount(ta.rho_in) == K - tb && __builtin_popcount(ta.rho_out) == K - tb &&
                   (ta.rho_in >> K) == 0 && (ta.rho_out >> K) == 0,
               QSLHIP_EINVAL, "register layout does not fit the tile");
    ta.basis = ~0ull;
    ta.runs = nullptr;
    ta.tile_mask = p.tile_mask;
    ta.gvar = ((n >= 64) ? ~0ull : ((1ull << n) - 1)) & ~p.tile_mask;
    ta.gfix = 0;
    ta.zmask = 0;
    ta.zval = 0;
    ta.stream = n >= (unsigned)tuning().stream_min_qubits ? (uint32_t)tuning().pass_stream : 0u;
    int pos = 0, kin = 0, kout = 0;
    for (unsigned q = 0; q < 64; q++) {
        if (!((p.tile_mask >> q) & 1ull)) continue;
        if (pos < 32 && ((ta.rho_in >> pos) & 1u)) ta.kq[kin++ & 3] = 1ull << q;
        else ta.tlow |= 1ull << q;
        if (pos < 32 && ((ta.rho_out >> pos) & 1u)) ta.kq_out[kout++ & 3] = 1ull << q;
        else ta.tlow_out |= 1ull << q;
        pos++;
    }
    QH_REQUIRE(pos == K, QSLHIP_EINVAL, "tile mask does not hold the tile width");
    QH_REQUIRE(K - tb >= 2 && K - tb <= 4, QSLHIP_EINVAL, "2..4 register bits per thread");
    for (int k = 0; k < R; k++) ta.koff0[k] = swz_bytes(reg_tile_bits((uint32_t)k, ta.rho_in), A);
    QH_REQUIRE(((size_t)A << K) <= 64 * 1024, QSLHIP_EINVAL,
               "register tile needs more than 64 KiB of LDS");
    QH_REQUIRE(n >= (unsigned)K, QSLHIP_EINVAL, "tile wider than the state");
    return ta;
}

RegTileArgs pass_args(unsigned n, const TilePass &p, int tb, int A, const PassExtras &x,
                      uint64_t *groups)
{
    RegTileArgs ta = make_tile_args(n, p, tb, A, x.lay);
    ta.basis = x.basis;
    ta.runs = x.runs;
    QH_REQUIRE((x.zmask & ~p.tile_mask) == 0 && (x.zval & ~x.zmask) == 0, QSLHIP_EINVAL,
               "zero mask outside the tile");
    ta.zmask = x.zmask;
    ta.zval = x.zval;
    ta.gorder = x.gorder;
    if (x.gvar != ~0ull) {
        QH_REQUIRE((x.gvar & ~ta.gvar) == 0 && (x.gfix & ~ta.gvar) == 0 && (x.gvar & x.gfix) == 0,
                   QSLHIP_EINVAL, "group subset outside the pass's outside qubits");
        QH_REQUIRE(!x.runs || x.gvar == ta.gvar, QSLHIP_EINVAL, "run sums need every group");
        ta.gvar = x.gvar;
        ta.gfix = x.gfix;
    }
    *groups = 1ull << __builtin_popcountll(ta.gvar);
    if (*groups < 8) ta.gorder = 0;
    return ta;
}

uint64_t reg_grid(uint64_t groups, int gpw)
{
    // several groups per workgroup only when groups are plentiful
    if (gpw < 1 || groups < (uint64_t)gpw * 8192) gpw = 1;
    return (groups + (uint64_t)gpw - 1) / (uint64_t)gpw;
}

namespace {

template <typename T, int NT, int R, bool EX>
void launch_one(qslhip_state *h, void *amps, const TilePass &p, const RegOp *dops, uint32_t nops,
                int gpw, const PassExtras &x)
{
    constexpr int A = (int)sizeof(cplx<T>);
    constexpr int tb = NT == 256 ? 8 : 9;
    constexpr int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    const int K = tb + r;
    uint64_t groups = 0;
    const RegTileArgs ta = pass_args(h->n, p, tb, A, x, &groups);
    const size_t shmem = (size_t)A << K;
    uint64_t grid;
    if (gpw > 0) {
        grid = reg_grid(groups, gpw);
    } else {  // persistent: as many workgroups as the device holds at once
        static int resident = 0;  // per instantiation
        if (!resident) {
            int per_cu = 0, dev = 0, cus = 0;
            QH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rtile<T, NT, R, EX>,
                                                                  NT, shmem));
            QH_CHECK(hipGetDevice(&dev));
            QH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            resident = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        }
        grid = groups < (uint64_t)resident ? groups : (uint64_t)resident;
    }
    RegTileArgs targ = ta;
    if (grid % 8) targ.gorder = 0;
    hipLaunchKernelGGL((k_rtile<T, NT, R, EX>), dim3((unsigned)grid), dim3(NT), shmem, h->stream,
                       static_cast<cplx<T> *>(amps), h->n, targ, dops, (int)nops, groups);
    QH_CHECK(hipGetLastError());
}

template <typename T, bool EX>
void launch_t(qslhip_state *h, void *amps, const TilePass &p, int nt, const RegOp *dops,
              uint32_t nops, int gpw, const PassExtras &x)
{
    const int r = p.logt - (nt == 512 ? 9 : 8);
    if (nt == 256 && r == 2) launch_one<T, 256, 4, EX>(h, amps, p, dops, nops, gpw, x);
    else if (nt == 256 && r == 3) launch_one<T, 256, 8, EX>(h, amps, p, dops, nops, gpw, x);
    else if (nt == 256 && r == 4) launch_one<T, 256, 16, EX>(h, amps, p, dops, nops, gpw, x);
    else if (nt == 512 && r == 3) launch_one<T, 512, 8, EX>(h, amps, p, dops, nops, gpw, x);
    else if (nt == 512 && r == 4) launch_one<T, 512, 16, EX>(h, amps, p, dops, nops, gpw, x);
    else throw Error(QSLHIP_EINVAL, "no register-tile kernel for this tile width");
}

}  // namespace

bool reg_tile_ok(int logt, int threads, int prec)
{
    const int r = logt - (threads == 512 ? 9 : 8);
    if (threads != 256 && threads != 512) return false;
    if (r < 2 || r > 4 || (threads == 512 && r == 2)) return false;
    return ((size_t)(prec == 8 ? 16 : 8) << logt) <= 64 * 1024;
}

static std::vector<PassExtras> pass_extras(unsigned n, const std::vector<TilePass> &passes,
                                           bool exact, uint64_t basis, double *runs)
{
    std::vector<PassExtras> xs(passes.size());
    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    if (!xs.empty()) xs[0].basis = basis;  // only the first pass reads the state
    if (basis != ~0ull && !exact) {
        // Reordered mode from a basis state.  After passes 0..p-1 the state
        // can be nonzero only where it agrees with the basis on every qubit
        // no tile has held yet (gates never mix groups), so pass p runs only
        // the groups that agree there.  When every qubit is in some tile,
        // nothing is stored before it is computed: the first pass runs the
        // basis index's group alone, and a later pass reads only the
        // amplitudes an earlier one wrote (those agreeing with the basis on
        // its tile qubits no earlier tile held) and takes the rest as +0.
        // Otherwise the first pass stores every group, the zero ones
        // without their ops, and nothing is left unwritten.  Program order
        // runs every group: gates applied to zeros may leave -0 there.
        uint64_t all = 0;
        for (const TilePass &p : passes) all |= p.tile_mask;
        const bool fill = (all & full) != full;
        uint64_t support = 0;  // qubits the amplitudes may vary in so far
        for (size_t i = 0; i < passes.size(); i++) {
            PassExtras &x = xs[i];
            const uint64_t tile = passes[i].tile_mask, outside = full & ~tile;
            if (i > 0 || !fill) {
                x.gvar = outside & support;
                x.gfix = basis & outside & ~support;
            }
            if (i > 0 && !fill) {
                x.zmask = tile & ~support;
                x.zval = basis & x.zmask;
            }
            support |= tile;
        }
    }
    // the last pass writes the run sums when it covers every group
    if (!xs.empty() && (xs.back().gvar == ~0ull ||
                        xs.back().gvar == (full & ~passes.back().tile_mask)))
        xs.back().runs = runs;
    return xs;
}

namespace {

// The group order of a pass (RegTileArgs::gorder): tuning group_order 0 or 1
// for every pass; 2 (default) the XCD sweep for passes that read only part
// of each tile (from a basis state: the write-heavy late passes, measured
// 1-2% faster that way at 30 qubits) and the plain order for passes that
// read whole tiles (measured 2-3% faster plain).
uint32_t choose_group_order(const PassExtras &x)
{
    const int mode = tuning().group_order;
    if (mode == 2) return x.zmask != 0 ? 1u : 0u;
    return mode == 1 ? 1u : 0u;
}

}  // namespace

bool launch_reg_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                       const std::vector<TileOp> &ops, void *amps, int threads, int gpw,
                       bool exact, uint64_t basis, double *runs)
{
    const int tb = threads == 512 ? 9 : 8;
    const int A = h->prec == 8 ? 16 : 8;
    QH_REQUIRE(!runs || (!passes.empty() && (passes.back().tile_mask & 7) == 7), QSLHIP_EINVAL,
               "run sums need qubits 0..2 in the last tile");
    std::vector<RegOp> rops;
    std::vector<uint32_t> first(passes.size() + 1);
    std::vector<RegLayout> lays(passes.size());
    for (size_t i = 0; i < passes.size(); i++) {
        QH_REQUIRE(reg_tile_ok(passes[i].logt, threads, h->prec), QSLHIP_EINVAL,
                   "no register-tile kernel for this tile width");
        first[i] = (uint32_t)rops.size();
        plan_reg_ops(ops.data() + passes[i].op_first, passes[i].op_count, passes[i].logt, tb, A,
                     rops, &lays[i]);
    }
    first[passes.size()] = (uint32_t)rops.size();
    // coefficient tables of the compiled passes: (c, s) per gate op
    std::vector<double> coef;
    std::vector<size_t> cfirst(passes.size());
    if (tuning().jit)
        for (size_t i = 0; i < passes.size(); i++) {
            cfirst[i] = coef.size();
            jit_coefficients(rops.data() + first[i], first[i + 1] - first[i], coef);
        }
    const size_t bytes = rops.size() * sizeof(RegOp);
    const size_t cbytes = coef.size() * sizeof(double);
    // The pinned staging buffer may still be read by the previous copy.
    if (h->staged) QH_CHECK(hipEventSynchronize(h->staged));
    h->pin.ensure(bytes + cbytes);
    h->optab.ensure(bytes + cbytes);
    std::memcpy(h->pin.ptr, rops.data(), bytes);
    if (cbytes) std::memcpy(static_cast<char *>(h->pin.ptr) + bytes, coef.data(), cbytes);
    QH_CHECK(hipMemcpyAsync(h->optab.ptr, h->pin.ptr, bytes + cbytes, hipMemcpyHostToDevice,
                            h->stream));
    QH_CHECK(hipEventRecord(h->staged, h->stream));
    const RegOp *dops = static_cast<const RegOp *>(h->optab.ptr);
    const double *dcoef =
        reinterpret_cast<const double *>(static_cast<char *>(h->optab.ptr) + bytes);
    std::vector<PassExtras> xs = pass_extras(h->n, passes, exact, basis, runs);
    for (size_t i = 0; i < passes.size(); i++) xs[i].lay = lays[i];
    for (size_t i = 0; i < passes.size(); i++) {
        const uint32_t n_ops = first[i + 1] - first[i];
        const TilePass &p = passes[i];
        xs[i].gorder = choose_group_order(xs[i]);
        const PassExtras &x = xs[i];
        if (cbytes && jit_launch_pass(h, amps, p, threads, gpw, exact, rops.data() + first[i],
                                      n_ops, dcoef + cfirst[i], x))
            continue;
        const RegOp *d = dops + first[i];
        if (h->prec == 8) {
            if (exact) launch_t<double, true>(h, amps, p, threads, d, n_ops, gpw, x);
            else launch_t<double, false>(h, amps, p, threads, d, n_ops, gpw, x);
        } else {
            if (exact) launch_t<float, true>(h, amps, p, threads, d, n_ops, gpw, x);
            else launch_t<float, false>(h, amps, p, threads, d, n_ops, gpw, x);
        }
    }
    return xs.empty() || xs.back().runs != nullptr;
}

}  // namespace qslhip
