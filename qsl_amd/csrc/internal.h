// internal.h -- shared declarations of the qslhip library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "qslhip.h"
#include "cplx.h"

namespace qslhip {

// Thrown inside the library, turned into a status code at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define QH_CHECK(expr)                                                        \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::qslhip::Error(e_ == hipErrorOutOfMemory ? QSLHIP_ENOMEM   \
                                                            : QSLHIP_EDEVICE, \
                                  std::string(#expr) + ": " +                 \
                                      hipGetErrorString(e_));                 \
    } while (0)

#define QH_REQUIRE(cond, code, msg)                                           \
    do {                                                                      \
        if (!(cond)) throw ::qslhip::Error((code), (msg));                    \
    } while (0)

// Largest qubit count a handle accepts: 2^36 FP32 amplitudes = 512 GiB is
// beyond one MI355X (288 GB), so allocation fails first; the bound keeps
// every shift in range.
constexpr unsigned kMaxQubits = 40;

// Device scratch that grows on demand.  It owns its allocation: the
// destructor releases it, so a buffer added to a handle or batch cannot be
// leaked by a release list that misses it.  Not copyable.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void ensure(size_t b);
    void release();
};

// Host pinned scratch (results of reductions, staging of gate tables),
// owned the same way.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void ensure(size_t b);
    void release();
};


// A gate prepared for the device: kind, qubits, coefficients in the
// handle's precision (already rounded from the host Fp computation).
struct GateDesc {
    int kind;
    unsigned q0, q1;
    double c, s;
};

// One op inside an LDS tile pass (see fused.hip).  pa/pb are tile bit
// positions of q0/q1, or -1 when the qubit is outside the tile (then its
// value is fixed per group and read from the group's base index).
struct TileOp {
    int32_t kind;
    int32_t pa, pb;
    uint32_t q0, q1;
    uint32_t pad;
    double c, s;
};

// A planned tile pass: which qubits are in the tile, and a slice of the
// circuit's TileOp table.
struct TilePass {
    int logt;              // tile has 2^logt amplitudes
    uint64_t tile_mask;    // qubits in the tile
    uint32_t op_first, op_count;
};

struct NPTables;  // np_compact.hip

}  // namespace qslhip

struct qslhip_state {
    int device = 0;
    unsigned n = 0;        // qubits
    unsigned k = 0;        // number of ones (NP handles)
    bool np = false;       // compact number-preserving layout
    int prec = 8;          // 8 double, 4 float
    uint64_t namps = 0;    // device amplitudes
    uint64_t cap = 0;      // amplitudes the allocation holds (>= namps)
    bool resizable = false;  // qubits may be added and removed (Resize backend)
    void *amps = nullptr;  // device state
    // A basis state that reset/setBasis recorded but has not written yet
    // (kNoBasis: `amps` holds the state).  Every reader writes it first; a
    // register-tile gate pass writes it in place of its first load.
    static constexpr uint64_t kNoBasis = ~0ull;
    uint64_t pending_basis = kNoBasis;
    // The sampling sums of 8-amplitude runs (sample_run_sums) hold the
    // current state: a gate pass wrote them, in the group order of its tile
    // (sums_tile, 2^sums_logt amplitudes), and nothing has changed it since.
    bool sums_valid = false;
    uint64_t sums_tile = 0;
    int sums_logt = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    qslhip::DevBuf red;      // reduction partials + result slots
    qslhip::DevBuf cdf;      // sampling prefix sums
    qslhip::DevBuf shots;    // per-shot uniforms / outcomes
    qslhip::DevBuf optab;    // tile-pass op tables
    qslhip::PinnedBuf pin;   // pinned host staging
    hipEvent_t staged = nullptr;  // last copy out of `pin` has completed
    // sampling: host copies on a stream of their own, overlapping kernels
    hipStream_t copy_stream = nullptr;
    hipEvent_t xev[5] = {};
    qslhip::NPTables *npt = nullptr;
    size_t amp_bytes() const { return prec == 8 ? 16 : 8; }
};

namespace qslhip {

// ---- launchers (defined in the .hip files) -----------------------------------
// amps: apply to this buffer (same size and precision) instead of h->amps
void launch_gate(qslhip_state *h, const GateDesc &g, void *amps = nullptr);
// exact: the reference's per-gate arithmetic (bit-identical to one kernel
// per gate); false: tolerance-bound fused multiply-adds (QSLHIP_MODE_REORDER)
// basis: the state is that basis state, not yet written (pending_basis);
// the first register-tile pass synthesizes it, other paths write it first.
// runs: where the last register-tile pass may write the sampling sums of
// the result (sample_run_sums); returns whether it did.
bool launch_tile_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                        const std::vector<TileOp> &ops, void *amps = nullptr,
                        bool exact = true, uint64_t basis = qslhip_state::kNoBasis,
                        double *runs = nullptr);
// the same passes with the amplitudes held in registers, `threads` (256 or
// 512) per workgroup (regtile.hip); reg_tile_ok says whether a kernel
// exists for the tile width
// Returns whether the last pass wrote the run sums.
bool launch_reg_passes(qslhip_state *h, const std::vector<TilePass> &passes,
                       const std::vector<TileOp> &ops, void *amps, int threads, int gpw,
                       bool exact, uint64_t basis = qslhip_state::kNoBasis,
                       double *runs = nullptr);
bool reg_tile_ok(int logt, int threads, int prec);
void launch_fill_basis(qslhip_state *h, uint64_t index);
double device_prob(qslhip_state *h, unsigned t, unsigned outcome);
double device_norm2(qslhip_state *h);
void launch_collapse(qslhip_state *h, unsigned t, unsigned outcome,
                     double factor);
// where a gate pass writes the sampling sums of the state's 8-amplitude
// runs (null when the state is too small for the sampler's layout, or the
// handle has not sampled yet and holds no buffer for them)
double *sample_run_sums(qslhip_state *h);
void device_sample(qslhip_state *h, const double *u, uint64_t nshots,
                   uint64_t *out, bool strict = false);
// The same outcomes with every buffer in a static device array (no
// allocation): full-vector handles of at most kNoAllocMaxAmps amplitudes,
// at most kNoAllocMaxShots uniforms per call; sampleAll semantics.
constexpr uint64_t kNoAllocMaxAmps = 1ull << 24;
constexpr uint64_t kNoAllocMaxShots = 1ull << 18;
void device_sample_noalloc(qslhip_state *h, const double *u, uint64_t nshots, uint64_t *out);

// resizeable backend (resize.hip): addQubit and collapseOut, in place
// where the allocation allows (the state's basis record must be written)
void resize_add_qubit(qslhip_state *h, unsigned targ);
void resize_collapse_out(qslhip_state *h, unsigned targ, unsigned outcome, double factor);

// NP compact backend (np_compact.hip)
void np_init(qslhip_state *h);
void np_free(qslhip_state *h);
void np_launch_gate(qslhip_state *h, const GateDesc &g);
// a gate list; runs of diagonal gates go through one pass each, with
// results bit-identical to np_launch_gate per gate
void np_apply_gates(qslhip_state *h, const std::vector<GateDesc> &gates);
double np_prob(qslhip_state *h, unsigned t, unsigned outcome);
void np_collapse(qslhip_state *h, unsigned t, unsigned outcome, double factor);
void np_sample(qslhip_state *h, const double *u, uint64_t nshots,
               uint64_t *out, bool strict = false);
void np_scatter_full(qslhip_state *h, void *dev_full);
void np_gather_full(qslhip_state *h, const void *dev_full);
uint64_t np_rank_host(uint64_t x);

// expectation values and adjoint gradients (adjoint.hip)
double expval(qslhip_state *h, const qslhip_pauli *terms, uint64_t nterms);
void adjoint_plain(qslhip_state *h, const std::vector<GateDesc> &gates,
                   const qslhip_pauli *terms, uint64_t nterms, double *energy,
                   double *grads);
int guarded_call(const std::function<void()> &f);

// planner (plan.cpp)
void plan_tile_passes(unsigned n, int prec, const std::vector<GateDesc> &gates,
                      int logt, int low_bits, std::vector<TilePass> &passes,
                      std::vector<TileOp> &ops);
void plan_tile_passes_reorder(unsigned n, const std::vector<GateDesc> &gates, int logt,
                              int low_bits, std::vector<TilePass> &passes,
                              std::vector<TileOp> &ops, std::vector<uint32_t> &order);
GateDesc make_desc(int prec, int kind, unsigned q0, unsigned q1, double angle);
bool is_gate1(int kind);
bool is_gate2(int kind);
bool is_np_kind(int kind);

// Tunables picked by the autotuner (autotune.cpp), with safe defaults.
struct Tuning {
    int logt_f64 = 11, low_f64 = 4;  // 4: 256-byte FP64 runs per load (12% faster
                                     // headline step than 3 at 30 qubits)
    int logt_f32 = 11, low_f32 = 3;
    int lds_max_target_f64 = 3;  // per-gate path: targets below use a tile
    int lds_max_target_f32 = 4;
    int dense_f = 4;             // qubits per fused dense unitary (4 or 5); 4: half the
                                 // matrix-core flops of 5 for 7% more unitaries (34q
                                 // FP32 Haar layers: 11.8 s against 15.4 s at 5)
    int reg_tile = 1;            // fused passes hold amplitudes in registers
    int rthreads_f64 = 256;      // workgroup size of the register-tile kernel
    int rthreads_f32 = 256;
    int rgroups = 4;             // groups per register-tile workgroup (0: persistent)
    int reg_min_qubits = 18;     // smaller states: LDS tiles (latency bound there)
    int jit = 1;                 // register-tile passes compiled at run time: 0 never,
                                 // 1 once a pass structure repeats, 2 always
    int np_lowbonds = 1;         // NP: adjacent disjoint bonds below qubit 12 fused into
                                 // one pass (0: one pass per bond)
    int np_hibonds = 2;          // NP: disjoint bonds in the high part of the strings
                                 // fused into one pass: 1 up to 4 (orbits in registers),
                                 // 2 up to 8 (groups of 5-8 staged through LDS: one
                                 // pass per 32q layer instead of two, 2518 against 2186
                                 // gates/s); 0 one pass per bond
    int group_order = 2;         // register-tile passes: 0 group g in workgroup slot g,
                                 // 1 each XCD sweeping an eighth of the groups, 2 the
                                 // latter for passes that read part of a tile
    int pass_tune = 1;           // compiled FP64 passes that read whole tiles: try the four
                                 // launch shapes (one group or a loop, group order 0 or
                                 // 1) on their first launches and keep the fastest
    int jit_prefetch = 2;        // compiled passes: a workgroup takes several groups and
                                 // loads the next while computing one: 0 never (one group
                                 // each), 1 always, 2 for FP32 passes and passes that read
                                 // part of a tile
    int dense_reorder = 0;       // dense MFMA plans: 1 commutation-aware tile passes (34q
                                 // config 5: 46 passes instead of 113 but 372 dense
                                 // unitaries instead of 264, 90 instead of 75 TFLOP/s and
                                 // 197 instead of 231 gates/s), 0 program order
    int free_layout = 1;         // register-tile passes load and store in the layouts their
                                 // first and last gates use (0: registers on the top positions)
    int np_phase_lut = 1;        // NP: runs of phase gates through per-string masks of the
                                 // gates acting on it (0: every gate tested per amplitude)
    int np_bonds = 1;            // NP: swap-like gates on disjoint pairs per pass (1: one
                                 // each; grouping measured slower for config 3)
    int np_adj_max = 4;          // NP: adjacent swap-like gates on qubits q, q+1 with
                                 // q <= this (and q+1 in the low table part) run one
                                 // element per lane with cross-lane exchange (-1: never)
    int np_adj_policy = 2;       // that kernel's memory policy: 0 pair members only,
                                 // 1 members with streaming accesses, 2 whole chunks,
                                 // 3 whole chunks streaming
    int np_stream_min_q = 64;    // NP gates whose lowest qubit is at least this use
                                 // streaming accesses (states of stream_min_qubits+);
                                 // off: on C(32,16) FP64 they run 5-10% slower
    int pass_stream = 3;         // register-tile passes on such states: streaming
                                 // loads (bit 0) and stores (bit 1)
    int lane_gates = 1;          // one-qubit gates on targets 0..5 by cross-lane
                                 // exchange (0: one-op LDS tiles / streaming pairs)
    int stream_min_qubits = 26;  // states of at least this many qubits (1 GiB FP64)
                                 // use streaming loads and stores in the gate
                                 // and reduction kernels; smaller ones stay cached
    int jit_waves = 0;           // compiled passes: register budget for this many
                                 // waves per SIMD (0: the compiler's choice)
};
Tuning &tuning();
extern int g_jit_compiled, g_jit_failed;  // passes compiled at run time / failed
void jit_wait_idle();  // blocks until the compiler thread has no queued pass
extern int g_run_sums_used;  // samplings that took a gate pass's run sums
int pass_tune_pending();       // passes still trying launch shapes (pass_tune)

inline unsigned grid_for(uint64_t items, unsigned per_block)
{
    uint64_t g = (items + per_block - 1) / per_block;
    return (unsigned)(g ? g : 1);
}

}  // namespace qslhip
