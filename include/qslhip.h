/*
 * qslhip.h -- C ABI of the MI355X (gfx950) state-vector backend for qsl.
 *
 * Every entry point returns an int status (QSLHIP_OK == 0) and never throws.
 * A handle owns one state vector in device memory and one HIP stream; calls
 * on one handle are stream-ordered, and every call that returns host data
 * synchronises that stream first.  Handles are not thread-safe: one caller
 * thread per handle, as with qsl::Qubits (the reference is not thread-safe).
 *
 * Amplitudes are interleaved {re, im} of the handle's precision (prec 8 =
 * double, 4 = float): the layout of qsl::complex<Fp>
 * (include/qsl/utils/complex.hpp:40-49), so get/set are plain copies.
 *
 * Each declaration names the reference interface it replaces.  The C++20
 * class templates qsl::Qubits<qsl::Type::Hip, Fp> and
 * qsl::Qubits<qsl::Type::HipNP, Fp> (include/qsl/qubits/hip.hpp,
 * include/qsl/qubits/hip_np.hpp) are the reference-shaped API over it.
 */
#ifndef QSLHIP_H
#define QSLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
    QSLHIP_OK = 0,
    QSLHIP_EINVAL = 1,  /* bad argument (qubit index, kind, pointer, size)   */
    QSLHIP_ESIZE = 2,   /* state size mismatch: the reference's logic_error  */
    QSLHIP_ENOTNP = 3,  /* state not number preserving / bad nones           */
    QSLHIP_ENOMEM = 4,  /* device allocation failed                          */
    QSLHIP_ENODEV = 5,  /* no usable HIP device                              */
    QSLHIP_EDEVICE = 6, /* any other HIP runtime error                       */
    QSLHIP_EKIND = 7    /* gate not available on this handle type            */
};

/* ---- gate kinds ------------------------------------------------------------
 * One-qubit: include/qsl/qubits/default.hpp:208-333.
 * Two-qubit: include/qsl/qubits/default.hpp:354-608; for controlled gates q0
 * is the control and q1 the target; for swap/fswap/np* q0, q1 are q1, q2. */
enum {
    QSLHIP_PAULI_X = 0, QSLHIP_PAULI_Y = 1, QSLHIP_PAULI_Z = 2,
    QSLHIP_HADAMARD = 3, QSLHIP_PHASE = 4, QSLHIP_ROT_X = 5,
    QSLHIP_ROT_Y = 6, QSLHIP_ROT_Z = 7,
    QSLHIP_CNOT = 16, QSLHIP_CY = 17, QSLHIP_CZ = 18, QSLHIP_CROT_X = 19,
    QSLHIP_CROT_Y = 20, QSLHIP_CROT_Z = 21, QSLHIP_CPHASE = 22,
    QSLHIP_CHADAMARD = 23, QSLHIP_SWAP = 24, QSLHIP_FSWAP = 25,
    QSLHIP_NP_ROT_X = 26, QSLHIP_NP_ROT_Y = 27, QSLHIP_NP_HADAMARD = 28
};

/* One gate of a gate list (24 bytes). */
typedef struct {
    int32_t kind;
    uint32_t q0;
    uint32_t q1;   /* unused by one-qubit gates */
    uint32_t pad;  /* must be 0 */
    double angle;  /* unused by fixed gates; rounded to the handle's Fp */
} qslhip_op;

/* Execution modes of qslhip_apply_ops. */
enum {
    QSLHIP_MODE_PER_GATE = 0, /* one kernel per gate (the reference's loop)  */
    QSLHIP_MODE_FUSED = 1,    /* gates packed into LDS-tile passes; same
                                 per-amplitude arithmetic, bit-identical     */
    QSLHIP_MODE_DENSE = 2,    /* FP32 handles: gate runs fused into dense
                                 4/5-qubit unitaries (FP64 products rounded
                                 once) applied on matrix cores (MFMA)        */
    QSLHIP_MODE_REORDER = 3   /* tile passes after commutation-aware
                                 reordering (gates on disjoint qubits, or
                                 both diagonal, may swap) with fused
                                 multiply-add arithmetic: the same product
                                 in exact arithmetic, fewer HBM passes and
                                 FP64 operations; the amplitudes differ from
                                 program order by rounding only (FP64: far
                                 below 1e-12 per amplitude)                 */
};

typedef struct qslhip_state *qslhip_t;

/* ---- library --------------------------------------------------------------- */
const char *qslhip_status_string(int status);
/* Detail of the most recent failure on the calling thread ("" if none). */
const char *qslhip_last_error(void);
/* Number of visible HIP devices; QSLHIP_ENODEV when the runtime has none. */
int qslhip_device_count(int *count);

/* ---- lifecycle ---------------------------------------------------------------
 * qslhip_create    <- Qubits<Default,Fp>(unsigned nqubits)  default.hpp:114,
 *                     default.cpp:39-46 (allocates 2^n, resets to |0..0>).
 * qslhip_create_np <- Qubits<NP,Fp>(unsigned nqubits, unsigned nones)
 *                     np.hpp:107, np.cpp:42-55 (throws unless 1<=nones<n);
 *                     the device keeps only the C(n,nones) amplitudes of
 *                     weight nones, in colex-rank order.
 * qslhip_clone     <- copy constructor default.hpp:137 (device-to-device).
 * qslhip_assign    <- operator= default.cpp:94-105 (QSLHIP_ESIZE when the
 *                     qubit counts differ).
 * prec is 8 (double) or 4 (float). */
int qslhip_create(unsigned nqubits, int prec, qslhip_t *out);
int qslhip_create_np(unsigned nqubits, unsigned nones, int prec, qslhip_t *out);
int qslhip_destroy(qslhip_t h);
int qslhip_clone(qslhip_t src, qslhip_t *out);
int qslhip_assign(qslhip_t dst, qslhip_t src);
/* qslhip_create_resize <- Qubits<Resize,Fp>(unsigned nqubits)
 *                     src/qubits/resize/resize.cpp:39-46: a full-vector
 *                     handle whose qubits can be added and removed
 *                     (qslhip_add_qubit, qslhip_measure_out, ...). */
int qslhip_create_resize(unsigned nqubits, int prec, qslhip_t *out);
/* nones is 0 for a full-vector handle; namps is the device amplitude count. */
int qslhip_info(qslhip_t h, unsigned *nqubits, unsigned *nones, int *prec,
                uint64_t *namps);
/* Use `stream` (a hipStream_t, or NULL for the handle's own) for all later
 * work of this handle.  The caller keeps the stream alive. */
int qslhip_set_stream(qslhip_t h, void *stream);
/* The hipStream_t all of this handle's work is queued on (for event timing). */
int qslhip_get_stream(qslhip_t h, void **stream);
int qslhip_sync(qslhip_t h);

/* ---- state -------------------------------------------------------------------
 * reset      <- reset()  default.cpp:57-65 / np.cpp:68-79 (index 2^nones-1)
 * set_state  <- setState() default.cpp:67-78; count must be 2^n
 *               (QSLHIP_ESIZE otherwise).  NP: the full 2^n vector; its
 *               weight is re-derived as checkStateNP does (np.cpp:81-98).
 * get_state  <- getState() default.cpp:108-112; NP returns the full 2^n
 *               vector (np.cpp:135-139).
 * set_basis  <- setBasisState() default.cpp:80-90; NP recomputes nones by
 *               popcount (np.cpp:100-116).
 * set_num_ones <- setNumOnes() np.cpp:155-165 (resets the state).
 * get_amps / set_amps: raw device-layout access ([first, first+count) of the
 *               namps amplitudes; compact colex order for NP handles). */
int qslhip_reset(qslhip_t h);
int qslhip_set_state(qslhip_t h, const void *host, uint64_t count);
int qslhip_get_state(qslhip_t h, void *host, uint64_t count);
int qslhip_set_basis(qslhip_t h, uint64_t index);
int qslhip_set_num_ones(qslhip_t h, unsigned nones);
int qslhip_get_amps(qslhip_t h, void *host, uint64_t first, uint64_t count);
int qslhip_set_amps(qslhip_t h, const void *host, uint64_t first,
                    uint64_t count);

/* ---- gates -------------------------------------------------------------------
 * qslhip_gate1 <- the one-qubit members default.hpp:208-333 (rotateX/Y/Z,
 *                 pauliX/Y/Z, hadamard, phase) and np.hpp:158-168.
 * qslhip_gate2 <- the two-qubit members default.hpp:354-608 and
 *                 np.hpp:173-208.
 * The angle's cos/sin are taken on the host in the handle's Fp exactly as
 * the reference does (gates.cpp:124-125), so both sides multiply by the
 * same coefficients. */
int qslhip_gate1(qslhip_t h, int kind, unsigned targ, double angle);
int qslhip_gate2(qslhip_t h, int kind, unsigned q0, unsigned q1, double angle);
/* A gate list applied in order (no reference counterpart: the reference
 * calls one member per gate).  mode: QSLHIP_MODE_*. */
int qslhip_apply_ops(qslhip_t h, const qslhip_op *ops, uint64_t nops,
                     int mode);

/* Planner introspection (host only, no device work): how
 * qslhip_apply_ops(QSLHIP_MODE_FUSED) would split `ops` on an n-qubit state
 * of precision prec.  Writes up to `cap` tile-qubit masks and per-pass gate
 * counts; *npasses is the full pass count. */
int qslhip_plan(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops,
                uint64_t *npasses, uint64_t *tile_masks, uint64_t *op_counts,
                uint64_t cap);
/* The same for any tile mode (QSLHIP_MODE_FUSED or QSLHIP_MODE_REORDER),
 * also writing the execution order: order[i] = index in `ops` of the i-th
 * gate applied (nops entries; may be null). */
int qslhip_plan_mode(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops,
                     int mode, uint64_t *npasses, uint64_t *tile_masks,
                     uint64_t *op_counts, uint64_t cap, uint64_t *order);

/* Passes compiled at run time (no reference counterpart; host only).
 * qslhip_jit_source writes the device source generated for pass `pass` of
 * the register-tile plan of `ops` in `mode` (QSLHIP_MODE_FUSED or _REORDER)
 * with the current tuning (*len = its length); qslhip_jit_check compiles a
 * source with hipRTC for `arch` (e.g. "gfx950") and writes the compiler
 * log.  Neither touches a device. */
int qslhip_jit_source(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops, int mode,
                      uint64_t pass, char *buf, size_t cap, uint64_t *len);
int qslhip_jit_check(const char *src, const char *arch, char *log, size_t cap);
/* Passes compiled at run time (tuning "jit" = 1) are compiled on a host
 * thread of the library once their structure repeats; until a module is
 * ready the pass runs interpreted, with the same results.  qslhip_jit_wait
 * blocks until every queued structure is compiled (no reference
 * counterpart; for callers that time steady state). */
int qslhip_jit_wait(void);

/* ---- tuning (no reference counterpart) ------------------------------------
 * Keys: "logt_f64", "low_f64", "logt_f32", "low_f32" (LDS tile width and the
 * contiguous low run of fused passes), "lds_max_target_f64/_f32" (highest
 * target the per-gate path routes through a one-op tile; -1 = none),
 * "dense_f" (qubits per fused dense unitary, 4 (default) or 5).  Further
 * keys select kernel variants (number-preserving bond passes "np_lowbonds",
 * "np_hibonds" 0-2, "np_bonds"; launch shapes; streaming); an unknown key or
 * an out-of-range value returns QSLHIP_EINVAL.  qslhip_autotune times
 * the candidates on the device (scratch state of min(nqubits, 26) qubits),
 * keeps the fastest in the process-wide table and writes a JSON report. */
int qslhip_tuning_get(const char *key, int *value);
int qslhip_tuning_set(const char *key, int value);
int qslhip_autotune(int prec, unsigned nqubits, char *report, size_t cap);

/* Dense-plan introspection (host only): how QSLHIP_MODE_DENSE with fusion
 * width F would split `ops`.  pass_info: 3 u64 per pass (tile mask, first
 * op, op count); op_info: 9 i64 per op (kind 0 = the single gate gidx as a
 * tile op, 1 = dense unitary starting at gate gidx; nq; float offset of its
 * row-major (re, im) matrix in `mats`; its nq qubits, low bit first). */
int qslhip_plan_dense(unsigned nqubits, const qslhip_op *ops, uint64_t nops, int F,
                      uint64_t *npasses, uint64_t *nops_out, uint64_t *nfloats,
                      uint64_t *pass_info, int64_t *op_info, float *mats,
                      uint64_t cap_passes, uint64_t cap_ops, uint64_t cap_floats);

/* ---- measurement ---------------------------------------------------------------
 * Uniforms are injected (the reference draws them from an unseedable
 * mt19937, random.cpp:32-45); a handle of precision 4 compares them as
 * float, as the reference's Fp uniform does.
 * prob       <- prob(targ, outcome) measure.cpp:126-147 (FP64 accumulation)
 * collapse   <- collapse() measure.cpp:31-57 (private in the reference)
 * measure    <- measure(targ) measure.cpp:81-107
 * postselect <- postselect() measure.cpp:149-161
 * measure_all<- measureAll() measure.cpp:110-123
 * sample_all <- sampleAll(N) measure.cpp:220-234: one outcome per uniform
 *               (the C++ layer folds them into the std::map)
 * sample     <- sample(targ, N) measure.cpp:191-216: counts[0], counts[1]
 * norm       <- norm(getState()) quantum.cpp:53-64 (FP64) */
int qslhip_prob(qslhip_t h, unsigned targ, unsigned outcome, double *p);
int qslhip_collapse(qslhip_t h, unsigned targ, unsigned outcome,
                    double factor);
int qslhip_measure(qslhip_t h, unsigned targ, double u, int *outcome);
int qslhip_postselect(qslhip_t h, unsigned targ, unsigned outcome, double *p);
int qslhip_measure_all(qslhip_t h, double u, uint64_t *outcome);
int qslhip_sample_all(qslhip_t h, const double *u, uint64_t nshots,
                      uint64_t *out);
/* ---- resizeable handles (src/qubits/resize/) ----------------------------------
 * Only on handles made by qslhip_create_resize (QSLHIP_EKIND otherwise).
 * resize_set_state <- Resize::setState resize.cpp:70-81: any power-of-two
 *               count (QSLHIP_ESIZE otherwise); the qubit count follows.
 * add_qubit  <- addQubit(targ) resize.cpp:107-147: a new qubit in |0> at
 *               index targ <= n (targ == n: appendQubit, :101-105); qubits
 *               k >= targ move to k + 1.  When 2^(n+1) amplitudes exceed
 *               the allocation it doubles (out of place), else the move
 *               runs in place.
 * collapse_out <- collapseOut measure.cpp:62-94 (private in the reference):
 *               keep qubit targ == outcome, times factor, and remove targ.
 * measure_out <- measureOut(targ) measure.cpp:157-190 (measure's rule with
 *               the injected uniform u, prob0 and factor in FP64).
 * postselect_out <- postselectOut(targ, outcome) measure.cpp:235-250.
 * reallocate <- reallocateState() resize.cpp:93-98: the allocation trimmed
 *               to 2^n amplitudes.
 * capacity:   amplitudes the allocation holds (>= 2^n).
 * Removing the last qubit is refused (QSLHIP_EINVAL). */
int qslhip_resize_set_state(qslhip_t h, const void *host, uint64_t count);
int qslhip_add_qubit(qslhip_t h, unsigned targ);
int qslhip_collapse_out(qslhip_t h, unsigned targ, unsigned outcome, double factor);
int qslhip_measure_out(qslhip_t h, unsigned targ, double u, int *outcome);
int qslhip_postselect_out(qslhip_t h, unsigned targ, unsigned outcome, double *p);
int qslhip_reallocate(qslhip_t h);
int qslhip_capacity(qslhip_t h, uint64_t *cap);

/* sample_all_noalloc: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the outcomes of qslhip_sample_all,
 * with every device buffer in one static array of the library (about
 * 6.3 MB per device, held until the process ends) instead of the handle's
 * lazily allocated sampling buffers.  Full-vector handles of at most 2^24
 * amplitudes (24 qubits), at most 2^18 uniforms per call (QSLHIP_EINVAL
 * beyond either; QSLHIP_EKIND for number-preserving handles).  It calls no
 * allocating routine.  All calls on one device share the array: calls that
 * overlap on one device race. */
int qslhip_sample_all_noalloc(qslhip_t h, const double *u, uint64_t nshots,
                              uint64_t *out);
/* sampleAll2 <- sampleAll2(N) measure.cpp:237-267: counts[x] = number of
 *               uniforms u with cdf(x-1) <= u < cdf(x) (strict merge rule);
 *               uniforms beyond the total mass are not counted.  ncounts
 *               must be 2^nqubits (indexed by basis state, NP included). */
int qslhip_sample_all2(qslhip_t h, const double *u, uint64_t nshots,
                       uint64_t *counts, uint64_t ncounts);
int qslhip_sample(qslhip_t h, unsigned targ, const double *u, uint64_t nshots,
                  uint64_t *counts);
int qslhip_norm(qslhip_t h, double *norm);

/* ---- expectation values and adjoint gradients (no reference counterpart;
 * the reference has no observables or gradients) ----------------------------
 * An observable is a real combination of Pauli strings.  Qubit q of a term
 * is X if only bit q of xmask is set, Z if only bit q of zmask, Y if both:
 *   P|j> = i^nY (-1)^popcount(j & zmask) |j ^ xmask>. */
typedef struct {
    double coef;
    uint64_t xmask;
    uint64_t zmask;
} qslhip_pauli;

/* <psi|H|psi> of the handle's state, accumulated in FP64 (full-vector
 * handles). */
int qslhip_expval(qslhip_t h, const qslhip_pauli *terms, uint64_t nterms,
                  double *value);

/* Applies `ops` to the handle's state psi0 (the handle keeps psi = U psi0)
 * and returns E = <psi|H|psi> and dE/d(angle) of every op (0 for gates
 * without an angle) by the adjoint reverse sweep: one kernel per gate, in
 * the handle's precision, FP64 accumulation.  Differentiable kinds: rotateX/
 * Y/Z, phase, controlRotateX/Y/Z, controlPhase, npRotateX/Y. */
int qslhip_expval_grad(qslhip_t h, const qslhip_op *ops, uint64_t nops,
                       const qslhip_pauli *terms, uint64_t nterms,
                       double *energy, double *grads);

/* Batched variational circuits: ncirc circuits of nqubits[c] qubits, each
 * started from |0...0>, circuit c owning ops[op_offsets[c] ..
 * op_offsets[c+1]) and observable terms[term_offsets[c] ..
 * term_offsets[c+1]).  FP32 states, FP64 energies and gradients, one
 * persistent kernel captured in a hipGraph at creation.
 * qslhip_batch_run replays it with new angles (NULL keeps the previous
 * ones; otherwise `angles` must hold exactly op_offsets[ncirc] doubles, one
 * per op, since the call takes no count) and writes energies[ncirc] and
 * grads[op_offsets[ncirc]], grads[i] the derivative by op i's angle.  The
 * persistent grid is capped by free device memory (each workgroup owns a
 * slot of 2 x 2^max(nqubits) FP32 amplitudes); QSLHIP_ENOMEM when not even
 * one slot fits.  Circuits of more than 12 qubits may apply
 * commuting gates in another order (the same unitary; results differ by
 * FP32 rounding only). */
typedef struct qslhip_batch *qslhip_batch_t;
int qslhip_batch_create(int prec, uint32_t ncirc, const uint32_t *nqubits,
                        const uint64_t *op_offsets, const qslhip_op *ops,
                        const uint64_t *term_offsets, const qslhip_pauli *terms,
                        qslhip_batch_t *out);
int qslhip_batch_run(qslhip_batch_t b, const double *angles, double *energies,
                     double *grads);
/* The stream the batch graph is launched on (for event timing). */
int qslhip_batch_stream(qslhip_batch_t b, void **stream);
int qslhip_batch_destroy(qslhip_batch_t b);

/* ---- synthetic workloads and seeded uniforms (no reference counterpart;
 * the reference's generators are unseedable, src/utils/random.cpp:32-45) --
 * qslhip_uniforms: mt19937_64 seeded with splitmix64(seed), u =
 * (x >> 11) * 2^-53 in [0, 1). */
int qslhip_uniforms(uint64_t seed, uint64_t count, double *out);

#ifdef __cplusplus
}
#endif
#endif
