/*
 * qsl_oracle.h -- CPU oracle for the qsl state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * arithmetic of qsl's Default, Omp and NP backends (lanamineh/qsl), used as
 * the checker for the HIP backend and as the CPU baseline ("port") timed
 * beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Conventions (all pinned by tests/test_oracle_*.py):
 *   - state vectors are interleaved {re, im} arrays of 2^n complex numbers,
 *     qubit k == bit k of the basis index (src/qubits/default/gates.cpp:35-41);
 *   - every function exists in a double (_d) and a float (_f) flavour; the
 *     float flavour keeps the reference's float arithmetic (Fp accumulators,
 *     Fp uniforms) exactly as the src/qubits/default sources do;
 *   - the reference draws its uniforms from a random_device-seeded mt19937
 *     (src/utils/random.cpp:32-45); here every uniform is injected by the
 *     caller so that the GPU and the oracle see identical draws;
 *   - shifts are 64-bit (the reference's `1 << targ` is int, UB for targ>=31).
 *
 * Pinning status per function is listed in the header comment of each
 * implementation file; nothing here is "parity unpinned".
 */
#ifndef QSL_ORACLE_H
#define QSL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate kinds.  Numbering is shared with include/qslhip.h (checked by a test). */
enum {
    ORC_PAULI_X = 0, ORC_PAULI_Y = 1, ORC_PAULI_Z = 2, ORC_HADAMARD = 3,
    ORC_PHASE = 4, ORC_ROT_X = 5, ORC_ROT_Y = 6, ORC_ROT_Z = 7,
    ORC_CNOT = 16, ORC_CY = 17, ORC_CZ = 18, ORC_CROT_X = 19, ORC_CROT_Y = 20,
    ORC_CROT_Z = 21, ORC_CPHASE = 22, ORC_CHADAMARD = 23, ORC_SWAP = 24,
    ORC_FSWAP = 25, ORC_NP_ROT_X = 26, ORC_NP_ROT_Y = 27, ORC_NP_HADAMARD = 28
};

/* One entry of a gate list: kind, qubits (q1 unused for 1-qubit gates) and
 * angle (unused for fixed gates).  Layout identical to qslhip_op. */
typedef struct {
    int32_t kind;
    uint32_t q0;
    uint32_t q1;
    uint32_t pad;
    double angle;
} orc_op;

/* ---- Default semantics (src/qubits/default/{default,gates,measure}.cpp) -- */
#define ORC_DECL(R, S)                                                          \
    void orc_reset_##S(R *st, unsigned n);                                      \
    void orc_set_basis_##S(R *st, unsigned n, uint64_t index);                  \
    void orc_gate1_##S(R *st, unsigned n, int kind, unsigned t, R angle);       \
    void orc_gate2_##S(R *st, unsigned n, int kind, unsigned a, unsigned b,     \
                       R angle);                                                \
    void orc_apply_ops_##S(R *st, unsigned n, const orc_op *ops, size_t nops);  \
    R orc_prob_##S(const R *st, unsigned n, unsigned t, unsigned outcome);      \
    void orc_collapse_##S(R *st, unsigned n, unsigned t, unsigned outcome,      \
                          R factor);                                            \
    int orc_measure_##S(R *st, unsigned n, unsigned t, R u);                    \
    R orc_postselect_##S(R *st, unsigned n, unsigned t, unsigned outcome);      \
    size_t orc_gen_dist_##S(const R *st, unsigned n, uint64_t *idx,             \
                            double *cdf);                                       \
    uint64_t orc_draw_sample_##S(const uint64_t *idx, const double *cdf,        \
                                 size_t size, R u);                             \
    uint64_t orc_measure_all_##S(R *st, unsigned n, R u);                       \
    void orc_sample_all_##S(const R *st, unsigned n, const R *u, size_t nshots, \
                            uint64_t *out);                                     \
    void orc_sample_all2_##S(const R *st, unsigned n, const R *u,               \
                             size_t nshots, uint64_t *counts);                  \
    void orc_sample_##S(const R *st, unsigned n, unsigned t, const R *u,        \
                        size_t nshots, uint64_t *counts2);                      \
    R orc_norm_##S(const R *st, uint64_t dim);                                  \
    void orc_add_qubit_##S(R *st, unsigned n, unsigned targ);                   \
    void orc_collapse_out_##S(R *st, unsigned n, unsigned targ,                 \
                              unsigned outcome, R factor);                      \
    int orc_measure_out_##S(R *st, unsigned n, unsigned targ, R u);             \
    R orc_postselect_out_##S(R *st, unsigned n, unsigned targ, unsigned outcome); \
    R orc_normalise_##S(R *st, uint64_t dim);                                   \
    R orc_fubini_study_##S(const R *v, const R *w, uint64_t dim);               \
    int orc_check_state_np_##S(const R *st, uint64_t dim, unsigned *nones);     \
    /* Omp-backend loop structure (src/qubits/omp/{gates,measure}.cpp) */       \
    void orc_omp_gate1_##S(R *st, unsigned n, int kind, unsigned t, R angle,    \
                           int nthreads);                                       \
    void orc_omp_gate2_##S(R *st, unsigned n, int kind, unsigned a, unsigned b, \
                           R angle, int nthreads);                              \
    void orc_omp_apply_ops_##S(R *st, unsigned n, const orc_op *ops,            \
                               size_t nops, int nthreads);                      \
    R orc_omp_prob_##S(const R *st, unsigned n, unsigned t, unsigned outcome,   \
                       int nthreads);                                           \
    /* NP semantics on the full 2^n vector (src/qubits/np/{np,gates,measure}) */ \
    void orc_np_reset_##S(R *st, unsigned n, unsigned nones);                   \
    void orc_np_gate1_##S(R *st, unsigned n, unsigned nones, int kind,          \
                          unsigned t, R angle);                                 \
    void orc_np_gate2_##S(R *st, unsigned n, unsigned nones, int kind,          \
                          unsigned a, unsigned b, R angle);                     \
    R orc_np_prob_##S(const R *st, unsigned n, unsigned nones, unsigned t,      \
                      unsigned outcome);                                        \
    size_t orc_np_gen_dist_##S(const R *st, unsigned n, unsigned nones,         \
                               uint64_t *idx, R *cdf);                          \
    uint64_t orc_np_draw_sample_##S(const uint64_t *idx, const R *cdf,          \
                                    size_t size, R u);                          \
    void orc_np_sample_all_##S(const R *st, unsigned n, unsigned nones,         \
                               const R *u, size_t nshots, uint64_t *out);       \
    /* NP gates on the compact colex-ranked vector (length C(n,k)) */           \
    void orc_npc_gate1_##S(R *amp, unsigned n, unsigned k, int kind,            \
                           unsigned t, R angle, int nthreads);                  \
    void orc_npc_gate2_##S(R *amp, unsigned n, unsigned k, int kind,            \
                           unsigned a, unsigned b, R angle, int nthreads);      \
    void orc_npc_apply_ops_##S(R *amp, unsigned n, unsigned k,                  \
                               const orc_op *ops, size_t nops, int nthreads);   \
    void orc_npc_to_full_##S(const R *amp, unsigned n, unsigned k, R *full);    \
    void orc_npc_from_full_##S(const R *full, unsigned n, unsigned k, R *amp);

ORC_DECL(double, d)
ORC_DECL(float, f)
#undef ORC_DECL

/* ---- integer combinatorics (src/utils/misc.cpp:44-90) ------------------- */
void orc_next(uint64_t *x);                      /* Gosper: next same popcount */
unsigned orc_hamming_weight(uint64_t x);
uint64_t orc_choose(unsigned n, unsigned k);     /* u64 (reference: 32-bit)   */
uint64_t orc_choose_table(unsigned n, unsigned k); /* the same, from a table  */
size_t orc_generate_lookup(unsigned len, unsigned ones, uint64_t *out);
uint64_t orc_np_rank(uint64_t x);                /* colex rank of x           */
uint64_t orc_np_unrank(uint64_t r, unsigned k);  /* inverse of orc_np_rank    */

#ifdef __cplusplus
}
#endif
#endif
