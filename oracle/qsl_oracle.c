/*
 * qsl_oracle.c -- CPU oracle for the qsl state-vector hot path.
 * TEST INFRASTRUCTURE ONLY (see qsl_oracle.h).  Built by oracle/Makefile
 * with -O2 -ffp-contract=off, matching the reference build's separate
 * multiply and add (CMakeLists.txt:23: -O2, no -march, so no FMA).
 */
#define _GNU_SOURCE
#include "qsl_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---- integer combinatorics (src/utils/misc.cpp) ------------------------- */

/* next(): Gosper's hack, misc.cpp:44-52.  The reference uses the 32-bit
 * __builtin_ctz on a size_t; ctzll gives the same answer whenever the low
 * 32 bits are nonzero (every string of n <= 32 bits with k >= 1) and stays
 * defined above that. */
void orc_next(uint64_t *x)
{
    uint64_t v = *x;
    uint64_t t = v | (v - 1);
    *x = (t + 1) | (((~t & (0 - ~t)) - 1) >> (__builtin_ctzll(v) + 1));
}

unsigned orc_hamming_weight(uint64_t x) /* misc.cpp:54-62 */
{
    unsigned c = 0;
    while (x) { c += (unsigned)(x & 1u); x >>= 1; }
    return c;
}

/* choose(): misc.cpp:65-73 with a u64 accumulator.  The running product
 * C(n-k+i, i) is exact after each division, as in the reference; the
 * reference's 32-bit version overflows at C(32,16). */
uint64_t orc_choose(unsigned n, unsigned k)
{
    if (k > n) return 0;
    uint64_t ans = 1;
    for (unsigned i = 1; i <= k; i++) {
        ans *= (uint64_t)(n - k + i);
        ans /= i;
    }
    return ans;
}

/* generateLookup(): misc.cpp:75-90, all len-bit strings of weight `ones` in
 * increasing order.  Returns the count written to out. */
size_t orc_generate_lookup(unsigned len, unsigned ones, uint64_t *out)
{
    size_t m = 0;
    if (ones == 0) { out[0] = 0; return 1; }
    uint64_t x = (1ull << ones) - 1ull;
    uint64_t end = x << (len - ones);
    while (x <= end) {
        out[m++] = x;
        orc_next(&x);
        if (x == 0) break; /* wrapped past 2^64 (len == 64 only) */
    }
    return m;
}

/* orc_choose(n, k) for n, k < 64, filled once when the library loads (the
 * same values; tests compare the table with orc_choose), so that rank and
 * unrank cost table reads rather than a loop of divisions per binomial. */
static uint64_t g_choose[64][65];

__attribute__((constructor)) static void choose_table_init(void)
{
    for (unsigned n = 0; n < 64; n++)
        for (unsigned k = 0; k <= 64; k++) g_choose[n][k] = orc_choose(n, k);
}

uint64_t orc_choose_table(unsigned n, unsigned k)
{
    return (n < 64 && k <= 64) ? g_choose[n][k] : orc_choose(n, k);
}

/* Colex rank: sum over the set bits p_1 < ... < p_k of C(p_i, i).  Because
 * generateLookup enumerates in increasing numeric order, which is colex
 * order, lookup(n,k)[rank(x)] == x. */
uint64_t orc_np_rank(uint64_t x)
{
    uint64_t r = 0;
    unsigned i = 0;
    while (x) {
        unsigned p = (unsigned)__builtin_ctzll(x);
        i++;
        r += g_choose[p][i];
        x &= x - 1;
    }
    return r;
}

uint64_t orc_np_unrank(uint64_t r, unsigned k)
{
    uint64_t x = 0;
    for (unsigned i = k; i >= 1; i--) {
        unsigned p = i - 1; /* C(i-1, i) = 0 <= r always */
        while (p + 1 < 64 && g_choose[p + 1][i] <= r) p++;
        x |= 1ull << p;
        r -= g_choose[p][i];
    }
    return x;
}

/* ---- floating-point bodies, once per precision -------------------------- */
#define R double
#define S d
#define RCOS cos
#define RSIN sin
#define RSQRT sqrt
#define RSINCOS sincos
#include "oracle_impl.inc"
#undef R
#undef S
#undef RCOS
#undef RSIN
#undef RSQRT
#undef RSINCOS

#define R float
#define S f
#define RCOS cosf
#define RSIN sinf
#define RSQRT sqrtf
#define RSINCOS sincosf
#include "oracle_impl.inc"
#undef R
#undef S
#undef RCOS
#undef RSIN
#undef RSQRT
#undef RSINCOS
