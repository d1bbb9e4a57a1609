// Host check of div_sqrt2 (qsl_amd/csrc/gate_math.h): the library's
// three-operation division by the Fp-rounded sqrt(2) must equal the
// reference's plain division x / sqrt2 (src/qubits/default/gates.cpp:94-100)
// bit for bit.  FP32: every finite x.  FP64: a seeded sample over every
// exponent, denser where amplitudes live, plus the guard boundaries.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "gate_math.h"

using qslhip::div_sqrt2;
using qslhip::div_sqrt2_x4;

int main(int argc, char **argv)
{
    const long f64_samples = argc > 1 ? std::atol(argv[1]) : (1L << 26);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const float df = (float)std::sqrt(2.0);
    const double dd = std::sqrt(2.0);
    std::vector<long> bad32(nt), bad64(nt), n64(nt);
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; w++)
        th.emplace_back([&, w] {
            for (uint64_t u = w; u < (1ull << 32); u += nt) {
                uint32_t b = (uint32_t)u;
                float x;
                std::memcpy(&x, &b, 4);
                if (std::isnan(x)) continue;
                float want = x / df, got = div_sqrt2(x);
                if (std::memcmp(&want, &got, 4)) bad32[w]++;
                // the four-at-once form, mixed with values on both sides of the
                // guard, on every 16th x
                if (u & 15) continue;
                float a = x, bb = -x * 0x1p-110f, c = 0.0f, e = x * 0.75f;
                float wa = a / df, wb = bb / df, wc = c / df, we = e / df;
                div_sqrt2_x4(a, bb, c, e);
                if (std::memcmp(&wa, &a, 4) || std::memcmp(&wb, &bb, 4) ||
                    std::memcmp(&wc, &c, 4) || std::memcmp(&we, &e, 4))
                    bad32[w]++;
            }
            uint64_t s = 0x9E3779B97F4A7C15ull * (w + 1);
            for (long i = w; i < f64_samples; i += nt) {
                s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                uint64_t bits = s;
                if (i % 4) {  // exponents of amplitudes, 2^-64 .. 2^3
                    uint64_t e = 1023 - 64 + (s >> 56) % 68;
                    bits = (bits & 0x800FFFFFFFFFFFFFull) | (e << 52);
                } else if (i % 8 == 0) {  // around the guard thresholds
                    uint64_t e = (i % 16 == 0) ? 1023 - 900 - 2 + (s >> 60) % 4
                                               : 1023 + 900 - 2 + (s >> 60) % 4;
                    bits = (bits & 0x800FFFFFFFFFFFFFull) | (e << 52);
                }
                double x;
                std::memcpy(&x, &bits, 8);
                if (std::isnan(x)) continue;
                double want = x / dd, got = div_sqrt2(x);
                n64[w]++;
                if (std::memcmp(&want, &got, 8)) bad64[w]++;
                double a = x, b = -0.0, c = x * 0x1p-1000, e = -x * 3.0;
                double wa = a / dd, wb = b / dd, wc = c / dd, we = e / dd;
                div_sqrt2_x4(a, b, c, e);
                if (std::memcmp(&wa, &a, 8) || std::memcmp(&wb, &b, 8) ||
                    std::memcmp(&wc, &c, 8) || std::memcmp(&we, &e, 8))
                    bad64[w]++;
            }
        });
    for (auto &t : th) t.join();
    long b32 = 0, b64 = 0, c64 = 0;
    for (unsigned w = 0; w < nt; w++) {
        b32 += bad32[w];
        b64 += bad64[w];
        c64 += n64[w];
    }
    std::printf("fp32 all non-NaN x: %ld mismatches\nfp64 %ld samples: %ld mismatches\n", b32, c64,
                b64);
    return (b32 || b64) ? 1 : 0;
}
