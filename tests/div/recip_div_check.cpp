// Checks armon::div_by_recip (armon-kokkos_amd/csrc/device_math.hpp, the function the batched
// sweeps divide by a non-power-of-two cell size with) against IEEE division, bit for bit, on the
// host. Test infrastructure: built by tests/div/Makefile, run by tests/test_recip_div.py.
//
// Cases, each with x of both signs and of both orders |x| >= |d| and |x| < |d|:
//  1. random significands and exponents over the range recip_div_ok accepts, d in [2^-100, 2^100];
//  2. quotients nearest a rounding midpoint: for 53-bit odd B and small t, the A with
//     A * 2^54 - N * B = t for an odd N, so that A / B lies |t| / (B 2^54) from the midpoint
//     N / 2^54 (and the same for the other binade, 2^53);
//  3. every pair of significands within 8 ulps below 2 (the corner of the proof);
//  4. zeros of both signs, and the ends of the accepted exponent range.
// Usage: recip_div_check <random cases> <midpoint cases> [seed]. Exit status 0 when every case
// matches (cases recip_div_ok refuses are counted, not compared).
#include "device_math.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using armon::div_by_recip;
using armon::recip_div_ok;

static uint64_t rng_state = 0x243F6A8885A308D3ull;
static uint64_t next()
{
    // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double make(uint64_t mant52, int exp, int sign)
{
    const uint64_t bits = ((uint64_t) sign << 63) | ((uint64_t) (exp + 1023) << 52) | (mant52 & 0xFFFFFFFFFFFFFull);
    double d;
    std::memcpy(&d, &bits, sizeof d);
    return d;
}

static long compared = 0, refused = 0, mismatches = 0;

static void check(double x, double d)
{
    if (!recip_div_ok(x)) {
        refused++;
        return;
    }
    const double inv = 1.0 / d;
    const double got = div_by_recip(x, d, inv), want = x / d;
    compared++;
    if (std::memcmp(&got, &want, sizeof got) != 0) {
        if (mismatches < 10) std::printf("MISMATCH x=%a d=%a got %a want %a\n", x, d, got, want);
        mismatches++;
    }
}

// A * 2^k - N * B = t with A < 2^53, N odd: A = t * (2^k)^-1 mod B.
static unsigned __int128 inverse_pow2(int k, uint64_t B)
{
    // 2^k mod B, then its inverse by extended Euclid
    unsigned __int128 p = 1;
    for (int i = 0; i < k; i++) p = (p * 2) % B;
    __int128 r0 = (__int128) B, r1 = (__int128) p, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const __int128 q = r0 / r1;
        __int128 t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    if (s0 < 0) s0 += B;
    return (unsigned __int128) s0;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <random cases> <midpoint cases> [seed]\n", argv[0]);
        return 2;
    }
    const long n_random = std::atol(argv[1]), n_mid = std::atol(argv[2]);
    if (argc > 3) rng_state = std::strtoull(argv[3], nullptr, 10);

    // 1. random operands
    for (long i = 0; i < n_random; i++) {
        const int ed = (int) (next() % 201) - 100;
        const int ex = (int) (next() % 1801) - 900;
        const double d = make(next(), ed, 0);
        const double x = make(next(), ex, (int) (next() & 1));
        check(x, d);
        check(make(next(), ed, (int) (next() & 1)), d);   // x in d's binade: |x / d| in (1/2, 2)
    }
    // 2. near-midpoint quotients (both binades of the quotient)
    long made = 0;
    while (made < n_mid) {
        const uint64_t B = (next() & ((1ull << 52) - 1)) | (1ull << 52) | 1ull;   // odd, 53 bits
        const int k = (next() & 1) ? 54 : 53;                                    // binade [1/2, 1) or [1, 2)
        const int64_t t = (int64_t) (next() % 129) - 64;
        if (t == 0) continue;
        const unsigned __int128 inv = inverse_pow2(k, B);
        const __int128 tm = ((t % (int64_t) B) + (int64_t) B) % (int64_t) B;
        const uint64_t A = (uint64_t) ((tm * inv) % B);
        if (A < (1ull << 52) || A >= (1ull << 53)) continue;
        if ((k == 54) != (A < B)) continue;                                       // the quotient's binade
        const __int128 num = (__int128) A * ((__int128) 1 << k) - t;
        if (num % B != 0) continue;
        if (((num / B) & 1) == 0) continue;                                       // N odd: a midpoint
        const double x = std::ldexp((double) A, -52), d = std::ldexp((double) B, -52);
        const int ex = (int) (next() % 401) - 200, ed = (int) (next() % 101) - 50;
        check(std::ldexp(x, ex), std::ldexp(d, ed));
        check(-std::ldexp(x, ex), std::ldexp(d, ed));
        made++;
    }
    // 3. significands within 8 ulps below 2, every pair
    for (int i = 1; i <= 8; i++)
        for (int j = 1; j <= 8; j++) {
            const double x = 2.0 - i * 0x1p-52, d = 2.0 - j * 0x1p-52;
            for (int e = -3; e <= 3; e++) {
                check(std::ldexp(x, e), d);
                check(-std::ldexp(x, e), d);
            }
        }
    // 4. zeros and the range ends
    const double ds[3] = {0x1.5555555555555p-7, 0x1.fffffffffffffp+99, 0x1p-100};
    for (double d : ds) {
        check(0.0, d);
        check(-0.0, d);
        check(std::ldexp(1.0, -900), d);
        check(std::ldexp(1.9999999999999998, 900), d);
        check(-std::ldexp(1.2345, -900), d);
    }
    // signed zeros keep their sign
    const double nz = div_by_recip(-0.0, 3.0, 1.0 / 3.0), pz = div_by_recip(0.0, 3.0, 1.0 / 3.0);
    const bool zeros = std::signbit(nz) && !std::signbit(pz) && nz == 0 && pz == 0;
    // the guard refuses what the theorem does not cover
    const bool guard = !recip_div_ok(std::ldexp(1.0, -901)) && !recip_div_ok(std::ldexp(1.0, 901)) &&
                       !recip_div_ok(INFINITY) && !recip_div_ok(NAN) && !recip_div_ok(0x1p-1074) &&
                       recip_div_ok(-0.0) && recip_div_ok(std::ldexp(1.0, -900)) && recip_div_ok(std::ldexp(1.5, 900));
    std::printf("compared %ld, refused %ld, mismatches %ld, midpoint cases %ld, signed zeros %s, guard %s\n", compared,
                refused, mismatches, made, zeros ? "ok" : "WRONG", guard ? "ok" : "WRONG");
    return (mismatches == 0 && zeros && guard) ? 0 : 1;
}
