// Host-only check of the 32-bit-chain integer helpers of
// primus_amd/csrc/arith.hpp against plain 64-bit and 128-bit arithmetic:
//   mul_sub_lo(w, b, k, 2^64 - q) == w*b - k*q (mod 2^64)
//   sub_mul_lo(x, k, 2^64 - q)    == x - k*q   (mod 2^64)
//   sub_if_ge(x, m)               == (x >= m ? x - m : x)
//   mulhi64_approx(a, b)          in [floor(a*b/2^64) - 2, floor(a*b/2^64)]
//   shoup_approx(w, wq, b, q)     == w*b (mod q) and < 4q, for q < 2^62, w < q,
//                                    wq = floor(w*2^64/q) and any 64-bit b
//   grid_fits(groups, threads)    == (groups * threads < 2^32), the launchers'
//                                    bound on a grid's work-items
// on random operands and on the edges (0, 1, all-ones, x == m, q near 2^62).
// Prints "ok <cases>" and exits 0, or the first mismatch and exits 1.
#include <cstdio>
#include <random>

#include "../../primus_amd/csrc/arith.hpp"
#include "../../primus_amd/csrc/ntt_device.hpp"

using namespace ph;
typedef unsigned __int128 u128x;

static int fail(const char *what, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
    std::printf("mismatch %s %llu %llu %llu %llu\n", what, (unsigned long long)a, (unsigned long long)b,
                (unsigned long long)c, (unsigned long long)d);
    return 1;
}

int main() {
    std::mt19937_64 g(0x5052494D5553ull);
    const uint64_t edges[] = {0, 1, 2, 0xFFFFFFFFull, 0x100000000ull, ~0ull, ~0ull - 1, 1ull << 63, (1ull << 62) - 57};
    long cases = 0;
    for (long i = 0; i < 4000000; ++i) {
        uint64_t w = g(), b = g(), k = g(), q = g() >> (g() % 64);
        if (i < 81) w = edges[i % 9], b = edges[i / 9];
        if (i % 13 == 7) k = edges[i % 9];
        if (mul_sub_lo(w, b, k, 0 - q) != w * b - k * q) return fail("mul_sub_lo", w, b, k, q);
        if (sub_mul_lo(w, k, 0 - q) != w - k * q) return fail("sub_mul_lo", w, k, q, 0);
        uint64_t x = g(), m = g() >> (g() % 64);
        if (i % 3 == 0) m = x;
        if (i % 5 == 0) m = x + 1;
        if (i % 11 == 0) x = m - 1;
        if (i < 81) x = edges[i % 9], m = edges[i / 9];
        if (sub_if_ge(x, m) != (x >= m ? x - m : x)) return fail("sub_if_ge", x, m, 0, 0);
        const uint64_t exact = (uint64_t)(((u128x)w * b) >> 64), est = mulhi64_approx(w, b);
        if (est > exact || exact - est > 2) return fail("mulhi64_approx", w, b, est, exact);
        cases += 4;
    }
    for (long i = 0; i < 2000000; ++i) {
        uint64_t q = (g() >> (2 + g() % 60)) | 3;  // odd, below 2^62
        if (i % 97 == 0) q = (1ull << 62) - 57;
        const uint64_t w = i % 31 == 0 ? q - 1 : g() % q;
        const uint64_t wq = (uint64_t)(((u128x)w << 64) / q);
        uint64_t b = g();
        if (i % 4 == 1) b %= 16 * (u128x)q > ~0ull ? ~0ull : 16 * q;
        if (i % 17 == 0) b = ~0ull;
        const uint64_t r = shoup_approx(w, wq, b, q);
        if (r >= 4 * (u128x)q || r % q != (uint64_t)(((u128x)w * b) % q)) return fail("shoup_approx", w, b, q, r);
        ++cases;
    }
    for (uint64_t threads : {1ull, 64ull, 256ull, 512ull, 1024ull}) {
        const uint64_t edge[] = {0, 1, (1ull << 32) / threads - 1, (1ull << 32) / threads,
                                 (1ull << 32) / threads + 1, 1ull << 40, ~0ull / threads};
        for (uint64_t groups : edge) {
            const bool want = (u128x)groups * threads < ((u128x)1 << 32);
            if (grid_fits(groups, threads) != want) return fail("grid_fits", groups, threads, want, 0);
            ++cases;
        }
    }
    std::printf("ok %ld\n", cases);
    return 0;
}
