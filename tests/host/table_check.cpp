// Host-only checks of the table construction (primus_amd/csrc/host_tables.hpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py.  Every property is checked against its
// definition, not against the code that computes it:
//   * the root is a primitive 2N-th root and the smallest one (every odd
//     power of it, the set of all primitive 2N-th roots, is examined);
//   * root_powers[rev(k)] = root^k, inv_root_powers[rev(k)+1] = root^-(k+1)
//     (their product with the matching power is 1), ordinal powers by repeated
//     multiplication;
//   * Shoup quotients: floor(w 2^64 / q) by the defining inequalities;
//   * N^-1 N = 1, the scaled last root, floor(2^128 / q) and the one-word
//     Barrett constant by their inequalities;
//   * binary64 twiddles centred (|w'| <= q/2, w' = w mod q) with
//     |wp - w'/q| <= 2^-54 + 2^-64;
//   * the register orders are bijections that move values, never change them.
// Prints "ok <checks>" and exits 0, or names the first failure and exits 1.
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <vector>

#include "../../primus_amd/csrc/host_tables.hpp"

using namespace ph;

static unsigned long long g_checks = 0;

#define CHECK(cond, ...)                                 \
    do {                                                 \
        ++g_checks;                                      \
        if (!(cond)) {                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                \
            fprintf(stderr, "\n");                       \
            exit(1);                                     \
        }                                                \
    } while (0)

static bool is_prime(uint64_t q) {
    if (q < 2) return false;
    for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull})
        if (q % p == 0) return q == p;
    // deterministic Miller-Rabin for 64-bit q with the twelve prime bases
    uint64_t d = q - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (uint64_t a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
        uint64_t x = h_powmod(a, d, q);
        if (x == 1 || x == q - 1) continue;
        bool comp = true;
        for (int r = 1; r < s && comp; ++r) {
            x = h_mulmod(x, x, q);
            if (x == q - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

// largest prime q < 2^bits with q = 1 mod 2^(log_n+1)
static uint64_t ntt_prime(int bits, uint32_t log_n) {
    const uint64_t m = 2ull << log_n;
    uint64_t q = ((1ull << bits) - 1) / m * m + 1;
    if (q >= (1ull << bits)) q -= m;
    while (q > m && !is_prime(q)) q -= m;
    CHECK(q > m, "no prime below 2^%d for N = 2^%u", bits, log_n);
    return q;
}

static void check_shoup(const Shoup &s, uint64_t q) {
    CHECK(s.w < q, "value %llu not below q %llu", (unsigned long long)s.w, (unsigned long long)q);
    // wq q <= w 2^64 < (wq + 1) q
    const u128 lhs = (u128)s.wq * q, w64 = (u128)s.w << 64;
    CHECK(lhs <= w64 && w64 - lhs < q, "Shoup quotient of %llu mod %llu", (unsigned long long)s.w,
          (unsigned long long)q);
}

static void check_f64(const Shoup &s, const ShoupF &f, uint64_t q) {
    const long double wc = f.w;
    CHECK(std::fabs((double)wc) <= (double)(q / 2) + 0.5, "centred value out of range");
    const long long wi = (long long)f.w;
    CHECK((long double)wi == wc, "binary64 value not an integer");
    const uint64_t back = wi < 0 ? (uint64_t)(wi + (long long)q) : (uint64_t)wi;
    CHECK(back == s.w, "binary64 value not congruent to the twiddle");
    const long double exact = wc / (long double)q;
    // the long double quotient itself is within 2^-64 of w'/q
    CHECK(std::fabs((double)(exact - (long double)f.wp)) <= 0x1p-54 + 0x1p-64 + 0x1p-64,
          "binary64 quotient off by more than 2^-54 + 2^-64");
}

static void check_table(uint32_t log_n, uint64_t q) {
    HostTable h;
    const int st = host_table(log_n, q, h);
    CHECK(st == PH_OK, "host_table(%u, %llu) = %d", log_n, (unsigned long long)q, st);
    const uint64_t n = 1ull << log_n;
    // primitive 2N-th root: root^N = -1
    CHECK(h_powmod(h.root, n, q) == q - 1, "root^N != -1");
    // minimal: every primitive 2N-th root is an odd power of root
    uint64_t cur = h.root;
    const uint64_t r2 = h_mulmod(h.root, h.root, q);
    for (uint64_t k = 0; k < n; ++k) {
        CHECK(cur >= h.root, "odd power root^%llu = %llu is smaller than root %llu", (unsigned long long)(2 * k + 1),
              (unsigned long long)cur, (unsigned long long)h.root);
        cur = h_mulmod(cur, r2, q);
    }
    CHECK(h.ord.size() == 2 * n && h.fwd.size() == n && h.inv.size() == n, "table sizes");
    uint64_t p = 1;
    for (uint64_t k = 0; k < 2 * n; ++k) {
        CHECK(h.ord[k].w == p, "ordinal power %llu", (unsigned long long)k);
        check_shoup(h.ord[k], q);
        p = h_mulmod(p, h.root, q);
    }
    CHECK(h.inv_root == h.ord[2 * n - 1].w && h_mulmod(h.inv_root, h.root, q) == 1 % q, "inverse root");
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t rk = 0;
        for (uint32_t i = 0; i < log_n; ++i) rk |= ((k >> i) & 1ull) << (log_n - 1 - i);
        // ord[k] = root^k was checked above by repeated multiplication
        CHECK(h.fwd[rk].w == h.ord[k].w, "root_powers[rev(%llu)]", (unsigned long long)k);
        check_shoup(h.fwd[k], q);
        check_shoup(h.inv[k], q);
        if (k + 1 < n)
            CHECK(h_mulmod(h.inv[rk + 1].w, h.ord[k + 1].w, q) == 1 % q, "inv_root_powers[rev(%llu)+1]",
                  (unsigned long long)k);
    }
    CHECK(h.inv[0].w == 1 % q, "inv_root_powers[0]");
    CHECK(h_mulmod(h.inv_n, n % q, q) == 1 % q, "N^-1");
    check_shoup(Shoup{h.inv_n, h.inv_n_q}, q);
    CHECK(h.sr == h_mulmod(h.inv[n - 1].w, h.inv_n, q), "scaled last root");
    check_shoup(Shoup{h.sr, h.sr_q}, q);
    // floor(2^128 / q): 0 < 2^128 - r q < q (q odd, so never 0)
    const u128 r = ((u128)h.br1 << 64) | h.br0;
    const u128 rem = (u128)0 - r * (u128)q;
    CHECK(rem > 0 && rem < q && (u128)(r * (u128)q) / q == r, "Barrett ratio");
    if (q < (1ull << 59)) {
        const uint32_t b = 64 - __builtin_clzll(q);
        const u128 top = (u128)1 << (b - 1 + 64);
        CHECK(h.bar_sh == b - 1 && (u128)h.bar_mu * q <= top && top - (u128)h.bar_mu * q < q, "one-word Barrett");
    } else {
        CHECK(h.bar_mu == 0, "one-word Barrett constant set for q >= 2^59");
    }
    // binary64 forms and register orders (the single-pass launch shapes)
    if (q < (1ull << 50)) {
        std::vector<ShoupF> ff(n), fi(n);
        for (uint64_t k = 0; k < n; ++k) {
            ff[k] = mkf(h.fwd[k], q);
            fi[k] = mkf(h.inv[k], q);
            check_f64(h.fwd[k], ff[k], q);
            check_f64(h.inv[k], fi[k], q);
        }
        // tag entries with their index to see where register_order moves them
        std::vector<ShoupF> tf(n), ti(n);
        for (uint64_t k = 0; k < n; ++k) tf[k] = ShoupF{(double)k, 0.0}, ti[k] = ShoupF{(double)k, 0.0};
        for (int loge = 3; loge <= 4 && log_n <= 13; ++loge) {
            std::vector<ShoupF> pf, pi;
            register_order(tf, ti, log_n, loge, pf, pi);
            CHECK(pf.size() == n && pi.size() == n, "register order sizes");
            std::vector<int> seen(n, 0);
            for (uint64_t k = 1; k < n; ++k) seen[(uint64_t)pf[k].w]++;
            for (uint64_t k = 1; k < n; ++k) CHECK(seen[k] == 1, "forward register order not a bijection");
            // inverse: every stage but the last keeps its segment [1 + n - (n >> s), 1 + n - (n >> (s+1)))
            for (uint64_t k = 1; k + 1 < n; ++k) {
                const uint64_t src = (uint64_t)pi[k].w;
                uint32_t s = 0;
                while (1 + n - (n >> (s + 1)) <= k) ++s;
                CHECK(src >= 1 + n - (n >> s) && src < 1 + n - (n >> (s + 1)), "inverse register order leaves its stage");
            }
            std::fill(seen.begin(), seen.end(), 0);
            for (uint64_t k = 1; k + 1 < n; ++k) seen[(uint64_t)pi[k].w]++;
            for (uint64_t k = 1; k + 1 < n; ++k) CHECK(seen[k] == 1, "inverse register order not a bijection");
            // the operand order: a bijection per limb
            const uint32_t L = 2;
            std::vector<ShoupF> op(L * n);
            for (uint64_t k = 0; k < L * n; ++k) op[k] = ShoupF{(double)k, 0.0};
            const std::vector<ShoupF> ro = operand_register_order(op, L, log_n, loge);
            std::vector<int> cnt(L * n, 0);
            for (uint64_t k = 0; k < L * n; ++k) {
                const uint64_t src = (uint64_t)ro[k].w;
                CHECK(src / n == k / n, "operand order crosses limbs");
                cnt[src]++;
            }
            for (uint64_t k = 0; k < L * n; ++k) CHECK(cnt[k] == 1, "operand register order not a bijection");
        }
    }
}

int main() {
    // brute-force minimality on a small prime: every candidate below the root fails
    {
        const uint64_t q = 257;
        uint64_t root = 0;
        CHECK(minimal_primitive_root(8, q, &root) == PH_OK, "root mod 257");
        for (uint64_t w = 1; w < root; ++w) CHECK(h_powmod(w, 128, q) != q - 1, "smaller primitive root %llu", (unsigned long long)w);
        CHECK(h_powmod(root, 128, q) == q - 1, "root mod 257 not primitive");
    }
    // no root: 2N does not divide q - 1 (97 - 1 = 3 * 2^5)
    {
        uint64_t root = 0;
        CHECK(minimal_primitive_root(5, 97, &root) == PH_OK, "97 has a primitive 32nd root");
        CHECK(minimal_primitive_root(6, 97, &root) == PH_ERR_NO_PRIMITIVE_ROOT, "97 has no primitive 64th root");
        HostTable h;
        CHECK(host_table(4, 17, h) == PH_ERR_NO_PRIMITIVE_ROOT, "a table of N = 16 mod 17");
    }
    // inverses
    for (uint64_t q : {17ull, 257ull, 65537ull, 1152921504606584833ull}) {
        for (uint64_t a = 1; a < 200; ++a)
            if (a % q) CHECK(h_mulmod(inv_mod(a, q), a % q, q) == 1, "inv_mod(%llu, %llu)", (unsigned long long)a,
                             (unsigned long long)q);
    }
    CHECK(inv_mod(6, 9) == 0, "inv_mod of a non-unit");
    // tables over the supported range of degrees and modulus widths
    const int widths[] = {20, 31, 32, 45, 50, 52, 59, 60, 61, 62};
    for (uint32_t log_n = 1; log_n <= 17; ++log_n) {
        for (int bits : widths) {
            if ((int)log_n + 2 >= bits) continue;
            if (log_n > 13 && bits != 50 && bits != 59 && bits != 62) continue;  // keep the run short
            check_table(log_n, ntt_prime(bits, log_n));
        }
    }
    check_table(8, 7681);    // small classic NTT prime
    check_table(10, 12289);
    printf("ok %llu\n", g_checks);
    return 0;
}
