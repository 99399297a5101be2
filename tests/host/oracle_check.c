/* Host-only run of the CPU oracle (oracle/primus_oracle.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer, built by
 * tests/test_sanitizers.py.  It drives every oracle entry point over edge
 * sizes and moduli and checks identities that hold whatever the code does
 * inside: transform/inverse round trips, the transform product against the
 * schoolbook product, lazy ranges, HEXL and Uint agreement on canonical
 * outputs, Barrett and MultiplyFactor results against exact 128-bit
 * remainders, decomposition ranges, the external product's two output
 * forms, automorphism composition, base conversion against CRT.  Prints "ok <checks>" or names
 * the first failure and exits 1. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/primus_oracle.h"

typedef unsigned __int128 u128;
static unsigned long long checks = 0;

#define CHECK(cond, ...)                                      \
    do {                                                      \
        ++checks;                                             \
        if (!(cond)) {                                        \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                     \
            fprintf(stderr, "\n");                            \
            exit(1);                                          \
        }                                                     \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void) {  /* splitmix64 */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t below(uint64_t q) { return (uint64_t)(((u128)rnd() * q) >> 64); }
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

static int is_prime(uint64_t q) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (q < 2) return 0;
    for (int i = 0; i < 12; ++i)
        if (q % bases[i] == 0) return q == bases[i];
    uint64_t d = q - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (int i = 0; i < 12; ++i) {
        uint64_t x = po_pow_mod(bases[i], d, q);
        if (x == 1 || x == q - 1) continue;
        int comp = 1;
        for (int r = 1; r < s && comp; ++r) {
            x = mulmod(x, x, q);
            if (x == q - 1) comp = 0;
        }
        if (comp) return 0;
    }
    return 1;
}

static uint64_t ntt_prime(int bits, uint32_t log_n) {
    const uint64_t m = 2ull << log_n;
    uint64_t q = ((1ull << bits) - 1) / m * m + 1;
    if (q >= (1ull << bits)) q -= m;
    while (q > m && !is_prime(q)) q -= m;
    CHECK(q > m, "no prime below 2^%d for N = 2^%u", bits, log_n);
    return q;
}

static uint64_t *alloc(size_t n) {
    uint64_t *p = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    CHECK(p != NULL, "allocation");
    return p;
}

static void check_transforms(uint32_t log_n, uint64_t q) {
    po_table *t = NULL;
    CHECK(po_table_new(log_n, q, &t) == 0, "po_table_new(%u, %llu)", log_n, (unsigned long long)q);
    const size_t n = po_table_n(t);
    uint64_t *a = alloc(n), *b = alloc(n), *x = alloc(n), *y = alloc(n), *c = alloc(n), *w = alloc(n);
    for (size_t i = 0; i < n; ++i) a[i] = below(q), b[i] = below(q);
    /* lazy forward: [0,4q) in, [0,4q) out, congruent to the canonical forward */
    for (size_t i = 0; i < n; ++i) x[i] = a[i] + q * (rnd() & 3);
    memcpy(y, a, n * 8);
    po_lazy_transform(t, x);
    po_transform(t, y);
    for (size_t i = 0; i < n; ++i) {
        CHECK(x[i] < 4 * q && y[i] < q && x[i] % q == y[i], "lazy forward slot %zu", i);
    }
    /* inverse of the forward is the identity (inputs in [0,2q)) */
    for (size_t i = 0; i < n; ++i) x[i] = y[i] + q * (rnd() & 1);
    po_inverse_transform(t, x);
    for (size_t i = 0; i < n; ++i) CHECK(x[i] == a[i], "round trip slot %zu", i);
    memcpy(x, y, n * 8);
    po_lazy_inverse_transform(t, x);
    for (size_t i = 0; i < n; ++i) CHECK(x[i] % q == a[i] && x[i] < 2 * q, "lazy inverse slot %zu", i);
    /* the transform product is the negacyclic product (schoolbook for small N) */
    if (log_n <= 9) {
        memcpy(x, a, n * 8);
        memcpy(w, b, n * 8);
        po_transform(t, x);
        po_transform(t, w);
        po_pointwise_mul(x, w, c, n, q);
        po_inverse_transform(t, c);
        po_schoolbook_negacyclic(a, b, y, n, q);
        for (size_t i = 0; i < n; ++i) CHECK(c[i] == y[i], "negacyclic product slot %zu", i);
        /* multiply-accumulate */
        memcpy(c, x, n * 8);
        po_pointwise_mul_add(c, x, w, n, q);
        for (size_t i = 0; i < n; ++i) CHECK(c[i] == (x[i] + mulmod(x[i], w[i], q)) % q, "mul_add slot %zu", i);
    }
    /* HEXL radix-2: canonical outputs equal the Uint ones, lazy outputs in range and congruent */
    memcpy(x, a, n * 8);
    memcpy(y, a, n * 8);
    po_transform(t, x);
    po_hexl_forward(t, y, 1);
    for (size_t i = 0; i < n; ++i) CHECK(x[i] == y[i], "HEXL forward slot %zu", i);
    memcpy(y, a, n * 8);
    po_hexl_forward(t, y, 4);
    for (size_t i = 0; i < n; ++i) CHECK(y[i] < 4 * q && y[i] % q == x[i], "HEXL lazy forward slot %zu", i);
    memcpy(y, x, n * 8);
    po_hexl_inverse(t, y, 1);
    for (size_t i = 0; i < n; ++i) CHECK(y[i] == a[i], "HEXL inverse slot %zu", i);
    memcpy(y, x, n * 8);
    po_hexl_inverse(t, y, 2);
    for (size_t i = 0; i < n; ++i) CHECK(y[i] < 2 * q && y[i] % q == a[i], "HEXL lazy inverse slot %zu", i);
    /* monomials: c X^d transformed equals the transform of the coefficient vector */
    for (size_t d = 0; d < n; d += (n > 8 ? n / 8 : 1)) {
        const uint64_t coeff = below(q);
        memset(x, 0, n * 8);
        x[d] = coeff;
        po_transform(t, x);
        po_transform_monomial(t, coeff, d, y);
        for (size_t i = 0; i < n; ++i) CHECK(x[i] == y[i], "monomial %zu slot %zu", d, i);
        memset(x, 0, n * 8);
        x[d] = 1;
        po_transform(t, x);
        po_transform_coeff_one_monomial(t, d, y);
        for (size_t i = 0; i < n; ++i) CHECK(x[i] == y[i], "X^d monomial slot %zu", i);
        memset(x, 0, n * 8);
        x[d] = q - 1;
        po_transform(t, x);
        po_transform_coeff_minus_one_monomial(t, d, y);
        for (size_t i = 0; i < n; ++i) CHECK(x[i] == y[i], "-X^d monomial slot %zu", i);
    }
    /* NTT-domain automorphism x -> x^g then x^(g^-1 mod 2N) is the identity,
       and it agrees with the coefficient-domain one */
    if (n >= 2) {
        const uint64_t two_n = 2 * (uint64_t)n;
        const uint64_t g = 5 % two_n;
        uint64_t ginv = 1;
        for (uint64_t k = 1; k < two_n; k += 2)
            if ((k * g) % two_n == 1) ginv = k;
        memcpy(x, a, n * 8);
        po_transform(t, x);
        po_ntt_automorphism(x, y, n, g);
        po_ntt_automorphism(y, w, n, ginv);
        for (size_t i = 0; i < n; ++i) CHECK(w[i] == x[i], "automorphism inverse slot %zu", i);
        po_coeff_automorphism(a, c, n, g, q);
        po_transform(t, c);
        for (size_t i = 0; i < n; ++i) CHECK(c[i] == y[i], "automorphism domains slot %zu", i);
    }
    free(a), free(b), free(x), free(y), free(c), free(w);
    po_table_free(t);
}

static void check_scalar(uint64_t q) {
    uint64_t ratio[2];
    CHECK(po_barrett_ratio(q, ratio) == 0, "Barrett ratio of %llu", (unsigned long long)q);
    for (int i = 0; i < 2000; ++i) {
        const uint64_t a = below(q), b = below(q), c = below(q), v = rnd();
        CHECK(po_barrett_reduce_mul(a, b, q, ratio) == mulmod(a, b, q), "reduce_mul");
        CHECK(po_barrett_reduce_mul_add(a, b, c, q, ratio) == (uint64_t)(((u128)a * b + c) % q), "reduce_mul_add");
        const uint64_t lz = po_barrett_lazy_reduce(v, q, ratio);
        CHECK(lz < 2 * q && lz % q == v % q, "lazy_reduce");
        const uint64_t hi = below(q);
        const uint64_t lz2 = po_barrett_lazy_reduce_2limb(v, hi, q, ratio);
        CHECK(lz2 < 2 * q && lz2 % q == (uint64_t)((((u128)hi << 64) | v) % q), "lazy_reduce_2limb");
        const uint64_t wq = po_shoup_quotient(a, q);
        const uint64_t sl = po_shoup_lazy_mul(a, wq, v, q);
        CHECK(sl < 2 * q && sl % q == mulmod(a, v % q, q), "Shoup lazy");
        CHECK(po_shoup_mul(a, wq, v, q) == mulmod(a, v % q, q), "Shoup");
        if (a) CHECK(mulmod(po_inv_mod(a, q), a, q) == 1, "inv_mod");
        const uint32_t shifts[3] = {32, 52, 64};
        for (int k = 0; k < 3; ++k) {
            if (shifts[k] < 64 && (q >> (shifts[k] - 1)) != 0) continue;  /* shift must exceed the modulus bits */
            const uint64_t f = po_multiply_factor(b, shifts[k], q);
            const uint64_t r = po_multiply_mod_lazy(a, b, f, shifts[k], q);
            CHECK(r < 2 * q && r % q == mulmod(a, b, q), "MultiplyFactor shift %u", shifts[k]);
        }
    }
}

static void check_decompose(uint64_t q, uint32_t log_basis, uint32_t l) {
    po_basis b;
    if (po_basis_new(q, log_basis, l, &b) != 0) return;
    const size_t n = 64;
    uint64_t *v = alloc(n), *res = alloc((size_t)l * n);
    for (size_t i = 0; i < n; ++i) v[i] = below(q);
    v[0] = 0, v[1] = q - 1, v[2] = q / 2;
    po_decompose(&b, v, n, res);
    for (size_t k = 0; k < (size_t)l * n; ++k) CHECK(res[k] < q, "residue out of range");
    free(v), free(res);
}

static void check_ext_product(uint32_t log_n, uint64_t q, uint32_t log_basis, uint32_t l) {
    po_table *t = NULL;
    CHECK(po_table_new(log_n, q, &t) == 0, "table");
    po_basis b;
    CHECK(po_basis_new(q, log_basis, l, &b) == 0, "basis");
    const size_t n = po_table_n(t);
    uint64_t *key = alloc(4 * (size_t)l * n), *in = alloc(2 * n), *out = alloc(2 * n), *out2 = alloc(2 * n);
    for (size_t i = 0; i < 4 * (size_t)l * n; ++i) key[i] = below(q);
    for (size_t i = 0; i < 2 * n; ++i) in[i] = below(q);
    po_ext_product(t, &b, key, in, out, 1);
    po_ext_product(t, &b, key, in, out2, 0);
    for (size_t i = 0; i < 2 * n; ++i) CHECK(out[i] < q && out2[i] < q, "ext product range");
    /* the coefficient output is the inverse transform of the NTT output */
    po_inverse_transform(t, out);
    po_inverse_transform(t, out + n);
    for (size_t i = 0; i < 2 * n; ++i) CHECK(out[i] == out2[i], "ext product outputs slot %zu", i);
    free(key), free(in), free(out), free(out2);
    po_table_free(t);
}

static void check_base_conversion(void) {
    /* in base {p0, p1}, out base {r}: exact conversion of x < p0 p1 is x mod r */
    const uint64_t ib[2] = {ntt_prime(40, 4), ntt_prime(41, 4)};
    const uint64_t ob[2] = {ntt_prime(45, 4), ntt_prime(44, 4)};
    po_base_converter *c = NULL;
    CHECK(po_base_converter_new(ib, 2, ob, 2, &c) == 0, "converter");
    const size_t n = 257;
    uint64_t *in = alloc(2 * n), *out = alloc(2 * n), *fast = alloc(2 * n);
    u128 xs[257];
    const u128 P = (u128)ib[0] * ib[1];
    for (size_t i = 0; i < n; ++i) {
        u128 x = (((u128)rnd() << 64) | rnd()) % P;
        if (i == 0) x = 0;
        if (i == 1) x = P - 1;
        xs[i] = x;
        in[i] = (uint64_t)(x % ib[0]);
        in[n + i] = (uint64_t)(x % ib[1]);
    }
    po_exact_convert_array(c, in, out, n);
    po_fast_convert_array(c, in, fast, n);
    for (size_t i = 0; i < n; ++i)
        for (int j = 0; j < 2; ++j) {
            const uint64_t want = (uint64_t)(xs[i] % ob[j]);
            /* the exact conversion has one output modulus, the first, and
               gives the centred value: x - P when x > P/2 */
            const uint64_t centred = xs[i] > P / 2 ? (want + ob[j] - (uint64_t)(P % ob[j])) % ob[j] : want;
            if (j == 0) CHECK(out[i] == centred, "exact conversion element %zu", i);
            /* fast conversion: x + k P for some 0 <= k < 2 (in-base count) */
            const uint64_t pm = (uint64_t)(P % ob[j]);
            const uint64_t f = fast[j * n + i];
            CHECK(f == want || f == (want + pm) % ob[j], "fast conversion element %zu limb %d", i, j);
        }
    free(in), free(out), free(fast);
    po_base_converter_free(c);
}

static void check_rlwe(void) {
    uint64_t th[38];
    po_gauss_thresholds(th);
    for (int i = 1; i < 38; ++i) CHECK(th[i] >= th[i - 1], "Gaussian thresholds increase");
    const uint32_t log_n = 10, limbs = 2;
    const size_t n = 1u << log_n;
    const uint64_t moduli[2] = {ntt_prime(59, log_n), ntt_prime(58, log_n)};
    int8_t *s = (int8_t *)malloc(n), *e = (int8_t *)malloc(n);
    uint8_t *m = (uint8_t *)malloc(n);
    uint64_t *a = alloc(limbs * n), *out = alloc(n);
    CHECK(s && e && m, "allocation");
    po_rlwe_sample(7, 3, log_n, limbs, moduli, s, e, m, a);
    for (size_t i = 0; i < n; ++i) CHECK(s[i] >= -1 && s[i] <= 1 && e[i] >= -19 && e[i] <= 19, "sample ranges");
    for (uint32_t j = 0; j < limbs; ++j) {
        po_table *t = NULL;
        CHECK(po_table_new(log_n, moduli[j], &t) == 0, "table");
        for (size_t i = 0; i < n; ++i) CHECK(a[j * n + i] < moduli[j], "a_hat range");
        po_rlwe_encrypt(t, s, e, m, 12345 % moduli[j], a + j * n, out);
        for (size_t i = 0; i < n; ++i) CHECK(out[i] < moduli[j], "b_hat range");
        po_table_free(t);
    }
    free(s), free(e), free(m), free(a), free(out);
}

int main(void) {
    /* degrees 2^1 .. 2^13 over modulus widths up to 62 bits */
    const int widths[] = {17, 31, 32, 50, 59, 61, 62};
    for (uint32_t log_n = 1; log_n <= 13; ++log_n)
        for (size_t k = 0; k < sizeof(widths) / sizeof(widths[0]); ++k)
            if ((int)log_n + 3 < widths[k]) check_transforms(log_n, ntt_prime(widths[k], log_n));
    check_transforms(3, 17);
    const uint64_t scal[] = {3, 17, 65537, ntt_prime(31, 4), ntt_prime(50, 4), ntt_prime(59, 4), ntt_prime(62, 4)};
    for (size_t k = 0; k < sizeof(scal) / sizeof(scal[0]); ++k) check_scalar(scal[k]);
    const uint64_t q59 = ntt_prime(59, 10), q50 = ntt_prime(50, 10);
    check_decompose(q59, 7, 8);
    check_decompose(q59, 1, 59);
    check_decompose(q50, 10, 5);
    check_decompose(q50, 25, 2);
    check_ext_product(10, q59, 7, 8);
    check_ext_product(10, q50, 10, 5);
    check_ext_product(6, ntt_prime(31, 6), 4, 7);
    check_base_conversion();
    check_rlwe();
    {
        const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
        uint32_t out[4];
        po_philox4x32_10(ctr, key, out);
        /* Random123's known answer for the zero counter and key */
        CHECK(out[0] == 0x6627e8d5u && out[1] == 0xe169c58du && out[2] == 0xbc57ac4cu && out[3] == 0x9b00dbd8u,
              "Philox known answer");
    }
    printf("ok %llu\n", checks);
    return 0;
}
