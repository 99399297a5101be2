/*
 * primus_oracle.c -- CPU restatement of Primus's UintNttTable<u64>, ShoupFactor,
 * MultiplyFactor, BarrettModulus<u64> and the schoolbook negacyclic product.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libprimus_hip.so and the timed
 * CPU baseline.  Nothing in the product links or calls it.
 *
 * The restatement keeps the reference's arithmetic step for step (the same
 * Harvey lazy butterflies, the same Shoup quotients, the same twiddle order),
 * so lazy outputs as well as canonical outputs are reproduced.  Citations are
 * relative to the reference workspace root.
 */
#include "primus_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Shoup factor                                                              */
/* ------------------------------------------------------------------------- */

/* ShoupFactor::new: quotient = floor(w * 2^64 / q), requires w < q.
 * primus_factor/src/shoup_factor/mod.rs:44-55 */
uint64_t po_shoup_quotient(uint64_t w, uint64_t q) {
    return (uint64_t)(((u128)w << 64) / q);
}

/* LazyFactorMul::lazy_factor_mul_modulo: w*b - floor(w'*b / 2^64)*q (wrapping),
 * result in [0, 2q) for any b < 2^64.  primus_factor/src/shoup_factor/mod.rs:108-113 */
uint64_t po_shoup_lazy_mul(uint64_t w, uint64_t wq, uint64_t b, uint64_t q) {
    uint64_t hw = (uint64_t)(((u128)wq * b) >> 64);
    return w * b - q * hw;
}

/* FactorMul::factor_mul_modulo: lazy then one conditional subtraction.
 * primus_factor/src/shoup_factor/mod.rs:121-124 */
uint64_t po_shoup_mul(uint64_t w, uint64_t wq, uint64_t b, uint64_t q) {
    uint64_t t = po_shoup_lazy_mul(w, wq, b, q);
    return t >= q ? t - q : t;
}

/* ------------------------------------------------------------------------- */
/* MultiplyFactor (HEXL Barrett factor)                                       */
/* ------------------------------------------------------------------------- */

/* MultiplyFactor::new: floor((op << bit_shift) / q), bit_shift in {32,52,64}.
 * primus_factor/src/multiply.rs:18-41, 71-75 */
uint64_t po_multiply_factor(uint64_t op, uint32_t bit_shift, uint64_t q) {
    u128 n = (u128)op << bit_shift;
    return (uint64_t)(n / q);
}

/* hexl::number_theory::multiply_mod_lazy: op*x - ((x*factor) >> shift)*q.
 * primus_ntt/src/ntt/hexl/number_theory.rs:66-94 */
uint64_t po_multiply_mod_lazy(uint64_t x, uint64_t y, uint64_t y_barrett, uint32_t bit_shift,
                              uint64_t q) {
    uint64_t hw = (uint64_t)(((u128)x * y_barrett) >> bit_shift);
    return y * x - hw * q;
}

/* ------------------------------------------------------------------------- */
/* BarrettModulus<u64>                                                       */
/* ------------------------------------------------------------------------- */

/* BarrettModulus::new / new_unchecked: ratio = floor(2^128 / q) as [lo, hi];
 * panics (here: PO_ERR_INVALID) for q <= 1 or bit-length(q) >= 63.
 * primus_modulus/src/barrett/mod.rs:55-75 */
int po_barrett_ratio(uint64_t q, uint64_t ratio[2]) {
    if (q <= 1) return PO_ERR_INVALID;
    if ((64 - __builtin_clzll(q)) >= 63) return PO_ERR_INVALID;
    /* floor(2^128/q) = floor((2^128-1)/q) unless q | 2^128, impossible for q>1 odd
     * and handled exactly below for even q. */
    u128 all = ~(u128)0;
    u128 r = all / q;
    if (all % q == q - 1) r += 1;
    ratio[0] = (uint64_t)r;
    ratio[1] = (uint64_t)(r >> 64);
    return PO_OK;
}

/* LazyReduce<T>::lazy_reduce: value mod 2q via the 1-limb Barrett quotient.
 * primus_modulus/src/barrett/ops.rs:8-31 */
uint64_t po_barrett_lazy_reduce(uint64_t value, uint64_t q, const uint64_t ratio[2]) {
    uint64_t tmp = (uint64_t)(((u128)value * ratio[0]) >> 64);
    uint64_t qq = (uint64_t)(((u128)value * ratio[1] + tmp) >> 64);
    return value - qq * q;
}

/* BarrettModulus::lazy_reduce_2limb: Barrett of the 128-bit (hi, lo), [0, 2q).
 * primus_modulus/src/barrett/ops.rs:42-74 */
uint64_t po_barrett_lazy_reduce_2limb(uint64_t lo, uint64_t hi, uint64_t q,
                                      const uint64_t ratio[2]) {
    uint64_t ah = (uint64_t)(((u128)lo * ratio[0]) >> 64);
    u128 b = (u128)lo * ratio[1] + ah; /* carrying_mul */
    u128 c = (u128)hi * ratio[0];
    uint64_t d = hi * ratio[1];
    uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    uint64_t c0 = (uint64_t)c, c1 = (uint64_t)(c >> 64);
    uint64_t carry = (uint64_t)(b0 + c0 < b0);
    uint64_t bch = b1 + c1 + carry;
    uint64_t qq = d + bch;
    return lo - qq * q;
}

static inline uint64_t reduce_once(uint64_t x, uint64_t m) { return x >= m ? x - m : x; }

/* ReduceMul::reduce_mul = reduce(widening_mul) = reduce_once(lazy_reduce_2limb).
 * primus_modulus/src/barrett/ops.rs:158-175, 310-316 */
uint64_t po_barrett_reduce_mul(uint64_t a, uint64_t b, uint64_t q, const uint64_t ratio[2]) {
    u128 p = (u128)a * b;
    return reduce_once(po_barrett_lazy_reduce_2limb((uint64_t)p, (uint64_t)(p >> 64), q, ratio), q);
}

/* ReduceMulAdd::reduce_mul_add = reduce(a.carrying_mul(b, c)).
 * primus_modulus/src/barrett/ops.rs:331-337 */
uint64_t po_barrett_reduce_mul_add(uint64_t a, uint64_t b, uint64_t c, uint64_t q,
                                   const uint64_t ratio[2]) {
    u128 p = (u128)a * b + c;
    return reduce_once(po_barrett_lazy_reduce_2limb((uint64_t)p, (uint64_t)(p >> 64), q, ratio), q);
}

/* ReduceExp by square-and-multiply; any exact method yields the same residue.
 * primus_modulus/src/barrett/ops.rs:399-456 */
uint64_t po_pow_mod(uint64_t base, uint64_t exp, uint64_t q) {
    u128 r = 1 % q, b = base % q;
    while (exp) {
        if (exp & 1) r = (r * b) % q;
        b = (b * b) % q;
        exp >>= 1;
    }
    return (uint64_t)r;
}

/* UintModulus::reduce_inv via the extended Euclidean algorithm (the reference
 * calls primus_gcd's gcdinv); returns 0 when a is not invertible.
 * primus_modulus/src/unsigned_integer/ops.rs, primus_gcd/src/lib.rs */
uint64_t po_inv_mod(uint64_t a, uint64_t q) {
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        __int128 qt = r / nr, tmp;
        tmp = t - qt * nt; t = nt; nt = tmp;
        tmp = r - qt * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

/* ------------------------------------------------------------------------- */
/* Primitive root, bit reversal                                              */
/* ------------------------------------------------------------------------- */

/* ReverseLsbs::reverse_lsbs: primus_ntt/src/reverse.rs:128-141 */
uint64_t po_reverse_lsbs(uint64_t x, uint32_t bits) {
    if (x == 0 || bits == 0) return 0;
    uint64_t r = 0;
    for (uint32_t i = 0; i < 64; ++i) r |= ((x >> i) & 1ull) << (63 - i);
    return r >> (64 - bits);
}

/* PrimitiveRoot::is_primitive_root: w != 0 and w^(2^(log_degree-1)) == q-1.
 * primus_ntt/src/root.rs:27-44 */
int po_is_primitive_root(uint64_t w, uint32_t log_degree, uint64_t q) {
    if (w == 0 || log_degree == 0) return 0;
    u128 x = w;
    for (uint32_t i = 0; i + 1 < log_degree; ++i) x = (x * x) % q;
    return (uint64_t)x == q - 1;
}

/* PrimitiveRoot::try_minimal_primitive_root: the smallest primitive 2^log_degree-th
 * root of unity mod q.  primus_ntt/src/root.rs:46-111.
 *
 * The reference draws up to 200 random candidates r in [2, q-1] and tests
 * w = r^((q-1)/2^log_degree); then it minimises over the odd powers of w.  The
 * odd powers of any primitive root are exactly the set of primitive roots, so
 * the minimum does not depend on which candidate succeeded.  Here candidates
 * are r = 2, 3, ... (at most 1 << 20 of them), which makes the search
 * deterministic; for a prime q the first quadratic non-residue succeeds. */
int po_minimal_primitive_root(uint32_t log_degree, uint64_t q, uint64_t *root) {
    if (log_degree >= 64 || q < 2) return PO_ERR_INVALID;
    uint64_t qm1 = q - 1;
    uint64_t quotient = qm1 >> log_degree;
    if ((quotient << log_degree) != qm1) return PO_ERR_NO_PRIMITIVE_ROOT;
    uint64_t w = 0;
    int found = 0;
    for (uint64_t r = 2; r <= qm1 && r < (2ull + (1ull << 20)); ++r) {
        w = po_pow_mod(r, quotient, q);
        if (po_is_primitive_root(w, log_degree, q)) { found = 1; break; }
    }
    if (!found) return PO_ERR_NO_PRIMITIVE_ROOT;
    /* root.rs:97-110: walk w, w^3, w^5, ... (2^log_degree steps), keep the minimum */
    uint64_t best = w;
    uint64_t g2 = (uint64_t)(((u128)w * w) % q);
    uint64_t cur = w;
    uint64_t degree = 1ull << log_degree;
    for (uint64_t i = 0; i < degree; ++i) {
        if (cur < best) best = cur;
        cur = (uint64_t)(((u128)g2 * cur) % q);
    }
    *root = best;
    return PO_OK;
}

/* ------------------------------------------------------------------------- */
/* UintNttTable<u64>                                                         */
/* ------------------------------------------------------------------------- */

struct po_table {
    uint64_t q, root, inv_root;
    uint32_t log_n;
    size_t n;
    uint64_t inv_n, inv_n_q;
    uint64_t *rp, *rp_q;   /* root_powers (bit-reversed), N */
    uint64_t *irp, *irp_q; /* inv_root_powers (scrambled), N */
    uint64_t *op, *op_q;   /* ordinal_root_powers, 2N */
    uint64_t *rev;         /* reverse_lsbs, N */
};

void po_table_free(po_table *t) {
    if (!t) return;
    free(t->rp); free(t->rp_q); free(t->irp); free(t->irp_q);
    free(t->op); free(t->op_q); free(t->rev);
    free(t);
}

/* UintNttTable::new.  primus_ntt/src/ntt/primitive.rs:114-185 */
int po_table_new(uint32_t log_n, uint64_t q, po_table **out) {
    *out = NULL;
    uint64_t ratio[2];
    if (po_barrett_ratio(q, ratio) != PO_OK) return PO_ERR_INVALID; /* BarrettModulus::new */
    /* log_n = 0 makes the reference's inverse loop bound `log_n - 1` wrap; it is
     * rejected here, as are tables too large to hold in memory. */
    if (log_n == 0 || log_n > 30) return PO_ERR_INVALID;
    uint64_t root;
    int st = po_minimal_primitive_root(log_n + 1, q, &root);
    if (st != PO_OK) return st;
    size_t n = (size_t)1 << log_n;
    if ((uint64_t)n >= q) return PO_ERR_DEGREE_TOO_LARGE; /* primitive.rs:164-166 */
    po_table *t = (po_table *)calloc(1, sizeof(po_table));
    if (!t) return PO_ERR_NTT_TABLE;
    t->q = q; t->root = root; t->log_n = log_n; t->n = n;
    t->rp = malloc(n * 8); t->rp_q = malloc(n * 8);
    t->irp = malloc(n * 8); t->irp_q = malloc(n * 8);
    t->op = malloc(2 * n * 8); t->op_q = malloc(2 * n * 8);
    t->rev = malloc(n * 8);
    if (!t->rp || !t->rp_q || !t->irp || !t->irp_q || !t->op || !t->op_q || !t->rev) {
        po_table_free(t);
        return PO_ERR_NTT_TABLE;
    }
    /* ordinal_root_powers[k] = root^k, k < 2N (primitive.rs:128-137) */
    uint64_t root_q = po_shoup_quotient(root, q);
    t->op[0] = 1 % q; t->op_q[0] = po_shoup_quotient(t->op[0], q);
    t->op[1] = root; t->op_q[1] = root_q;
    uint64_t power = root;
    for (size_t k = 2; k < 2 * n; ++k) {
        power = po_shoup_mul(root, root_q, power, q);
        t->op[k] = power; t->op_q[k] = po_shoup_quotient(power, q);
    }
    t->inv_root = t->op[2 * n - 1]; /* primitive.rs:139 */
    for (size_t i = 0; i < n; ++i) t->rev[i] = po_reverse_lsbs(i, log_n); /* :143 */
    /* root_powers[rev(k)] = root^k (primitive.rs:145-149) */
    for (size_t k = 0; k < n; ++k) { t->rp[t->rev[k]] = t->op[k]; t->rp_q[t->rev[k]] = t->op_q[k]; }
    t->rp[0] = t->op[0]; t->rp_q[0] = t->op_q[0];
    /* inv_root_powers[rev(k)+1] = root^{-(k+1)} = ordinal[2N-1-k], k < N-1 (:151-159) */
    t->irp[0] = t->op[0]; t->irp_q[0] = t->op_q[0];
    for (size_t k = 0; k + 1 < n; ++k) {
        size_t src = 2 * n - 1 - k;
        t->irp[t->rev[k] + 1] = t->op[src];
        t->irp_q[t->rev[k] + 1] = t->op_q[src];
    }
    /* inv_n = N^{-1} mod q (:161-168) */
    t->inv_n = po_inv_mod((uint64_t)n % q, q);
    t->inv_n_q = po_shoup_quotient(t->inv_n, q);
    *out = t;
    return PO_OK;
}

size_t po_table_n(const po_table *t) { return t->n; }
uint64_t po_table_modulus(const po_table *t) { return t->q; }
uint64_t po_table_root(const po_table *t) { return t->root; }
uint64_t po_table_inv_root(const po_table *t) { return t->inv_root; }
void po_table_inv_n(const po_table *t, uint64_t *value, uint64_t *quotient) {
    *value = t->inv_n; *quotient = t->inv_n_q;
}
void po_table_powers(const po_table *t, int which, uint64_t *values, uint64_t *quotients) {
    const uint64_t *v = which == 0 ? t->rp : which == 1 ? t->irp : t->op;
    const uint64_t *w = which == 0 ? t->rp_q : which == 1 ? t->irp_q : t->op_q;
    size_t cnt = which == 2 ? 2 * t->n : t->n;
    if (values) memcpy(values, v, cnt * 8);
    if (quotients) memcpy(quotients, w, cnt * 8);
}

/* lazy_transform_slice: Cooley-Tukey, normal in, bit-reversed out, [0,4q) -> [0,4q).
 * primus_ntt/src/ntt/primitive.rs:211-232 */
void po_lazy_transform(const po_table *t, uint64_t *poly) {
    const uint64_t q = t->q, q2 = q << 1;
    size_t ri = 1;
    for (int s = (int)t->log_n - 1; s >= 0; --s) {
        size_t gap = (size_t)1 << s;
        for (size_t base = 0; base < t->n; base += gap << 1) {
            uint64_t w = t->rp[ri], wq = t->rp_q[ri];
            ++ri;
            for (size_t i = base; i < base + gap; ++i) {
                uint64_t u = reduce_once(poly[i], q2);
                uint64_t v = po_shoup_lazy_mul(w, wq, poly[i + gap], q);
                poly[i] = u + v;
                poly[i + gap] = u + q2 - v;
            }
        }
    }
}

/* transform_slice: lazy transform then reduce [0,4q) -> [0,q).
 * primus_ntt/src/ntt/primitive.rs:234-244 */
void po_transform(const po_table *t, uint64_t *poly) {
    po_lazy_transform(t, poly);
    const uint64_t q = t->q, q2 = q << 1;
    for (size_t i = 0; i < t->n; ++i) poly[i] = reduce_once(reduce_once(poly[i], q2), q);
}

/* lazy_inverse_transform_slice: Gentleman-Sande, bit-reversed in, normal out,
 * [0,2q) -> [0,q); the last stage folds N^{-1}.  primus_ntt/src/ntt/primitive.rs:246-286 */
void po_lazy_inverse_transform(const po_table *t, uint64_t *values) {
    const uint64_t q = t->q, q2 = q << 1;
    const uint32_t log_n = t->log_n;
    size_t ri = 1;
    for (uint32_t s = 0; s + 1 < log_n; ++s) {
        size_t gap = (size_t)1 << s;
        for (size_t base = 0; base < t->n; base += gap << 1) {
            uint64_t w = t->irp[ri], wq = t->irp_q[ri];
            ++ri;
            for (size_t i = base; i < base + gap; ++i) {
                uint64_t u = values[i], v = values[i + gap];
                values[i] = reduce_once(u + v, q2);
                values[i + gap] = po_shoup_lazy_mul(w, wq, u + q2 - v, q);
            }
        }
    }
    size_t gap = (size_t)1 << (log_n - 1);
    uint64_t s = t->inv_n, sq = t->inv_n_q;
    uint64_t sr = po_shoup_mul(t->irp[ri], t->irp_q[ri], s, q);
    uint64_t srq = po_shoup_quotient(sr, q);
    for (size_t i = 0; i < gap; ++i) {
        uint64_t u = values[i], v = values[i + gap];
        values[i] = po_shoup_mul(s, sq, u + v, q);
        values[i + gap] = po_shoup_mul(sr, srq, u + q2 - v, q);
    }
}

/* inverse_transform_slice: lazy inverse then reduce_once.  primitive.rs:288-296 */
void po_inverse_transform(const po_table *t, uint64_t *values) {
    po_lazy_inverse_transform(t, values);
    for (size_t i = 0; i < t->n; ++i) values[i] = reduce_once(values[i], t->q);
}

/* HexlNttTable's portable radix-2 path (the one compute_forward and
 * compute_inverse take without AVX-512, primus_ntt/src/ntt/hexl/mod.rs:530-538,
 * 623-632) over the same tables (hexl/mod.rs:188-204 build them as UintNttTable
 * does), with its own primitives: MultiplyFactor precons at shift 64
 * (primus_factor/src/multiply.rs) and hexl::number_theory::multiply_mod_lazy.
 *
 * forward_transform_to_bit_reverse_radix2_inplace, hexl/radix2.rs:52-194:
 * Harvey butterflies, one twiddle per group taken in table order from index 1;
 * output_mod_factor 1 ends with reduce_twice. */
void po_hexl_forward(const po_table *t, uint64_t *x, uint32_t output_mod_factor) {
    const uint64_t q = t->q, q2 = q << 1;
    const size_t n = t->n;
    size_t wi = 1; /* w_iter.next() skips w[0] */
    for (size_t gap = n >> 1, m = 1; m < n; gap >>= 1, m <<= 1) {
        for (size_t base = 0; base < n; base += gap << 1) {
            const uint64_t w = t->rp[wi];
            const uint64_t wp = po_multiply_factor(w, 64, q);
            ++wi;
            for (size_t i = base; i < base + gap; ++i) {
                /* fwd_butterfly_radix2, radix2.rs:40-50 */
                const uint64_t tx = reduce_once(x[i], q2);
                const uint64_t tt = po_multiply_mod_lazy(x[i + gap], w, wp, 64, q);
                x[i] = tx + tt;
                x[i + gap] = tx + q2 - tt;
            }
        }
    }
    if (output_mod_factor == 1)
        for (size_t i = 0; i < n; ++i) {
            uint64_t v = x[i]; /* reduce_twice, hexl/number_theory.rs:27-38 */
            if (v >= q2) v -= q2;
            if (v >= q) v -= q;
            x[i] = v;
        }
}

/* inverse_transform_from_bit_reverse_radix2_inplace, hexl/radix2.rs:220-382:
 * Harvey inverse butterflies with inverse twiddles in table order from index
 * 1, then the last stage with N^-1 and N^-1 * w folded in, both products lazy
 * ([0, 2q)) and the sum reduced once first; output_mod_factor 1 ends with
 * reduce_once. */
void po_hexl_inverse(const po_table *t, uint64_t *x, uint32_t output_mod_factor) {
    const uint64_t q = t->q, q2 = q << 1;
    const size_t n = t->n;
    size_t wi = 1;
    for (size_t gap = 1, m = n >> 1; m > 1; gap <<= 1, m >>= 1) {
        for (size_t base = 0; base < n; base += gap << 1) {
            const uint64_t w = t->irp[wi];
            const uint64_t wp = po_multiply_factor(w, 64, q);
            ++wi;
            for (size_t i = base; i < base + gap; ++i) {
                /* inv_butterfly_radix2, radix2.rs:210-218 */
                const uint64_t tx = x[i] + x[i + gap];
                const uint64_t y = x[i] + q2 - x[i + gap];
                x[i] = reduce_once(tx, q2);
                x[i + gap] = po_multiply_mod_lazy(y, w, wp, 64, q);
            }
        }
    }
    const uint64_t w = t->irp[wi];
    const uint64_t inv_n = t->inv_n, inv_n_p = po_multiply_factor(inv_n, 64, q);
    /* MultiplyFactor::mul_modulo of w by N^-1: canonical */
    const uint64_t inv_n_w = reduce_once(po_multiply_mod_lazy(w, inv_n, inv_n_p, 64, q), q);
    const uint64_t inv_n_w_p = po_multiply_factor(inv_n_w, 64, q);
    const size_t h = n >> 1;
    for (size_t i = 0; i < h; ++i) {
        const uint64_t tx = reduce_once(x[i] + x[i + h], q2);
        const uint64_t ty = x[i] + q2 - x[i + h];
        x[i] = po_multiply_mod_lazy(tx, inv_n, inv_n_p, 64, q);
        x[i + h] = po_multiply_mod_lazy(ty, inv_n_w, inv_n_w_p, 64, q);
    }
    if (output_mod_factor == 1)
        for (size_t i = 0; i < n; ++i) x[i] = reduce_once(x[i], q);
}

/* transform_monomial: out[k] = coeff * psi^(((2 rev(k) + 1) * degree) mod 2N).
 * primus_ntt/src/ntt/primitive.rs:298-347 */
void po_transform_monomial(const po_table *t, uint64_t coeff, size_t degree, uint64_t *values) {
    size_t n = t->n;
    if (coeff == 0) { memset(values, 0, n * 8); return; }
    if (degree == 0) { for (size_t i = 0; i < n; ++i) values[i] = coeff; return; }
    size_t mask = (2 * n) - 1;
    for (size_t k = 0; k < n; ++k) {
        size_t idx = ((2 * t->rev[k] + 1) * degree) & mask;
        if (coeff == 1) values[k] = t->op[idx];
        else if (coeff == t->q - 1) values[k] = t->op[idx ^ n];
        else values[k] = po_shoup_mul(t->op[idx], t->op_q[idx], coeff, t->q);
    }
}

/* transform_coeff_one_monomial.  primitive.rs:349-372 */
void po_transform_coeff_one_monomial(const po_table *t, size_t degree, uint64_t *values) {
    size_t n = t->n;
    if (degree == 0) { for (size_t i = 0; i < n; ++i) values[i] = 1; return; }
    size_t mask = (2 * n) - 1;
    for (size_t k = 0; k < n; ++k) values[k] = t->op[((2 * t->rev[k] + 1) * degree) & mask];
}

/* transform_coeff_minus_one_monomial.  primitive.rs:374-397 */
void po_transform_coeff_minus_one_monomial(const po_table *t, size_t degree, uint64_t *values) {
    size_t n = t->n;
    if (degree == 0) { for (size_t i = 0; i < n; ++i) values[i] = t->q - 1; return; }
    size_t mask = (2 * n) - 1;
    for (size_t k = 0; k < n; ++k) values[k] = t->op[(((2 * t->rev[k] + 1) * degree) & mask) ^ n];
}

/* ------------------------------------------------------------------------- */
/* Element-wise and schoolbook                                               */
/* ------------------------------------------------------------------------- */

/* NttPolynomial::mul_assign / mul_inplace (Barrett reduce_mul per element).
 * primus_poly/src/ntt/mul.rs:87-119 */
void po_pointwise_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n, uint64_t q) {
    uint64_t ratio[2];
    po_barrett_ratio(q, ratio);
    for (size_t i = 0; i < n; ++i) out[i] = po_barrett_reduce_mul(a[i], b[i], q, ratio);
}

/* NttPolynomial::add_mul_assign: acc = (a*b + acc) mod q.
 * primus_poly/src/ntt/mod.rs:97-108 -> primus_modulus/src/barrett/slice.rs:175-188 */
void po_pointwise_mul_add(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t n,
                          uint64_t q) {
    uint64_t ratio[2];
    po_barrett_ratio(q, ratio);
    for (size_t i = 0; i < n; ++i) acc[i] = po_barrett_reduce_mul_add(a[i], b[i], acc[i], q, ratio);
}

/* Per-element Shoup multiply x[i] = w[i]*x[i] mod q (canonical).
 * primus_poly/src/dcrt/mul.rs:119-155, primus_factor/src/shoup_factor/mod.rs:121-124 */
void po_shoup_pointwise(const uint64_t *w, const uint64_t *wq, uint64_t *x, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; ++i) x[i] = po_shoup_mul(w[i], wq[i], x[i], q);
}

/* Polynomial::naive_mul_inplace: c = a*b mod (X^N + 1, q), c starts at zero.
 * primus_poly/src/poly/mul.rs:111-142 */
void po_schoolbook_negacyclic(const uint64_t *a, const uint64_t *b, uint64_t *c, size_t n,
                              uint64_t q) {
    uint64_t ratio[2];
    po_barrett_ratio(q, ratio);
    memset(c, 0, n * 8);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j <= i; ++j) c[i] = po_barrett_reduce_mul_add(a[j], b[i - j], c[i], q, ratio);
    for (size_t i = n; i + 1 < 2 * n; ++i) {
        size_t k = i - n;
        for (size_t j = i - n + 1; j < n; ++j) {
            uint64_t p = po_barrett_reduce_mul(a[j], b[i - j], q, ratio);
            c[k] = c[k] >= p ? c[k] - p : c[k] + q - p; /* reduce_sub_assign, ops.rs:268-276 */
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Batched CPU baseline                                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    po_table *const *tables;
    uint32_t L;
    uint64_t *polys;
    size_t begin, end;
    const uint64_t *operand, *operand_q;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    for (size_t i = j->begin; i < j->end; ++i) {
        const po_table *t = j->tables[i % j->L];
        size_t n = t->n;
        uint64_t *p = j->polys + i * n;
        const uint64_t *w = j->operand + (i % j->L) * n;
        const uint64_t *wq = j->operand_q + (i % j->L) * n;
        po_transform(t, p);
        po_shoup_pointwise(w, wq, p, n, t->q);
        po_inverse_transform(t, p);
    }
    return NULL;
}

/* UintCrtNttTable::transform_slice / inverse_transform_slice over limb-major
 * chunks (primus_ntt/src/dcrt/primitive.rs:117-141) with a Shoup pointwise
 * multiply between them, split over host threads. */
void po_fwd_mul_inv_batch(po_table *const *tables, uint32_t L, uint64_t *polys, size_t count,
                          const uint64_t *operand, const uint64_t *operand_q, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    batch_job jobs[256];
    size_t per = (count + threads - 1) / threads;
    int started = 0;
    for (int k = 0; k < threads; ++k) {
        size_t b = (size_t)k * per, e = b + per;
        if (b >= count) break;
        if (e > count) e = count;
        jobs[k] = (batch_job){tables, L, polys, b, e, operand, operand_q};
        if (threads == 1) { batch_worker(&jobs[k]); continue; }
        pthread_create(&tid[k], NULL, batch_worker, &jobs[k]);
        ++started;
    }
    for (int k = 0; k < started; ++k) pthread_join(tid[k], NULL);
}

/* ------------------------------------------------------------------------- */
/* Gadget decomposition and the external product                             */
/* ------------------------------------------------------------------------- */

/* BigUintApproxSignedBasis::new for a one-limb modulus.
 * primus_decompose/src/big_integer/basis.rs:34-195 */
int po_basis_new(uint64_t q, uint32_t log_basis, uint32_t l, po_basis *out) {
    if (q < 2 || log_basis == 0 || log_basis >= 64 || l == 0) return PO_ERR_INVALID;
    uint32_t bits = 64 - (uint32_t)__builtin_clzll(q);
    uint32_t full = bits / log_basis;
    if (full < l) return PO_ERR_INVALID; /* basis.rs:57: decompose_length >= reverse_len */
    po_basis b;
    memset(&b, 0, sizeof b);
    b.q = q;
    b.log_basis = log_basis;
    b.l = l;
    b.modulus_bits = bits;
    b.drop_bits = bits - l * log_basis;
    b.basis = 1ull << log_basis;
    b.basis_minus_one = b.basis - 1;
    b.carry_mask = log_basis == 1 ? 2ull : ((1ull << log_basis) | (1ull << (log_basis - 1))); /* :75-79 */
    /* split_value (:81-128) */
    u128 value = 0;
    if (log_basis == 1) {
        if (b.drop_bits == 0) {
            b.has_split = 0;
        } else {
            for (uint32_t i = 0; i < l; ++i) value = (value << 1) | 1;
            value = (value << 1) | 1;
            value <<= (b.drop_bits - 1);
            b.has_split = value < q;
        }
    } else {
        for (uint32_t i = 0; i < l; ++i) value = (value << log_basis) | (b.basis_minus_one >> 1);
        if (b.drop_bits > 0) {
            value = (value << 1) | 1;
            value <<= (b.drop_bits - 1);
        } else {
            value += 1;
        }
        b.has_split = value < q;
    }
    b.split = (uint64_t)value;
    /* next_pow_of_2_sub_modulus = (2^bits - 1) - (q - 1) (:134-144) */
    b.pow2_sub_modulus = (uint64_t)((((u128)1) << bits) - q);
    *out = b;
    return PO_OK;
}

/* basis.rs:307-331, common.rs:141-152,184-201, primus_rns/src/base.rs:201-236 */
void po_decompose(const po_basis *b, const uint64_t *values, size_t n, uint64_t *residues) {
    const uint64_t half = (b->basis + 1) / 2;
    for (size_t i = 0; i < n; ++i) {
        uint64_t adj = values[i];
        if (b->has_split && adj >= b->split) adj += b->pow2_sub_modulus;
        int carry = b->drop_bits > 0 ? (int)((adj >> (b->drop_bits - 1)) & 1) : 0;
        for (uint32_t d = 0; d < b->l; ++d) {
            uint32_t shift = b->drop_bits + d * b->log_basis;
            uint64_t temp = ((adj >> shift) & b->basis_minus_one) + (uint64_t)carry;
            carry = (temp & b->carry_mask) != 0;
            uint64_t digit = temp & b->basis_minus_one;
            uint64_t r;
            if (b->basis != 2) r = digit < half ? digit : (b->q - b->basis) + digit;
            else r = digit;
            residues[(size_t)d * n + i] = r;
        }
    }
}

void po_decompose_unsigned(const po_basis *b, const uint64_t *values, size_t n, uint64_t *residues) {
    for (size_t i = 0; i < n; ++i)
        for (uint32_t d = 0; d < b->l; ++d)
            residues[(size_t)d * n + i] =
                ((values[i] >> (b->drop_bits + d * b->log_basis)) & b->basis_minus_one) % b->q;
}

static void ext_product_digits(const po_table *t, const po_basis *b, const uint64_t *key, const uint64_t *in,
                               uint64_t *out, int ntt_output, int unsigned_digits) {
    const size_t n = t->n;
    uint64_t *res = (uint64_t *)malloc(b->l * n * 8);
    memset(out, 0, 2 * n * 8);
    for (int p = 0; p < 2; ++p) {
        if (unsigned_digits) po_decompose_unsigned(b, in + p * n, n, res);
        else po_decompose(b, in + p * n, n, res);
        for (uint32_t d = 0; d < b->l; ++d) {
            uint64_t *x = res + (size_t)d * n;
            po_transform(t, x); /* table.transform_slice(multi_residues), dcrt.rs:238 */
            for (int j = 0; j < 2; ++j) /* add_dcrt_glwe_mul_dcrt_polynomial_assign, dcrt.rs:101-119 */
                po_pointwise_mul_add(out + j * n, x, key + (((size_t)p * b->l + d) * 2 + j) * n, n, t->q);
        }
    }
    if (!ntt_output)
        for (int j = 0; j < 2; ++j) po_inverse_transform(t, out + j * n);
    free(res);
}

void po_ext_product(const po_table *t, const po_basis *b, const uint64_t *key, const uint64_t *in,
                    uint64_t *out, int ntt_output) {
    ext_product_digits(t, b, key, in, out, ntt_output, 0);
}

void po_ext_product_unsigned(const po_table *t, const po_basis *b, const uint64_t *key, const uint64_t *in,
                             uint64_t *out, int ntt_output) {
    ext_product_digits(t, b, key, in, out, ntt_output, 1);
}

/* ------------------------------------------------------------------------- */
/* RLWE encryption with supplied samples                                     */
/* ------------------------------------------------------------------------- */

/* NttRlweSecretKey::encrypt_inplace_with_embedding, unsigned embedding, with
 * the samples given instead of drawn (primus_fhe_core/src/secret_key/rlwe.rs:227-268):
 *   b = e mod q                          (sample_gaussian_values_inplace, :245)
 *   b = b + (delta * m mod q) mod q      (add_mul_factor_assign, Shoup, :247-260)
 *   b = NTT(b)                           (transform_slice, :263)
 *   b = b + a (.) s_hat (mod q)          (add_mul_assign, Barrett, :267)
 * s_hat = NTT(s mod q) is the NTT-form secret of the key.  a_hat is the uniform
 * polynomial already in NTT form (:265). */
void po_rlwe_encrypt(const po_table *t, const int8_t *s, const int8_t *e, const uint8_t *m, uint64_t delta,
                     const uint64_t *a_hat, uint64_t *b_hat) {
    const size_t n = t->n;
    const uint64_t q = t->q;
    uint64_t *sh = (uint64_t *)malloc(n * 8);
    for (size_t i = 0; i < n; ++i) sh[i] = s[i] >= 0 ? (uint64_t)s[i] : q - (uint64_t)(-(int)s[i]);
    po_transform(t, sh);
    uint64_t dq = po_shoup_quotient(delta, q);
    for (size_t i = 0; i < n; ++i) {
        uint64_t b = e[i] >= 0 ? (uint64_t)e[i] : q - (uint64_t)(-(int)e[i]);
        b = reduce_once(b + po_shoup_mul(delta, dq, m[i], q), q);
        b_hat[i] = b;
    }
    po_transform(t, b_hat);
    po_pointwise_mul_add(b_hat, a_hat, sh, n, q);
    free(sh);
}

/* ------------------------------------------------------------------------- */
/* NTT-domain automorphism                                                   */
/* ------------------------------------------------------------------------- */

/* x -> x^degree on a polynomial in bit-reversed NTT order, degree odd.
 * generate_ntt_permutation (primus_fhe_core/src/automorphism/glwe/dcrt.rs:43-61):
 * natural slot i holds the evaluation at psi^(2i+1); sigma(f) evaluated there
 * is f at psi^(degree*(2i+1) mod 2N), natural slot (j-1)/2; in bit-reversed
 * storage out[rev(i)] = in[rev((j-1)/2)].  ntt_poly_auto_inplace (:97-112)
 * applies the gather. */
void po_ntt_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t degree) {
    uint32_t log_n = 0;
    while (((size_t)1 << log_n) < n) ++log_n;
    const uint64_t two_n = 2 * (uint64_t)n;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t j = (uint64_t)(((unsigned __int128)(degree % two_n) * (2 * i + 1)) % two_n);
        const uint64_t target = (j - 1) / 2;
        out[po_reverse_lsbs(i, log_n)] = in[po_reverse_lsbs(target, log_n)];
    }
}

/* The same automorphism in the coefficient domain, a(X) -> a(X^degree) mod
 * (X^N + 1, q): coefficient j moves to j*degree mod 2N, negated when that is
 * N or more (X^N = -1).  Used to pin po_ntt_automorphism through the
 * transform: NTT(a(X^k)) must equal the permuted NTT(a). */
void po_coeff_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t degree, uint64_t q) {
    const uint64_t two_n = 2 * (uint64_t)n;
    for (size_t j = 0; j < n; ++j) {
        const uint64_t e = (uint64_t)(((unsigned __int128)(degree % two_n) * j) % two_n);
        const uint64_t v = in[j] % q;
        if (e < n) out[e] = v;
        else out[e - n] = v ? q - v : 0;
    }
}

/* ------------------------------------------------------------------------- */
/* RNS base conversion                                                       */
/* ------------------------------------------------------------------------- */

struct po_base_converter {
    uint32_t in_count, out_count;
    uint64_t *q, *p;          /* input / output moduli */
    uint64_t *inv, *inv_q;    /* [(Q/q_i)^-1]_{q_i} as ShoupFactor */
    uint64_t *m;              /* [out][in]: (Q/q_i) mod p_j */
    uint64_t *q_mod_p;        /* [out]: Q mod p_j */
};

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((unsigned __int128)a * b) % m);
}

void po_base_converter_free(po_base_converter *c) {
    if (!c) return;
    free(c->q);
    free(c->p);
    free(c->inv);
    free(c->inv_q);
    free(c->m);
    free(c->q_mod_p);
    free(c);
}

/* BaseConverter::new (primus_rns/src/converter.rs:21-46) over RNSBase::new
 * (primus_rns/src/base.rs:50-90): the input moduli pairwise coprime, the
 * punctured products Q/q_i kept modulo each modulus they meet. */
int po_base_converter_new(const uint64_t *ibase, uint32_t in_count, const uint64_t *obase, uint32_t out_count,
                          po_base_converter **out) {
    if (!out || !ibase || !obase || in_count == 0 || out_count == 0) return PO_ERR_INVALID;
    *out = NULL;
    for (uint32_t i = 0; i < in_count; ++i) {
        if (ibase[i] <= 1 || ibase[i] >= (1ull << 62)) return PO_ERR_INVALID;
        for (uint32_t k = i + 1; k < in_count; ++k)
            if (gcd_u64(ibase[i], ibase[k]) != 1) return PO_ERR_INVALID; /* RNSError::CoPrimeError */
    }
    for (uint32_t j = 0; j < out_count; ++j)
        if (obase[j] <= 1 || obase[j] >= (1ull << 62)) return PO_ERR_INVALID;
    po_base_converter *c = (po_base_converter *)calloc(1, sizeof *c);
    if (!c) return PO_ERR_INVALID;
    c->in_count = in_count;
    c->out_count = out_count;
    c->q = (uint64_t *)malloc(in_count * 8);
    c->p = (uint64_t *)malloc(out_count * 8);
    c->inv = (uint64_t *)malloc(in_count * 8);
    c->inv_q = (uint64_t *)malloc(in_count * 8);
    c->m = (uint64_t *)malloc((size_t)in_count * out_count * 8);
    c->q_mod_p = (uint64_t *)malloc(out_count * 8);
    for (uint32_t i = 0; i < in_count; ++i) {
        c->q[i] = ibase[i];
        uint64_t punct = 1 % ibase[i];
        for (uint32_t k = 0; k < in_count; ++k)
            if (k != i) punct = mulmod_u64(punct, ibase[k] % ibase[i], ibase[i]);
        c->inv[i] = po_inv_mod(punct, ibase[i]); /* try_inv_modulo, base.rs:78 */
        c->inv_q[i] = po_shoup_quotient(c->inv[i], ibase[i]);
    }
    for (uint32_t j = 0; j < out_count; ++j) {
        c->p[j] = obase[j];
        uint64_t all = 1 % obase[j];
        for (uint32_t k = 0; k < in_count; ++k) all = mulmod_u64(all, ibase[k] % obase[j], obase[j]);
        c->q_mod_p[j] = all;
        for (uint32_t i = 0; i < in_count; ++i) {
            uint64_t punct = 1 % obase[j];
            for (uint32_t k = 0; k < in_count; ++k)
                if (k != i) punct = mulmod_u64(punct, ibase[k] % obase[j], obase[j]);
            c->m[(size_t)j * in_count + i] = punct; /* m_i.modulo(modulus), converter.rs:35-41 */
        }
    }
    *out = c;
    return PO_OK;
}

/* y_i = [x_i * (Q/q_i)^-1]_{q_i}: fill_fast_convert_array_buffer
 * (converter.rs:103-146); a factor of one reduces x instead of multiplying. */
static uint64_t bc_scaled(const po_base_converter *c, uint32_t i, uint64_t x) {
    if (c->inv[i] == 1) return x % c->q[i];
    return po_shoup_mul(c->inv[i], c->inv_q[i], x, c->q[i]);
}

/* BaseConverter::fast_convert_array (converter.rs:148-177): in [in_count][n],
 * out [out_count][n]; out_j[k] = sum_i y_i[k] * ((Q/q_i) mod p_j) mod p_j
 * (reduce_dot_product: the canonical dot product). */
void po_fast_convert_array(const po_base_converter *c, const uint64_t *in, uint64_t *out, size_t n) {
    uint64_t *y = (uint64_t *)malloc(c->in_count * 8);
    for (size_t k = 0; k < n; ++k) {
        for (uint32_t i = 0; i < c->in_count; ++i) y[i] = bc_scaled(c, i, in[(size_t)i * n + k]);
        for (uint32_t j = 0; j < c->out_count; ++j) {
            unsigned __int128 acc = 0;
            for (uint32_t i = 0; i < c->in_count; ++i)
                acc = (acc + (unsigned __int128)y[i] * c->m[(size_t)j * c->in_count + i]) % c->p[j];
            out[(size_t)j * n + k] = (uint64_t)acc;
        }
    }
    free(y);
}

/* BaseConverter::exact_convert_array (converter.rs:222-316), one output
 * modulus p: v = round(sum_i y_i / q_i) in binary64 (each term divided as
 * written, summed in limb order, plus 0.5, truncated), then
 * out = (sum_i y_i * ((Q/q_i) mod p) - v * (Q mod p)) mod p. */
void po_exact_convert_array(const po_base_converter *c, const uint64_t *in, uint64_t *out, size_t n) {
    const uint64_t p = c->p[0];
    uint64_t *y = (uint64_t *)malloc(c->in_count * 8);
    for (size_t k = 0; k < n; ++k) {
        double sum = -0.0;
        for (uint32_t i = 0; i < c->in_count; ++i) {
            /* the else-branch multiplies with a full 128-bit remainder (mul_modulo,
             * converter.rs:268-276); both equal the Shoup product's canonical value */
            y[i] = bc_scaled(c, i, in[(size_t)i * n + k]);
            sum += (double)y[i] / (double)c->q[i];
        }
        const uint64_t v = (uint64_t)(sum + 0.5);
        unsigned __int128 acc = 0;
        for (uint32_t i = 0; i < c->in_count; ++i) acc = (acc + (unsigned __int128)y[i] * c->m[i]) % p;
        const uint64_t vq = mulmod_u64(v % p, c->q_mod_p[0], p);
        const uint64_t s = (uint64_t)acc;
        out[k] = s >= vq ? s - vq : s + p - vq;
    }
    free(y);
}

/* ------------------------------------------------------------------------ */
/* Sampling of RLWE encryption inputs                                        */
/* ------------------------------------------------------------------------ */

void po_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t a = (uint64_t)0xD2511F53u * x0; /* Philox multipliers */
        const uint64_t b = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        const uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        x1 = (uint32_t)b;
        x3 = (uint32_t)a;
        x0 = y0;
        x2 = y2;
        k0 += 0x9E3779B9u; /* Weyl key increments */
        k1 += 0xBB67AE85u;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

void po_gauss_thresholds(uint64_t t[38]) {
    const double sigma = 3.2;
    double w[39], z = 0.0;
    for (int k = -19; k <= 19; ++k) {
        w[k + 19] = exp(-(double)(k * k) / (2.0 * sigma * sigma));
        z += w[k + 19];
    }
    double cum = 0.0;
    for (int i = 0; i < 38; ++i) {
        cum += w[i];
        t[i] = (uint64_t)ldexp(cum / z, 64);
    }
}

static void draw(uint64_t seed, uint32_t request, uint32_t purpose, uint32_t k, uint64_t *lo, uint64_t *hi) {
    const uint32_t ctr[4] = {k, request, purpose, 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    po_philox4x32_10(ctr, key, o);
    *lo = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
    *hi = (uint64_t)o[2] | ((uint64_t)o[3] << 32);
}

void po_rlwe_sample(uint64_t seed, uint32_t request, uint32_t log_n, uint32_t limbs, const uint64_t *moduli,
                    int8_t *s, int8_t *e, uint8_t *m, uint64_t *a_hat) {
    const uint32_t n = 1u << log_n;
    uint64_t t[38];
    po_gauss_thresholds(t);
    for (uint32_t k = 0; k < n; ++k) {
        uint64_t lo, hi;
        draw(seed, request, 0, k, &lo, &hi);
        s[k] = (int8_t)((int)(((unsigned __int128)lo * 3u) >> 64) - 1); /* floor(3u / 2^64) - 1 */
        draw(seed, request, 1, k, &lo, &hi);
        int v = -19; /* inversion of the cumulative table */
        while (v < 19 && lo >= t[v + 19]) ++v;
        e[k] = (int8_t)v;
        draw(seed, request, 2, k, &lo, &hi);
        m[k] = (uint8_t)(lo & 255u);
        for (uint32_t j = 0; j < limbs; ++j) {
            draw(seed, request, 3 + j, k, &lo, &hi);
            /* floor((hi:lo) * q / 2^128) */
            const unsigned __int128 prod_hi = (unsigned __int128)hi * moduli[j];
            const unsigned __int128 prod_lo = ((unsigned __int128)lo * moduli[j]) >> 64;
            a_hat[(size_t)j * n + k] = (uint64_t)((prod_hi + prod_lo) >> 64);
        }
    }
}
