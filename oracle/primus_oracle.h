/*
 * primus_oracle.h -- CPU restatement of Primus's negacyclic NTT and the
 * modular arithmetic under it.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library (libprimus_hip.so) is compared against, and the timed CPU baseline
 * in bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  No product code path calls into it.
 *
 * Every function cites the reference file:line it restates.  The reference
 * is a Rust workspace (serendipity-crypto/primus); paths are relative to its
 * root.
 */
#ifndef PRIMUS_ORACLE_H
#define PRIMUS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical numbering to include/primus_hip.h. */
#define PO_OK 0
#define PO_ERR_NO_PRIMITIVE_ROOT 1  /* NttError::NoPrimitiveRoot    primus_ntt/src/error.rs:9 */
#define PO_ERR_DEGREE_CONVERSION 2  /* NttError::DegreeConversionErr primus_ntt/src/error.rs:17 */
#define PO_ERR_DEGREE_TOO_LARGE 3   /* NttError::DegreeTooLarge     primus_ntt/src/error.rs:25 */
#define PO_ERR_NTT_TABLE 4          /* NttError::NttTableErr        primus_ntt/src/error.rs:33 */
#define PO_ERR_INVALID 5            /* bad argument (BarrettModulus::new panics) */

/* ---- Shoup factor: primus_factor/src/shoup_factor/mod.rs:32-125 ---- */
uint64_t po_shoup_quotient(uint64_t w, uint64_t q);
uint64_t po_shoup_lazy_mul(uint64_t w, uint64_t wq, uint64_t b, uint64_t q);
uint64_t po_shoup_mul(uint64_t w, uint64_t wq, uint64_t b, uint64_t q);

/* ---- MultiplyFactor (HEXL Barrett factor): primus_factor/src/multiply.rs:4-76 ---- */
uint64_t po_multiply_factor(uint64_t op, uint32_t bit_shift, uint64_t q);
uint64_t po_multiply_mod_lazy(uint64_t x, uint64_t y, uint64_t y_barrett, uint32_t bit_shift, uint64_t q);

/* ---- BarrettModulus<u64>: primus_modulus/src/barrett/{mod.rs,ops.rs} ---- */
int po_barrett_ratio(uint64_t q, uint64_t ratio[2]);
uint64_t po_barrett_lazy_reduce(uint64_t value, uint64_t q, const uint64_t ratio[2]);
uint64_t po_barrett_lazy_reduce_2limb(uint64_t lo, uint64_t hi, uint64_t q, const uint64_t ratio[2]);
uint64_t po_barrett_reduce_mul(uint64_t a, uint64_t b, uint64_t q, const uint64_t ratio[2]);
uint64_t po_barrett_reduce_mul_add(uint64_t a, uint64_t b, uint64_t c, uint64_t q,
                                   const uint64_t ratio[2]);
uint64_t po_pow_mod(uint64_t base, uint64_t exp, uint64_t q);
uint64_t po_inv_mod(uint64_t a, uint64_t q);

/* ---- PrimitiveRoot / ReverseLsbs: primus_ntt/src/{root.rs,reverse.rs} ---- */
int po_is_primitive_root(uint64_t w, uint32_t log_degree, uint64_t q);
int po_minimal_primitive_root(uint32_t log_degree, uint64_t q, uint64_t *root);
uint64_t po_reverse_lsbs(uint64_t x, uint32_t bits);

/* ---- UintNttTable<u64>: primus_ntt/src/ntt/primitive.rs:34-398 ---- */
typedef struct po_table po_table;
int po_table_new(uint32_t log_n, uint64_t q, po_table **out);
void po_table_free(po_table *t);
size_t po_table_n(const po_table *t);
uint64_t po_table_modulus(const po_table *t);
uint64_t po_table_root(const po_table *t);
uint64_t po_table_inv_root(const po_table *t);
void po_table_inv_n(const po_table *t, uint64_t *value, uint64_t *quotient);
/* which: 0 = root_powers (N), 1 = inv_root_powers (N), 2 = ordinal_root_powers (2N) */
void po_table_powers(const po_table *t, int which, uint64_t *values, uint64_t *quotients);

void po_lazy_transform(const po_table *t, uint64_t *poly);
void po_transform(const po_table *t, uint64_t *poly);
void po_lazy_inverse_transform(const po_table *t, uint64_t *values);
void po_inverse_transform(const po_table *t, uint64_t *values);
/* HexlNttTable's radix-2 path (hexl/radix2.rs): forward output_mod_factor 1
 * or 4, inverse output_mod_factor 1 or 2. */
void po_hexl_forward(const po_table *t, uint64_t *x, uint32_t output_mod_factor);
void po_hexl_inverse(const po_table *t, uint64_t *x, uint32_t output_mod_factor);
void po_transform_monomial(const po_table *t, uint64_t coeff, size_t degree, uint64_t *values);
void po_transform_coeff_one_monomial(const po_table *t, size_t degree, uint64_t *values);
void po_transform_coeff_minus_one_monomial(const po_table *t, size_t degree, uint64_t *values);

/* ---- primus_poly element-wise ops and the schoolbook product ---- */
void po_pointwise_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n, uint64_t q);
void po_pointwise_mul_add(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t n, uint64_t q);
void po_shoup_pointwise(const uint64_t *w, const uint64_t *wq, uint64_t *x, size_t n, uint64_t q);
void po_schoolbook_negacyclic(const uint64_t *a, const uint64_t *b, uint64_t *c, size_t n,
                              uint64_t q);

/* ---- Batched CPU baselines (the timed CPU leg) ---- */
/* polys: count limb-polynomials of N coefficients, limb-major: poly i uses table[i % L].
 * operand / operand_q: L*N Shoup-prepared NTT-domain operand.  Runs forward,
 * Shoup pointwise multiply and inverse for each limb-polynomial on `threads`
 * host threads. */
void po_fwd_mul_inv_batch(po_table *const *tables, uint32_t L, uint64_t *polys, size_t count,
                          const uint64_t *operand, const uint64_t *operand_q, int threads);


/* ---- Approximate signed gadget basis for one u64 limb ----
 * BigUintApproxSignedBasis::new(q, log_basis, Some(l), rns) with a single
 * modulus (primus_decompose/src/big_integer/basis.rs:34-195). */
typedef struct {
    uint64_t q;
    uint32_t log_basis, l, drop_bits, modulus_bits;
    uint64_t basis, basis_minus_one, carry_mask;
    uint64_t split;            /* valid when has_split */
    int has_split;
    uint64_t pow2_sub_modulus; /* 2^modulus_bits - q */
} po_basis;
int po_basis_new(uint64_t q, uint32_t log_basis, uint32_t l, po_basis *out);
/* init_value_carry_slice_inplace + l x (unsigned_decompose_slice_inplace ->
 * wrapping_decompose_small_values_inplace): residues[l][n] in [0, q).
 * basis.rs:307-331, common.rs:184-201, primus_rns/src/base.rs:201-236 */
void po_decompose(const po_basis *b, const uint64_t *values, size_t n, uint64_t *residues);
/* External product RLWE x RGSW (k = 1), single modulus:
 * CrtGlwe::mul_dcrt_ggsw_inplace (primus_lattice/src/glwe/crt.rs:202-229) ->
 * DcrtGlwe::add_dcrt_glev_mul_crt_poly_assign (primus_lattice/src/glwe/dcrt.rs:170-247),
 * then the caller's inverse transform when ntt_output == 0.
 * in: [2][N] coefficients < q; key: [2][l][2][N] NTT form < q; out: [2][N]. */
void po_ext_product(const po_table *t, const po_basis *b, const uint64_t *key, const uint64_t *in,
                    uint64_t *out, int ntt_output);
/* The literal reading of "l unsigned digits of base B": digit d of a value
 * v < q is (v >> (drop_bits + d*log_basis)) & (B - 1), with no rounding carry,
 * no split adjustment and no signed mapping (no reference counterpart: the
 * reference's only external product uses the approximate signed digits
 * above).  residues[l][n], each digit reduced mod q. */
void po_decompose_unsigned(const po_basis *b, const uint64_t *values, size_t n, uint64_t *residues);
/* po_ext_product with these digits. */
void po_ext_product_unsigned(const po_table *t, const po_basis *b, const uint64_t *key, const uint64_t *in,
                             uint64_t *out, int ntt_output);

#ifdef __cplusplus
}
#endif
#endif
#ifndef PRIMUS_ORACLE_AUTO_H
#define PRIMUS_ORACLE_AUTO_H
#ifdef __cplusplus
extern "C" {
#endif
/* NTT-domain automorphism x -> x^degree (degree odd), bit-reversed order:
 * primus_fhe_core/src/automorphism/glwe/dcrt.rs:43-61, 97-112 */
void po_ntt_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t degree);
/* The coefficient-domain automorphism a(X) -> a(X^degree) mod (X^N+1, q) (for pinning) */
void po_coeff_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t degree, uint64_t q);
#ifdef __cplusplus
}
#endif
#endif
#ifndef PRIMUS_ORACLE_RNS_H
#define PRIMUS_ORACLE_RNS_H
#ifdef __cplusplus
extern "C" {
#endif
/* BaseConverter (primus_rns/src/converter.rs): new :21-46, fast_convert_array
 * :148-177, exact_convert_array :222-316 (one output modulus).  Limb-major
 * [in_count][n] -> [out_count][n]. */
typedef struct po_base_converter po_base_converter;
int po_base_converter_new(const uint64_t *ibase, uint32_t in_count, const uint64_t *obase, uint32_t out_count,
                          po_base_converter **out);
void po_base_converter_free(po_base_converter *c);
void po_fast_convert_array(const po_base_converter *c, const uint64_t *in, uint64_t *out, size_t n);
void po_exact_convert_array(const po_base_converter *c, const uint64_t *in, uint64_t *out, size_t n);
#ifdef __cplusplus
}
#endif
#endif
#ifndef PRIMUS_ORACLE_RLWE_H
#define PRIMUS_ORACLE_RLWE_H
#ifdef __cplusplus
extern "C" {
#endif
/* RLWE encryption with supplied samples: b_hat = NTT(e + delta*m) + a_hat (.) NTT(s).
 * primus_fhe_core/src/secret_key/rlwe.rs:227-268 */
void po_rlwe_encrypt(const po_table *t, const int8_t *s, const int8_t *e, const uint8_t *m, uint64_t delta,
                     const uint64_t *a_hat, uint64_t *b_hat);

/* Sampling of the encryption inputs from a counter-based generator (the batch
 * workload's distributions; the reference samples on the host with rand's
 * thread RNG, primus_distr/src/common.rs:56-121).  Philox4x32-10 as published
 * by Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2,
 * 3" (SC 2011). */
void po_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The 38 cumulative thresholds of the rounded Gaussian, sigma 3.2, |e| <= 19. */
void po_gauss_thresholds(uint64_t t[38]);
/* One request's s, e, m [N] and a_hat [limbs][N]: element k of what-is-drawn p
 * (0 s, 1 e, 2 m, 3 + j a_hat limb j) comes from the block of counter
 * (k, request, p, 0) under key (seed low, seed high). */
void po_rlwe_sample(uint64_t seed, uint32_t request, uint32_t log_n, uint32_t limbs, const uint64_t *moduli,
                    int8_t *s, int8_t *e, uint8_t *m, uint64_t *a_hat);
#ifdef __cplusplus
}
#endif
#endif
