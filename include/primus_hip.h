/*
 * primus_hip.h -- C ABI of libprimus_hip.so: Primus's negacyclic NTT tables
 * and the transform pipelines over them, running on an AMD MI355X (gfx950).
 *
 * The entry points mirror the reference's Rust traits (serendipity-crypto/primus):
 *   trait NttTable   primus_ntt/src/ntt/mod.rs:20-117, impl UintNttTable
 *                    primus_ntt/src/ntt/primitive.rs:110-398
 *   trait DcrtTable  primus_ntt/src/dcrt/mod.rs:17-129, impl UintCrtNttTable
 *                    primus_ntt/src/dcrt/primitive.rs:18-142
 * with the reduction semantics of primus_reduce / primus_modulus (Barrett) and
 * primus_factor (Shoup).  Each declaration cites the reference item it replaces.
 *
 * Conventions
 *   - Every fallible entry point returns an int status (PH_OK = 0).  Codes 1-4
 *     mirror NttError (primus_ntt/src/error.rs:7-35); the reference panics on a
 *     bad modulus (primus_modulus/src/barrett/mod.rs:55-60) and debug-asserts
 *     slice lengths, both reported here as PH_ERR_INVALID.  Nothing aborts.
 *   - Tables are immutable after creation and may be shared between threads
 *     (the Rust traits are Send + Sync).  Their twiddles live in device memory.
 *   - Host-slice entry points (`*_slice`, `*_monomial`) take host memory, run on
 *     the GPU and return when the result is back in the slice.
 *   - Device entry points (`*_device`) take device pointers and a HIP stream
 *     (hipStream_t passed as void*, NULL = the default stream); they only
 *     enqueue work.
 *   - Coefficients are uint64_t; multi-limb data is limb-major [L][N]
 *     (primus_poly/src/dcrt/mod.rs:70-76).  In the batched device forms
 *     limb-polynomial i of a batch uses modulus i % L.
 *   - Forward transforms take normal order to bit-reversed order, inverse
 *     transforms the reverse; slot k of a forward output is the evaluation at
 *     psi^(2*rev(k)+1), psi the smallest primitive 2N-th root of unity.
 *   - Supported ring degrees: N = 2^1 .. 2^17; modulus 1 < q < 2^62.
 */
#ifndef PRIMUS_HIP_H
#define PRIMUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PH_OK 0
#define PH_ERR_NO_PRIMITIVE_ROOT 1 /* NttError::NoPrimitiveRoot     primus_ntt/src/error.rs:9 */
#define PH_ERR_DEGREE_CONVERSION 2 /* NttError::DegreeConversionErr primus_ntt/src/error.rs:17 */
#define PH_ERR_DEGREE_TOO_LARGE 3  /* NttError::DegreeTooLarge      primus_ntt/src/error.rs:25 */
#define PH_ERR_NTT_TABLE 4         /* NttError::NttTableErr         primus_ntt/src/error.rs:33 */
#define PH_ERR_INVALID 5           /* bad argument: modulus, length, range, null pointer */
#define PH_ERR_DEVICE 6            /* a HIP call failed */
#define PH_ERR_UNSUPPORTED 7       /* ring degree outside 2^1 .. 2^17 */
#define PH_ERR_NO_MEMORY 8         /* host or device allocation failed */
#define PH_ERR_CANCELLED 9         /* serving: the request was cancelled before it was launched */

/* Table creation flags.  By default canonical-output transforms over moduli
 * q < 2^50 run on exact binary64 FMA arithmetic (full-rate on gfx950) and all
 * others on 64-bit integer Shoup arithmetic built from 32-bit multiplies;
 * both give identical canonical results.  Lazy-output entry points always use
 * the integer butterflies, whose lazy values match the reference's.
 * PH_TABLE_INTEGER_ONLY forces the integer arithmetic everywhere. */
#define PH_TABLE_INTEGER_ONLY 1u

typedef struct ph_ntt_table ph_ntt_table;   /* UintNttTable<u64> */
typedef struct ph_dcrt_table ph_dcrt_table; /* UintCrtNttTable<u64> */
typedef struct ph_operand ph_operand;       /* device-resident Shoup-prepared NTT-domain operand */

/* ---- diagnostics ---- */
const char *ph_status_string(int status);
/* Text of the last failure on the calling thread ("" if none). */
const char *ph_last_error(void);
/* Debug builds only (no reference counterpart; the reference debug_assert!s
 * its value ranges, primus_ntt/src/ntt/hexl/mod.rs:452-465, 546-558, and
 * primus_reduce/src/slice_ops.rs:12-17).  The library variant built with
 * PH_RANGE_CHECK=1 (libprimus_hip_rc.so) checks every value at the transform
 * kernels' register-round boundaries against its arithmetic's range -- the
 * integer butterflies' [0, 4q) forward, [0, 2q) inverse and [0, q) canonical
 * outputs, and the wider ranges of the canonical-only arithmetics -- and counts
 * violations without changing any result.  report = {violations, first site
 * (1 forward input, 2 forward round, 3 inverse input, 4 inverse round, 5
 * inverse output), the first offending value's bits, its modulus}; clear != 0
 * zeroes the counters.  Synchronise the streams that ran the kernels first.
 * Default builds return PH_ERR_UNSUPPORTED. */
int ph_debug_range_report(uint64_t report[4], int clear);
/* Measurement builds only (no reference counterpart).  The library variant
 * built with PH_CLOCK_PROBE=1 (libprimus_hip_probe.so) logs every wave of the
 * transform kernels' single-pass chunk launches as it ends, 4 words per wave:
 * {shader-clock counter, 100 MHz counter, HW_ID register | XCC_ID register <<
 * 32, workgroup index}.  Each of PH_CLOCK_PROBE_SLOTS hardware wave slots
 * keeps its last PH_CLOCK_PROBE_DEPTH entries as a ring (wave c of the slot
 * at index c mod depth) and counts its waves.  entries (slots x depth x 4
 * words) and counts (slots) receive them when not null; clear != 0 restarts
 * the counts.  Successive entries of one slot over back-to-back launches give
 * the engine clock the kernel ran at: the ratio of the two counters' advances
 * times 100 MHz.  Synchronise first.  Default builds return
 * PH_ERR_UNSUPPORTED. */
#define PH_CLOCK_PROBE_SLOTS 65536
#define PH_CLOCK_PROBE_DEPTH 32
int ph_debug_clock_report(uint64_t *entries, uint32_t *counts, int clear);

/* ---- NttTable ---- */

/* NttTable::new -> UintNttTable::new (primus_ntt/src/ntt/primitive.rs:114-185):
 * psi = minimal primitive 2N-th root (primus_ntt/src/root.rs:89-111), Shoup
 * twiddle tables, N^-1.  `modulus` plays BarrettModulus::new (q in (1, 2^62)). */
int ph_ntt_table_new(uint32_t log_n, uint64_t modulus, ph_ntt_table **out);
/* ph_ntt_table_new with PH_TABLE_* flags (no reference counterpart: the
 * reference has one CPU arithmetic). */
int ph_ntt_table_new_ex(uint32_t log_n, uint64_t modulus, uint32_t flags, ph_ntt_table **out);
void ph_ntt_table_free(ph_ntt_table *table);
/* 1 if canonical transforms of this table run on binary64 FMA arithmetic, 0 if
 * on integer arithmetic, -1 for a null table. */
int ph_ntt_table_arith(const ph_ntt_table *table);

/* NttTable::poly_length (primitive.rs:188-190) */
size_t ph_ntt_table_poly_length(const ph_ntt_table *table);
/* UintNttTable::{log_n, modulus, root, inv_root} (primitive.rs:51-71) */
uint32_t ph_ntt_table_log_n(const ph_ntt_table *table);
uint64_t ph_ntt_table_modulus(const ph_ntt_table *table);
uint64_t ph_ntt_table_root(const ph_ntt_table *table);
uint64_t ph_ntt_table_inv_root(const ph_ntt_table *table);
/* UintNttTable::inv_n as ShoupFactor {value, quotient} (primitive.rs:81-83) */
int ph_ntt_table_inv_n(const ph_ntt_table *table, uint64_t *value, uint64_t *quotient);
/* UintNttTable::{root_powers (which=0, N), inv_root_powers (1, N),
 * ordinal_root_powers (2, 2N)} (primitive.rs:87-101); either output may be NULL. */
int ph_ntt_table_powers(const ph_ntt_table *table, int which, uint64_t *values, uint64_t *quotients);

/* NttTable::lazy_transform_slice: [0,4q) -> [0,4q) (primitive.rs:211-232) */
int ph_lazy_transform_slice(const ph_ntt_table *table, uint64_t *poly, size_t len);
/* NttTable::transform_slice: [0,4q) -> [0,q) (primitive.rs:234-244) */
int ph_transform_slice(const ph_ntt_table *table, uint64_t *poly, size_t len);
/* NttTable::lazy_inverse_transform_slice: [0,2q) -> [0,q) for this table (primitive.rs:246-286) */
int ph_lazy_inverse_transform_slice(const ph_ntt_table *table, uint64_t *values, size_t len);
/* NttTable::inverse_transform_slice: [0,2q) -> [0,q) (primitive.rs:288-296) */
int ph_inverse_transform_slice(const ph_ntt_table *table, uint64_t *values, size_t len);
/* HexlNttTable::compute_forward (primus_ntt/src/ntt/hexl/mod.rs:447-540), the
 * entry the HEXL table's NttTable methods call (hexl/mod.rs:321-338).  The
 * HEXL table computes the transform of UintNttTable on the same roots, and
 * the reference asserts equal outputs (primus_ntt/tests/hexl.rs), so this
 * runs the kernels above on a table made by ph_ntt_table_new.
 * input_mod_factor 1, 2 or 4: operand in [0, factor*q).  output_mod_factor 1:
 * [0,q), bit-identical to HEXL's; 4: a value in [0,4q) congruent to it (this
 * table's lazy value, not necessarily HEXL's own lazy bit pattern).  Other
 * factors are PH_ERR_INVALID (the reference debug_asserts them). */
int ph_hexl_compute_forward(const ph_ntt_table *table, uint64_t *operand, size_t len, uint64_t input_mod_factor,
                            uint64_t output_mod_factor);
/* HexlNttTable::compute_inverse (hexl/mod.rs:542-635): input_mod_factor 1 or 2,
 * output_mod_factor 1 ([0,q), bit-identical to HEXL's) or 2 (in [0,2q); this
 * table's lazy inverse is already in [0,q)). */
int ph_hexl_compute_inverse(const ph_ntt_table *table, uint64_t *operand, size_t len, uint64_t input_mod_factor,
                            uint64_t output_mod_factor);
/* NttTable::transform_monomial (primitive.rs:298-347) */
int ph_transform_monomial(const ph_ntt_table *table, uint64_t coeff, size_t degree, uint64_t *values,
                          size_t len);
/* NttTable::transform_coeff_one_monomial (primitive.rs:349-372) */
int ph_transform_coeff_one_monomial(const ph_ntt_table *table, size_t degree, uint64_t *values,
                                    size_t len);
/* NttTable::transform_coeff_minus_one_monomial (primitive.rs:374-397) */
int ph_transform_coeff_minus_one_monomial(const ph_ntt_table *table, size_t degree, uint64_t *values,
                                          size_t len);

/* ---- DcrtTable ---- */

/* DcrtTable::new -> UintCrtNttTable::new (primus_ntt/src/dcrt/primitive.rs:24-43) */
int ph_dcrt_table_new(uint32_t log_n, const uint64_t *moduli, uint32_t moduli_count, ph_dcrt_table **out);
int ph_dcrt_table_new_ex(uint32_t log_n, const uint64_t *moduli, uint32_t moduli_count, uint32_t flags,
                         ph_dcrt_table **out);
void ph_dcrt_table_free(ph_dcrt_table *table);
/* A DcrtTable over existing NttTables of one ring degree (the
 * UintCrtNttTable { ntt_tables } of primus_ntt/src/dcrt/primitive.rs:9-14
 * without rebuilding them).  The view does not own the tables, which must
 * outlive it. */
int ph_dcrt_table_view(const ph_ntt_table *const *tables, uint32_t moduli_count, ph_dcrt_table **out);
/* As ph_ntt_table_arith, for all limbs together. */
int ph_dcrt_table_arith(const ph_dcrt_table *table);
/* DcrtTable::{poly_length, moduli_count, crt_poly_length} (dcrt/primitive.rs:57-70) */
size_t ph_dcrt_poly_length(const ph_dcrt_table *table);
size_t ph_dcrt_moduli_count(const ph_dcrt_table *table);
size_t ph_dcrt_crt_poly_length(const ph_dcrt_table *table);
/* DcrtTable::ntt_tables()[i] (dcrt/primitive.rs:47-49); owned by the DCRT table */
const ph_ntt_table *ph_dcrt_ntt_table(const ph_dcrt_table *table, uint32_t i);
/* DcrtTable::{lazy_,}transform_slice, {lazy_,}inverse_transform_slice over
 * limb-major [L][N] host slices (dcrt/primitive.rs:107-141) */
int ph_dcrt_lazy_transform_slice(const ph_dcrt_table *table, uint64_t *poly, size_t len);
int ph_dcrt_transform_slice(const ph_dcrt_table *table, uint64_t *poly, size_t len);
int ph_dcrt_lazy_inverse_transform_slice(const ph_dcrt_table *table, uint64_t *poly, size_t len);
int ph_dcrt_inverse_transform_slice(const ph_dcrt_table *table, uint64_t *poly, size_t len);

/* ---- batched device forms ---- */

/* `count` limb-polynomials at d_polys; limb-polynomial i uses modulus i % L.
 * The slice semantics of transform_slice (lazy = 0) / lazy_transform_slice (lazy = 1)
 * and inverse_transform_slice, applied to every limb-polynomial in one launch. */
int ph_dcrt_transform_device(const ph_dcrt_table *table, uint64_t *d_polys, size_t count, int lazy,
                             void *stream);
int ph_dcrt_inverse_transform_device(const ph_dcrt_table *table, uint64_t *d_polys, size_t count,
                                     void *stream);
/* Single-modulus conveniences (L = 1). */
int ph_ntt_transform_device(const ph_ntt_table *table, uint64_t *d_polys, size_t count, int lazy,
                            void *stream);
int ph_ntt_inverse_transform_device(const ph_ntt_table *table, uint64_t *d_polys, size_t count,
                                    void *stream);

/* Device-resident NTT-domain operand, one canonical [0,q_i) row of N per limb
 * ([L][N] host array), held with its Shoup quotients as ShoupFactor values
 * (primus_factor/src/shoup_factor/mod.rs:44-55). */
int ph_operand_new(const ph_dcrt_table *table, const uint64_t *host_values, ph_operand **out);
void ph_operand_free(ph_operand *op);

/* Fused forward transform -> Shoup multiply by `op` -> inverse transform of each
 * limb-polynomial, in place, inputs in [0,4q), outputs in [0,q).  Equals
 * transform_slice, NttPolynomial::mul_assign (primus_poly/src/ntt/mul.rs:87-95),
 * inverse_transform_slice on every limb. */
int ph_dcrt_fwd_mul_inv_device(const ph_dcrt_table *table, uint64_t *d_polys, size_t count,
                               const ph_operand *op, void *stream);
/* Negacyclic product d_a <- d_a * d_b mod (X^N + 1, q_i) per limb-polynomial,
 * inputs in [0,q); the NTT route to Polynomial::naive_mul_inplace
 * (primus_poly/src/poly/mul.rs:111-142). */
int ph_dcrt_negacyclic_mul_device(const ph_dcrt_table *table, uint64_t *d_a, const uint64_t *d_b,
                                  size_t count, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ph_dcrt_negacyclic_mul_device with the scratch that
 * ring degrees N >= 2^14 need (count * N words) taken from a fixed device
 * array of PH_NOALLOC_SCRATCH_BYTES held by each GPU for the life of the
 * process, so no call allocates or frees anything.  Larger batches are
 * refused with PH_ERR_INVALID.  Every call on one GPU shares that GPU's
 * array: calls whose work overlaps in time on one GPU (other streams, other
 * threads) race and must be ordered by the caller. */
#define PH_NOALLOC_SCRATCH_BYTES (64u << 20)
int ph_dcrt_negacyclic_mul_device_noalloc(const ph_dcrt_table *table, uint64_t *d_a, const uint64_t *d_b,
                                          size_t count, void *stream);
/* The same product over limb-polynomials of uneven limb counts (no reference
 * counterpart: a batched form of the call above): row i of d_a and d_b is
 * reduced by the table's modulus d_limb_of[i] (< the table's limb count)
 * instead of i % L, so requests of one ring degree and any limb counts share
 * one launch.  d_limb_of is device memory, one byte per row; a row whose
 * index is L or more is left unchanged. */
int ph_dcrt_negacyclic_mul_rows_device(const ph_dcrt_table *table, uint64_t *d_a, const uint64_t *d_b,
                                       const uint8_t *d_limb_of, size_t count, void *stream);
/* x <- factors[i] * x (mod q_i) over `count` limb-polynomials (limb i = poly % L)
 * through MultiplyFactor (primus_factor/src/multiply.rs:4-76) at bit_shift 64
 * (ShoupFactor: NttPolynomial::mul_factor_assign, primus_poly/src/ntt/mul.rs:63-68),
 * 52 or 32 (HEXL's factor forms, hexl/mod.rs:231-271); lazy != 0 leaves
 * lazy_mul_modulo's [0, 2q) values, else mul_modulo's [0, q).  factors[i] <= q_i
 * (host memory); at shifts 32 and 52 every modulus and input must be below
 * 2^bit_shift (checked for the moduli). */
int ph_dcrt_mul_factor_device(const ph_dcrt_table *table, uint64_t *d_polys, size_t count, const uint64_t *factors,
                              uint32_t bit_shift, int lazy, void *stream);
/* NttPolynomial::mul_inplace (primus_poly/src/ntt/mul.rs:105-119): d_out = d_a * d_b. */
int ph_dcrt_pointwise_mul_device(const ph_dcrt_table *table, const uint64_t *d_a, const uint64_t *d_b,
                                 uint64_t *d_out, size_t count, void *stream);
/* NttPolynomial::add_mul_assign (primus_poly/src/ntt/mod.rs:97-108): d_acc += d_a * d_b. */
int ph_dcrt_pointwise_mul_add_device(const ph_dcrt_table *table, uint64_t *d_acc, const uint64_t *d_a,
                                     const uint64_t *d_b, size_t count, void *stream);

/* NTT-domain automorphism x -> x^degree (degree odd) of `count` limb-polynomials
 * in bit-reversed NTT order, out of place (d_in and d_out must not overlap):
 * dcrt_poly_ntt_auto_inplace (primus_fhe_core/src/automorphism/glwe/dcrt.rs:119-131)
 * with the permutation of generate_ntt_permutation (:43-61), computed per
 * element instead of read from a table.  The permutation is the same for
 * every limb; the table only fixes N. */
int ph_dcrt_automorphism_device(const ph_dcrt_table *table, const uint64_t *d_in, uint64_t *d_out, size_t count,
                                uint64_t degree, void *stream);

/* ---- RNS base conversion ---- */

typedef struct ph_base_converter ph_base_converter; /* BaseConverter<u64, BarrettModulus<u64>> */

/* BaseConverter::new (primus_rns/src/converter.rs:21-46) over RNSBase::new
 * (primus_rns/src/base.rs:50-90): input moduli pairwise coprime (else
 * PH_ERR_INVALID, RNSError::CoPrimeError), all moduli in (1, 2^62), at most 32
 * per base.  [(Q/q_i)^-1]_{q_i} and (Q/q_i) mod p_j stay on the device. */
int ph_base_converter_new(const uint64_t *ibase, uint32_t in_count, const uint64_t *obase, uint32_t out_count,
                          ph_base_converter **out);
void ph_base_converter_free(ph_base_converter *converter);
/* BaseConverter::fast_convert_array (converter.rs:148-177) on `count` CRT
 * polynomials of 2^log_n coefficients: d_in [count][in_count][N] ->
 * d_out [count][out_count][N], canonical.  Buffers must not overlap. */
int ph_fast_convert_array_device(const ph_base_converter *converter, const uint64_t *d_in, uint64_t *d_out,
                                 uint32_t log_n, size_t count, void *stream);
/* BaseConverter::exact_convert_array (converter.rs:222-316): one output modulus;
 * the binary64 rounding of sum_i y_i/q_i is the reference's, operation for
 * operation, so the centred result matches bit for bit. */
int ph_exact_convert_array_device(const ph_base_converter *converter, const uint64_t *d_in, uint64_t *d_out,
                                  uint32_t log_n, size_t count, void *stream);

/* ---- external product (RLWE x RGSW, k = 1, one modulus) ---- */

typedef struct ph_ext_key ph_ext_key; /* device-resident RGSW/GGSW key in NTT form */

/* The key of CrtGlwe::mul_dcrt_ggsw_inplace (primus_lattice/src/glwe/crt.rs:202-229)
 * with the gadget basis BigUintApproxSignedBasis::new(q, log_basis,
 * Some(decompose_length), rns) for a one-limb q
 * (primus_decompose/src/big_integer/basis.rs:34-195).  key_ntt holds the GGSW
 * (primus_lattice/src/ggsw/dcrt.rs:10-21) as [2 GLev][l GLWE][2 polys][N], NTT
 * form, canonical.  The table must outlive the key; 2^6 <= N <= 2^12. */
int ph_ext_key_new(const ph_ntt_table *table, uint32_t log_basis, uint32_t decompose_length,
                   const uint64_t *key_ntt, ph_ext_key **out);
void ph_ext_key_free(ph_ext_key *key);
/* log2 of the key's ring degree (0 for a null key). */
uint32_t ph_ext_key_log_n(const ph_ext_key *key);
/* `count` independent external products: d_in [count][2][N] RLWE ciphertexts
 * (a, b) in coefficient form, d_out [count][2][N].  Per product and input
 * polynomial: approximate signed decomposition into l digits
 * (primus_decompose/src/big_integer/common.rs:150-201, primus_rns/src/base.rs:201-236),
 * forward transform of each digit, multiply-accumulate against the key
 * (DcrtGlwe::add_dcrt_glev_mul_crt_poly_assign, primus_lattice/src/glwe/dcrt.rs:170-247);
 * then, unless ntt_output, the inverse transform of both accumulators.  One
 * fused launch. */
int ph_ext_product_device(const ph_ext_key *key, const uint64_t *d_in, uint64_t *d_out, size_t count,
                          int ntt_output, void *stream);
/* The key of ph_ext_key_new with a choice of digits.  PH_DIGITS_APPROX_SIGNED
 * is the reference's approximate signed decomposition, as ph_ext_key_new.
 * PH_DIGITS_UNSIGNED is the literal reading of "l unsigned digits of base
 * 2^log_basis": digit d of a coefficient v < q is
 * (v >> (drop_bits + d * log_basis)) & (2^log_basis - 1), with no rounding
 * carry, no split adjustment and no signed mapping.  It has no counterpart in
 * the reference, whose only external product uses the approximate signed
 * digits (primus_decompose/src/big_integer/common.rs:150-201). */
#define PH_DIGITS_APPROX_SIGNED 0
#define PH_DIGITS_UNSIGNED 1
int ph_ext_key_new_digits(const ph_ntt_table *table, uint32_t log_basis, uint32_t decompose_length, int digits,
                          const uint64_t *key_ntt, ph_ext_key **out);
/* The same on host memory, synchronous. */
int ph_ext_product(const ph_ext_key *key, const uint64_t *in, uint64_t *out, size_t count, int ntt_output);

/* The external product over several limbs: CrtGlwe::mul_dcrt_ggsw_inplace
 * (primus_lattice/src/glwe/crt.rs:201-229) with k = 1.  The key is a DCRT GGSW
 * [2][l][2][L][N] in NTT form over the table's L limbs (1 <= L <= 4), values
 * below their limb's modulus, with the gadget basis
 * BigUintApproxSignedBasis::new(Q, log_basis, Some(l)) over Q = q_0 ... q_{L-1}
 * (primus_decompose/src/big_integer/basis.rs:34-195).  The table must outlive
 * the key. */
typedef struct ph_ext_key_rns ph_ext_key_rns;
int ph_ext_key_rns_new(const ph_dcrt_table *table, uint32_t log_basis, uint32_t decompose_length,
                       const uint64_t *key_ntt, ph_ext_key_rns **out);
void ph_ext_key_rns_free(ph_ext_key_rns *key);
/* d_in [count][2][L][N] residues, each below its limb's modulus (as the
 * reference's DCRT ciphertexts hold them), d_out [count][2][L][N]: per input polynomial
 * each coefficient's residues composed into one integer below Q
 * (primus_rns/src/base.rs:396-450), its gadget digits, each digit wrapped
 * into every limb (base.rs:201-236), transformed, and multiply-accumulated
 * with the key (glwe/dcrt.rs:170-247); then, unless ntt_output, the inverse
 * transform of the accumulators.  For 2^6 <= N <= 2^12 with the digits
 * inside one 64-bit window of the composed value (l * log_basis, plus one
 * bit when bits are dropped, at most 64) this is one fused launch per batch
 * that allocates nothing and can be captured in a HIP graph; otherwise a digit
 * pass, the batched transforms and a multiply-accumulate pass, with scratch
 * by hipMallocAsync on the stream. */
int ph_ext_product_rns_device(const ph_ext_key_rns *key, const uint64_t *d_in, uint64_t *d_out, size_t count,
                              int ntt_output, void *stream);

/* ---- RLWE encryption, uneven batches ---- */

typedef struct ph_rlwe_ctx ph_rlwe_ctx; /* device-resident tables for a grid of (ring degree, limb) */

/* One request of an uneven batch: ring degree 2^log_n, `limbs` RNS limbs.
 * a_offset: word offset of its [limbs][N] a_hat block in the a buffer (the
 * b_hat block goes to the same offset of the output); small_offset: element
 * offset of its N secret / error / message coefficients. */
typedef struct {
    uint32_t log_n;
    uint32_t limbs;
    uint64_t a_offset;
    uint64_t small_offset;
} ph_rlwe_request;

/* Tables for ring degrees 2^min_log_n .. 2^max_log_n (within 2^10 .. 2^16),
 * limb i of degree 2^l using moduli[(l - min_log_n) * max_limbs + i]. */
int ph_rlwe_ctx_new(uint32_t min_log_n, uint32_t max_log_n, uint32_t max_limbs, const uint64_t *moduli,
                    ph_rlwe_ctx **out);
void ph_rlwe_ctx_free(ph_rlwe_ctx *ctx);
uint64_t ph_rlwe_ctx_modulus(const ph_rlwe_ctx *ctx, uint32_t log_n, uint32_t limb);
/* NttRlweSecretKey::encrypt_inplace_with_embedding, unsigned embedding, with the
 * samples supplied (primus_fhe_core/src/secret_key/rlwe.rs:227-268): per request
 * and limb, b_hat = NTT(e + Delta*m) + a_hat (.) NTT(s) mod q_i, with
 * Delta = floor(Q/256) and Q the product of the request's limb moduli.  s, e
 * are int8 (ternary secret, rounded Gaussian error), m uint8; a_hat is
 * uniform in NTT form.  All requests run in one persistent launch. */
int ph_rlwe_encrypt_device(const ph_rlwe_ctx *ctx, const ph_rlwe_request *requests, uint32_t n_requests,
                           const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e, const uint8_t *d_m,
                           uint64_t *d_out, void *stream);

/* A batch layout prepared once (no reference counterpart: launch planning):
 * the per-limb work items with their Delta, sorted, uploaded, and the
 * launch's scratch allocated.  ph_rlwe_encrypt_plan_device then only
 * enqueues the persistent kernel -- no allocation, copy or synchronisation,
 * so it can be captured in a HIP graph -- for any buffers laid out as the
 * plan's requests.  The context must outlive the plan. */
typedef struct ph_rlwe_plan ph_rlwe_plan;
int ph_rlwe_plan_new(const ph_rlwe_ctx *ctx, const ph_rlwe_request *requests, uint32_t n_requests,
                     ph_rlwe_plan **out);
void ph_rlwe_plan_free(ph_rlwe_plan *plan);
int ph_rlwe_encrypt_plan_device(const ph_rlwe_plan *plan, const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e,
                                const uint8_t *d_m, uint64_t *d_out, void *stream);
/* Times persistent grids of 1..4 workgroups per compute unit on the given
 * buffers (outputs are overwritten) and keeps the fastest in the plan. */
int ph_rlwe_plan_autotune(ph_rlwe_plan *plan, const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e,
                          const uint8_t *d_m, uint64_t *d_out, int reps, uint32_t *grid, float *best_ms);
/* Workgroups of the plan's launch. */
uint32_t ph_rlwe_plan_grid(const ph_rlwe_plan *plan);
/* Draws the encryption inputs on the device, laid out as for
 * ph_rlwe_encrypt_device: s ternary uniform in {-1, 0, 1}, e the rounded
 * discrete Gaussian of sigma 3.2 cut at |e| <= 19, m uniform in [0, 256) and
 * a_hat uniform mod q_i.  Replaces the host sampling of
 * NttRlweSecretKey::encrypt_inplace_with_embedding
 * (primus_fhe_core/src/secret_key/rlwe.rs:227-268, samplers in
 * primus_distr/src/common.rs:56-121 and primus_distr/src/discrete_gaussian/mod.rs)
 * with the distributions of the batch workload.  Every value is a function of
 * (seed, request index, what is drawn, element) through Philox4x32-10, so the
 * result is the same however the launch is arranged. */
int ph_rlwe_sample_device(const ph_rlwe_ctx *ctx, const ph_rlwe_request *requests, uint32_t n_requests,
                          uint64_t seed, uint64_t *d_a, int8_t *d_s, int8_t *d_e, uint8_t *d_m, void *stream);
/* The same for a plan's requests: one kernel enqueue (capturable, like
 * ph_rlwe_encrypt_plan_device). */
int ph_rlwe_sample_plan_device(const ph_rlwe_plan *plan, uint64_t seed, uint64_t *d_a, int8_t *d_s, int8_t *d_e,
                               uint8_t *d_m, void *stream);

/* ---- serving: continuous batching (no reference counterpart; the reference is a
 * single-threaded library, this is the request loop around the batched forms) ---- */

#define PH_REQ_POLY_MUL 0     /* negacyclic product, [limbs][N] in a and b */
#define PH_REQ_RLWE_ENCRYPT 1 /* as ph_rlwe_encrypt_device, one request */
#define PH_REQ_EXT_PRODUCT 2  /* as ph_ext_product_device with the server's key */

typedef struct {
    int kind;
    uint32_t log_n;
    uint32_t limbs;      /* poly-multiply / RLWE limbs; ignored for external products */
    const uint64_t *in0; /* poly-mul: a [limbs][N]; RLWE: a_hat [limbs][N]; ext: (a, b) [2][N] */
    const uint64_t *in1; /* poly-mul: b [limbs][N] */
    const int8_t *s;     /* RLWE: ternary secret [N] */
    const int8_t *e;     /* RLWE: error [N] */
    const uint8_t *m;    /* RLWE: message [N] */
    uint64_t *out;       /* poly-mul, RLWE: [limbs][N]; ext: [2][N] */
} ph_serve_request;

typedef struct {
    uint64_t hbm_budget_bytes; /* admission limit on the device memory of in-flight requests */
    uint32_t max_batch;        /* requests per batch; 1 = the plain one-launch-per-request path */
    uint32_t streams;          /* batches in flight at once: one stream each, plus one stream
                                * for all input copies and one for all result copies.  Each
                                * wants a hardware queue of its own (HIP's
                                * GPU_MAX_HW_QUEUES >= streams + 2, set before HIP starts):
                                * streams sharing a queue wait behind each other's event
                                * waits. */
    uint32_t min_log_n, max_log_n, max_limbs;
    const uint64_t *moduli;    /* [(max_log_n - min_log_n + 1)][max_limbs] limb moduli per ring degree */
    const ph_ext_key *ext_key; /* key of the external-product requests (NULL: none accepted) */
    uint32_t per_poly_launches; /* 1: one launch per limb-polynomial of a poly-multiply; with
                                 * max_batch 1 and one stream this is the plain path */
} ph_server_config;

typedef struct {
    uint64_t completed, batches, launches, peak_inflight_bytes, cancelled;
    /* host seconds, summed: inside ph_server_submit; the dispatcher enqueueing
     * batches, and waiting for a free slot; the completion thread waiting for
     * batches to finish on the device */
    double submit_s, dispatch_s, slot_wait_s, device_wait_s;
} ph_server_stats;

typedef struct ph_server ph_server;

/* Builds the resident tables and starts the dispatcher and completion threads. */
int ph_server_new(const ph_server_config *config, ph_server **out);
/* Stops taking requests, launches every pending one, waits until all
 * in-flight batches have completed and their results are in the callers'
 * buffers, then frees the server. */
void ph_server_free(ph_server *server);
/* Queues a request; host buffers must stay valid until it completes.  Rejects
 * malformed requests (PH_ERR_INVALID) and ones larger than the budget
 * (PH_ERR_NO_MEMORY). */
int ph_server_submit(ph_server *server, const ph_serve_request *request, uint64_t *ticket);
/* Blocks until the ticket completes; *status receives its PH_* result. */
int ph_server_wait(ph_server *server, uint64_t ticket, int *status);
/* Seconds from submission to completion of a finished ticket. */
int ph_server_latency(ph_server *server, uint64_t ticket, double *seconds);
int ph_server_drain(ph_server *server);
/* Cancels a request that is still queued: its ticket completes at once with
 * PH_ERR_CANCELLED and its buffers are never touched.  PH_ERR_INVALID if it
 * is already in a batch or finished. */
int ph_server_cancel(ph_server *server, uint64_t ticket);
/* paused != 0: the dispatcher forms no new batch (submissions queue, batches
 * already launched complete); 0 resumes.  Freeing the server resumes it. */
int ph_server_pause(ph_server *server, int paused);
/* Failure injection for tests of error handling: the next `batches` batches
 * are retired with PH_ERR_DEVICE without launching anything, as if a HIP call
 * of theirs had failed; later batches run normally. */
int ph_server_inject_fault(ph_server *server, uint32_t batches);
int ph_server_stats_get(ph_server *server, ph_server_stats *out);
/* Open-loop load: submits requests[i] at Poisson arrival times of the given
 * rate, waits for all, reports per-request latency and the elapsed time. */
int ph_server_run_poisson(ph_server *server, const ph_serve_request *requests, size_t n, double rate_per_s,
                          uint64_t seed, double *latency_s, double *elapsed_s);

/* ---- autotuning (no reference counterpart: launch-shape selection for the GPU) ---- */

/* Operation codes for the tuning table. */
#define PH_OP_FORWARD_LAZY 0
#define PH_OP_FORWARD 1
#define PH_OP_INVERSE 2
#define PH_OP_FWD_MUL_INV 3
#define PH_OP_NEGACYCLIC_MUL 4

/* Process-wide tuning table: for (log_n, op, arith: 0 integer / 1 binary64) the
 * launch shape of the transform kernels, as bit fields (0 = the default shape):
 *   - log2 of the coefficients each thread holds, 3 or 4 (radix-8 or radix-16
 *     register rounds; 0: 4);
 *   - PH_SHAPE_STREAM: the persistent streaming kernel (forward-first
 *     operations, 2^10 <= N <= 2^13), with PH_SHAPE_LPB(k), k = 0..3, the
 *     limb-polynomials each workgroup walks: k = 0 sizes the grid to one
 *     resident generation of workgroups, k > 0 to ceil(count / 2^k) of them;
 *   - PH_SHAPE_TILE(c), c = 11..13 (N >= 2^14, N / 2^c <= 32): the two-pass
 *     schedule's LDS tile, chunks of 2^c coefficients after columns of
 *     N / 2^c (0: 4096).
 * ph_tuning_set rejects a shape whose fields do not apply at that degree.
 * Results never depend on the shape. */
#define PH_SHAPE_STREAM 16
#define PH_SHAPE_LPB(k) ((k) << 5)
#define PH_SHAPE_TILE(c) ((c) << 8)
int ph_tuning_set(uint32_t log_n, int op, int arith, int shape);
int ph_tuning_get(uint32_t log_n, int op, int arith);
/* The shapes ph_autotune tries for (log_n, op): writes up to `max` of them to
 * shapes and returns how many there are (negative status on a bad query). */
int ph_tuning_shapes(uint32_t log_n, int op, int *shapes, int max);
/* Save / load the table as text lines "log_n op arith shape".  The file the
 * environment variable PH_TUNING_FILE names, if it exists, is loaded once
 * when the process creates its first table; an invalid entry in it makes
 * table creation fail with PH_ERR_INVALID. */
int ph_tuning_save(const char *path);
int ph_tuning_load(const char *path);
/* Times every launch shape of `op` for this table (ph_tuning_shapes) on
 * `count` limb-polynomials (reps timed launches each), stores the fastest in
 * the tuning table and reports it. */
int ph_autotune(const ph_dcrt_table *table, int op, size_t count, int reps, int *chosen, float *best_ms);

/* ---- device memory helpers (for callers without their own HIP allocator) ---- */
int ph_device_alloc(size_t bytes, void **d_ptr);
int ph_device_free(void *d_ptr);
int ph_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
int ph_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);
int ph_stream_synchronize(void *stream);
/* Page-locked host memory (hipHostMalloc).  Serving requests whose buffers
 * all come from here are copied to and from the device directly, without
 * the server's staging copies. */
int ph_host_alloc(size_t bytes, void **h_ptr);
int ph_host_free(void *h_ptr);

#ifdef __cplusplus
}
#endif
#endif /* PRIMUS_HIP_H */
