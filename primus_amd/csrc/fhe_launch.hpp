// fhe_launch.hpp -- host-side view of the fused external-product and
// RLWE-encryption kernels (fhe_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntt_device.hpp"
#include "sample.hpp"

namespace ph {

constexpr uint32_t kExtMinLogN = 6;
// Coefficients per thread of the external-product kernel (2^kExtLogE); the
// key rows are stored permuted for this register layout.
constexpr uint32_t kExtLogE = 3;
constexpr uint32_t kExtMaxLogN = 12;

// External product arguments: gadget constants of
// BigUintApproxSignedBasis::new(q, log_basis, Some(l)) for a one-limb q, the
// resident key and the batch.
struct ExtArgs {
    const uint64_t *in;     // [count][2][N] coefficients < q
    uint64_t *out;          // [count][2][N]
    const Shoup *key;       // [2][l][2][N] NTT form, permuted for coalescing (integer path)
    const ShoupF *key_f;    // the same, centred binary64 form (q < 2^50)
    const DevTable *tab;
    uint64_t split;         // values >= split are shifted by pow2_sub_q (has_split)
    uint64_t pow2_sub_q;    // 2^bits(q) - q
    uint64_t basis_minus_one, carry_mask, half, q_minus_basis;
    uint32_t drop_bits, log_basis, l;
    uint32_t has_split, basis_is_two, ntt_output;
    // binary64 keys: the accumulators are reduced after this many key
    // products (ext_mac_reduce_interval); integer keys ignore it
    uint32_t mac_reduce_every;
    // 1: the literal unsigned digits (v >> (drop_bits + d*log_basis)) & (B - 1)
    // of each coefficient, no carry, no split, no signed mapping
    // (PH_DIGITS_UNSIGNED; has_split is 0 then)
    uint32_t unsigned_digits;
};

hipError_t launch_ext_product(const ExtArgs &g, uint32_t log_n, bool f64, uint64_t count, hipStream_t s);

// Key products a binary64 accumulator takes between reductions.  Each product
// is an exact integer below 0.7503 q in magnitude (a forward output below
// 3.2 q times a centred key entry, F64Arith::mulmod), a reduced accumulator is
// at most q/2 + 1, and the sum stays exact while below 2^53; 0.76 q per product
// and q for the reduced start leave a margin.  At least 9 for q < 2^50, so an
// external product with 2 l <= 9 terms never reduces before its end.
inline uint32_t ext_mac_reduce_interval(uint64_t q) {
    const double r = (9007199254740992.0 - (double)q) / (0.76 * (double)q);
    return r >= 1048576.0 ? 1048576u : (uint32_t)r;
}

// External product over several limbs (CrtGlwe::mul_dcrt_ggsw_inplace): the
// gadget basis is over Q = q_0 ... q_{L-1}, at most kExtRnsMaxLimbs limbs, so
// Q and every constant below fit kExtRnsWords little-endian words.
constexpr uint32_t kExtRnsMaxLimbs = 4;
constexpr uint32_t kExtRnsWords = 4;

// The digit pass: each coefficient's L residues composed into one integer
// below Q, its gadget digits, and each digit's residue per limb.
struct ExtRnsArgs {
    const uint64_t *in;  // [count][2][L][N] residues
    uint64_t *digits;    // [count][2][l][L][N] digit residues
    uint64_t q[kExtRnsMaxLimbs];
    uint64_t inv[kExtRnsMaxLimbs], inv_q[kExtRnsMaxLimbs];  // (Q/q_i)^-1 mod q_i as Shoup pairs
    uint64_t punct[kExtRnsMaxLimbs][kExtRnsWords];          // Q/q_i
    uint64_t Q[kExtRnsWords], split[kExtRnsWords], pow2_sub[kExtRnsWords];
    uint64_t basis, basis_minus_one, carry_mask, half;
    uint64_t count;
    uint32_t L, l, log_basis, drop_bits, has_split, log_n;
    uint32_t punct_words;  // the most nonzero words of any Q/q_i
    uint32_t mac_reduce_every;  // as ExtArgs (the fused kernel's binary64 accumulators)
};

// The multiply-accumulate of the transformed digits with the key.
struct ExtRnsMacArgs {
    const uint64_t *digits;  // [count][2][l][L][N], NTT form
    const Shoup *key;        // [2][l][2][L][N], NTT form, canonical, with Shoup quotients
    uint64_t *out;           // [count][2][L][N]
    const DevTable *tabs;    // L tables
    uint64_t count;
    uint32_t L, l, log_n;
};

hipError_t launch_ext_rns_digits(const ExtRnsArgs &a, hipStream_t s);
hipError_t launch_ext_rns_mac(const ExtRnsMacArgs &a, hipStream_t s);

// The fused form: one workgroup per (product, output limb) composes each
// coefficient's residues, takes its digits, transforms each digit's residues
// for its limb, multiply-accumulates them with the key and inverse-transforms
// the two accumulators, with no intermediate in device memory.  Ring degrees
// kExtMinLogN .. kExtMaxLogN; E = 2^kExtLogE coefficients per thread.
//
// Its composition is Garner's mixed-radix form of the CRT value, which for
// residues r_i < q_i is the same integer below Q that compose_inplace
// (primus_rns/src/base.rs:396-420) sums:
//   x = v_0 + v_1 q_0 + v_2 q_0 q_1 + ...,  v_0 = r_0,
//   v_i = (r_i - (v_0 + v_1 q_0 + ... + v_{i-1} q_0..q_{i-2})) (q_0..q_{i-1})^-1 mod q_i,
// each v_i below q_i, so x < Q with no reduction by Q.
struct ExtRnsGarner {
    uint64_t c[kExtRnsMaxLimbs], c_q[kExtRnsMaxLimbs];  // (q_0 ... q_{i-1})^-1 mod q_i, Shoup pair (i >= 1)
    // q_i ceil(2^63 / q_i): added before the partial value (below 2^63) is
    // subtracted, so the Shoup operand r_i + off_i - s stays a 64-bit value
    // congruent to r_i - s
    uint64_t off[kExtRnsMaxLimbs];
    // q_j mod q_i with its Shoup quotient, for the partial value's Horner
    // steps mod q_i (j <= i - 2)
    uint64_t qm[kExtRnsMaxLimbs][kExtRnsMaxLimbs], qm_q[kExtRnsMaxLimbs][kExtRnsMaxLimbs];
};

struct ExtRnsFusedArgs {
    ExtRnsArgs g;              // gadget constants and the input (g.digits unused)
    ExtRnsGarner garner;
    uint64_t *out;             // [count][2][L][N]
    const Shoup *key;          // [2][l][2][L] rows of N, slot t*E + j at j*(N/E) + t (integer arithmetic)
    const ShoupF *key_f;       // the same as centred binary64 pairs (every limb below 2^50)
    const DevTable *tabs;      // L tables
    uint32_t ntt_output;
    uint32_t pad_;
};

// arith: kArithF64 / kArithCanonInt bits (ntt_launch.hpp) of the limbs; the
// kernel is instantiated per limb count a.g.L (1..kExtRnsMaxLimbs), Q held in
// L words.  Input residues must lie below their limb's modulus.
hipError_t launch_ext_rns_fused(const ExtRnsFusedArgs &a, int arith, hipStream_t s);

// RLWE encryption: ring degrees 2^10 .. 2^16 in one persistent launch.
constexpr uint32_t kRlweMinLogN = 10;
constexpr uint32_t kRlweMaxLogN = 16;
constexpr uint32_t kRlweThreads = 256;

// One (request, limb) work item.
struct RlweItem {
    uint64_t row;        // word offset of this limb's row in a_hat / b_hat
    uint64_t small;      // element offset of the request's s, e, m
    uint64_t delta;      // Delta mod q_i
    uint64_t delta_q;    // its Shoup quotient
    uint32_t log_n;
    uint32_t tab;        // index into the table array
};

struct RlweArgs {
    const RlweItem *items;  // sorted by decreasing ring degree
    uint32_t n_items;       // work items; the list holds one more, the work queue
    uint32_t pad_;
    const DevTable *tabs;
    const uint64_t *a;      // a_hat rows (NTT form)
    uint64_t *out;          // b_hat rows
    const int8_t *s;        // ternary secret
    const int8_t *e;        // error
    const uint8_t *m;       // message
    uint64_t *scratch;      // scratch_words per workgroup (ring degrees above 2^12)
    uint64_t scratch_words; // the largest ring degree of the launch
    uint32_t canon_int;     // every modulus below 2^60: IntCanonArith kernel, else IntArith
    uint32_t pad2_;
    // Work queue: {next item, workgroups done}, the item list's trailing
    // element (zero when uploaded; the launch's last workgroup zeroes it again,
    // so launches of one list run one after the other, never at once).
    uint32_t *queue;
};

hipError_t launch_rlwe_encrypt(const RlweArgs &a, uint32_t grid, hipStream_t s);

// Sampling of RLWE encryption inputs (sample.hpp): one workgroup per chunk of
// up to kSampleChunk consecutive elements of one request's row for one
// purpose.  dst: element offset in s / e / m, or word offset in a_hat.
constexpr uint32_t kSampleChunk = 4096;
struct SampleChunk {
    uint64_t dst;
    uint64_t q;          // the row's modulus (a_hat rows)
    uint32_t request;    // counter word 1
    uint32_t purpose;    // counter word 2: kSampleSecret, kSampleError, kSampleMessage, kSampleA + limb
    uint32_t k0, count;  // first element (counter word 0) and element count
};

struct SampleArgs {
    const SampleChunk *chunks;
    uint32_t n_chunks;
    uint32_t seed_lo, seed_hi;  // Philox key
    const uint64_t *gauss;      // kGaussThresholds cumulative thresholds
    uint64_t *a;
    int8_t *s;
    int8_t *e;
    uint8_t *m;
};

hipError_t launch_rlwe_sample(const SampleArgs &a, hipStream_t st);

}  // namespace ph
