// arith.hpp -- 64-bit modular arithmetic for gfx950 (device) and for table
// set-up (host).
//
// All arithmetic is exact unsigned integer work.  The device forms compile to
// 32-bit VALU chains (v_mul_hi_u32 / v_mad_u64_u32 / v_mul_lo_u32): CDNA has
// no 64x64 multiplier.  Semantics follow the reference's reduction contract:
//   * Shoup lazy multiply: primus_factor/src/shoup_factor/mod.rs:108-113
//     (result in [0, 2q) for any 64-bit b), full: :121-124.
//   * Barrett 2-limb reduction of a 128-bit product:
//     primus_modulus/src/barrett/ops.rs:42-74, reduce_mul :310-316,
//     reduce_mul_add :331-337.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host_tables.hpp"  // u128, Shoup, ShoupF and the host arithmetic

namespace ph {

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// floor(a*b / 2^64) - e with e in {0, 1, 2}: a1*b1 plus the high halves of
// the two middle products, without a0*b0 and without the low halves of the
// middle products (together below 3 * 2^64, so the floor loses at most 2).
// Two multiply-highs and one multiply-add instead of four multiplies and
// their carries.  The sum stays below 2^64: (2^32-1)^2 + 2 (2^32-1) < 2^64.
__host__ __device__ __forceinline__ uint64_t mulhi64_approx(uint64_t a, uint64_t b) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint32_t h1 = (uint32_t)(((uint64_t)a1 * b0) >> 32), h2 = (uint32_t)(((uint64_t)a0 * b1) >> 32);
    return (uint64_t)a1 * b1 + h1 + h2;
}

// x >= m ? x - m : x.  The subtraction's own borrow selects (sub, sub-with-
// borrow, two selects) where a 64-bit comparison would add a fifth
// instruction.
__host__ __device__ __forceinline__ uint64_t sub_if_ge(uint64_t x, uint64_t m) {
    unsigned c0, c1;
    const unsigned d0 = __builtin_subc((unsigned)x, (unsigned)m, 0u, &c0);
    const unsigned d1 = __builtin_subc((unsigned)(x >> 32), (unsigned)(m >> 32), c0, &c1);
    return c1 ? x : (((uint64_t)d1 << 32) | d0);
}

// w*b - k*q (mod 2^64), given nq = 2^64 - q: w*b + k*nq, whose two low
// products chain through one 64-bit multiply-add (w0*b0, then k0*nq0 added)
// and whose four cross products only reach the high word.  Two multiply-adds,
// four multiply-lows and two adds, against two full 64-bit products and a
// 64-bit subtraction.
__host__ __device__ __forceinline__ uint64_t mul_sub_lo(uint64_t w, uint64_t b, uint64_t k, uint64_t nq) {
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint32_t k0 = (uint32_t)k, k1 = (uint32_t)(k >> 32), n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
    const uint32_t cross = w1 * b0 + w0 * b1 + k1 * n0 + k0 * n1;
    const uint64_t low = (uint64_t)k0 * n0 + (uint64_t)w0 * b0;
    return low + ((uint64_t)cross << 32);
}

// x - k*q (mod 2^64), given nq = 2^64 - q: the low product k0*nq0 as a
// multiply-add onto x, the two cross products into the high word.
__host__ __device__ __forceinline__ uint64_t sub_mul_lo(uint64_t x, uint64_t k, uint64_t nq) {
    const uint32_t k0 = (uint32_t)k, k1 = (uint32_t)(k >> 32), n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
    return (uint64_t)k0 * n0 + x + ((uint64_t)(k1 * n0 + k0 * n1) << 32);
}

// w*b - floor(wq*b / 2^64)*q (mod 2^64), in [0, 2q); wq = floor(w*2^64/q).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t w, uint64_t wq, uint64_t b, uint64_t q) {
    return mul_sub_lo(w, b, mulhi64(wq, b), 0 - q);
}

__device__ __forceinline__ uint64_t shoup_full(uint64_t w, uint64_t wq, uint64_t b, uint64_t q) {
    return sub_if_ge(shoup_lazy(w, wq, b, q), q);
}

// shoup_lazy with the approximate quotient: the quotient is at most two below
// Shoup's, so the result is congruent to w*b and lies in [0, 4q) for any
// 64-bit b (q < 2^62).  Used only where no caller sees lazy values.
__host__ __device__ __forceinline__ uint64_t shoup_approx(uint64_t w, uint64_t wq, uint64_t b, uint64_t q) {
    return mul_sub_lo(w, b, mulhi64_approx(wq, b), 0 - q);
}

// Canonical (a*b + c) mod q by the 2-limb Barrett quotient with ratio
// floor(2^128/q) = (r1:r0).  Inputs a, b < 2^64 with a*b + c < 2^128.
__device__ __forceinline__ uint64_t barrett_reduce128(uint64_t lo, uint64_t hi, uint64_t q,
                                                      uint64_t r0, uint64_t r1) {
    uint64_t ah = mulhi64(lo, r0);
    uint64_t b_lo = lo * r1;
    uint64_t b_hi = mulhi64(lo, r1);
    b_lo += ah;
    b_hi += (b_lo < ah);
    uint64_t c_lo = hi * r0;
    uint64_t c_hi = mulhi64(hi, r0);
    uint64_t d = hi * r1;
    uint64_t s = b_lo + c_lo;
    uint64_t bch = b_hi + c_hi + (s < b_lo);
    uint64_t qq = d + bch;
    return sub_if_ge(lo - qq * q, q);
}

// The same quotient for any (hi:lo) < q * 2^64 (hi < q): the estimate is then
// at most three below the true quotient (one from the ratio's truncation,
// two from the dropped low partial products), so up to three subtractions
// make the remainder canonical.
__device__ __forceinline__ uint64_t barrett_reduce128_wide(uint64_t lo, uint64_t hi, uint64_t q, uint64_t r0,
                                                           uint64_t r1) {
    uint64_t ah = mulhi64(lo, r0);
    uint64_t b_lo = lo * r1;
    uint64_t b_hi = mulhi64(lo, r1);
    b_lo += ah;
    b_hi += (b_lo < ah);
    uint64_t c_lo = hi * r0;
    uint64_t c_hi = mulhi64(hi, r0);
    uint64_t d = hi * r1;
    uint64_t s = b_lo + c_lo;
    uint64_t bch = b_hi + c_hi + (s < b_lo);
    uint64_t qq = d + bch;
    uint64_t r = lo - qq * q;
    r = sub_if_ge(r, q);
    r = sub_if_ge(r, q);
    return sub_if_ge(r, q);
}

__device__ __forceinline__ uint64_t barrett_mul(uint64_t a, uint64_t b, uint64_t q, uint64_t r0,
                                                uint64_t r1) {
    return barrett_reduce128(a * b, mulhi64(a, b), q, r0, r1);
}

__device__ __forceinline__ uint64_t barrett_mul_add(uint64_t a, uint64_t b, uint64_t c, uint64_t q,
                                                    uint64_t r0, uint64_t r1) {
    uint64_t lo = a * b;
    uint64_t hi = mulhi64(a, b);
    lo += c;
    hi += (lo < c);
    return barrett_reduce128(lo, hi, q, r0, r1);
}

// (a + b) mod q for a, b < q.
__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) {
    return sub_if_ge(a + b, q);
}

}  // namespace ph
