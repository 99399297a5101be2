// ntt_device.hpp -- register/LDS building blocks of the negacyclic NTT on gfx950.
//
// The transform is the reference's Harvey-lazy Cooley-Tukey forward
// (normal -> bit-reversed, primus_ntt/src/ntt/primitive.rs:211-232) and
// Gentleman-Sande inverse (bit-reversed -> normal, :246-286), butterfly for
// butterfly: every butterfly pairs the same two slots with the same twiddle.
//
// Only the schedule differs.  A thread holds a group of 2^R slots in VGPRs and
// runs R consecutive stages on them (radix-2^R); groups are re-formed between
// register rounds through LDS.  Slot numbering inside a round:
//
//   forward round, global stages [S, S+R):  k = (c << (logN-S)) + (j << (logN-S-R)) + o
//   inverse round, global stages [P, P+R):  k = (c << (P+R)) + (j << P) + o
//
// with j < 2^R the register index, c the chunk the group lives in at the
// round's first (forward) or last (inverse) stage and o the offset inside it.
// The twiddle of a butterfly is the one the CPU loop would fetch:
//   forward stage s, chunk c_s:  root_powers[2^s + c_s]
//   inverse stage p, chunk c_p:  inv_root_powers[1 + N - (N >> p) + c_p]
//   (the inverse's last stage, gap N/2, folds N^-1 as in primitive.rs:270-285).
//
// Two arithmetic policies run the same schedule:
//
// * IntArith: uint64 lanes, the reference's Shoup butterflies exactly
//   (lazy [0,4q) forward, [0,2q) inverse); 32-bit multiply chains; any
//   q < 2^62.  Lazy outputs are bit-identical to the CPU table.
// * F64Arith: for q < 2^50, coefficients are exact integers held in binary64
//   lanes in a signed range, and a*w mod q is computed with full-rate FMAs:
//       h = x*w;  l = fma(x, w, -h)            (x*w = h + l exactly)
//       k = fma(x, w/q, C) - C                 (C = 1.5*2^52: k = round(x*(w/q)))
//       r = fma(-k, q, h) + l                  (both steps exact)
//   For |x| <= 2^51 and |w/q - wp| <= 2^-54 + 2^-64, |x*w/q - k| <= 0.6251,
//   so r = x*w - k*q with |r| <= 0.6251 q; every intermediate is an integer of
//   magnitude < 2^53 and is therefore exact.  Values entering a product are
//   kept below 2^51 by centred reductions (a - round(a/q)*q, |.| <= q/2 + 1).
//   The arithmetic is exact, so canonical outputs equal the CPU table's;
//   intermediate (lazy) values differ, so lazy entry points use IntArith.
#pragma once

#include "arith.hpp"

namespace ph {

// A launch's grid in one dimension is a 32-bit count of work-items (the
// dispatch packet's grid size): launchers refuse larger batches with an error
// instead of truncating the workgroup count.
inline bool grid_fits(uint64_t groups, uint64_t threads_per_group) {
    return groups <= 0xffffffffull / threads_per_group;
}

// Pointer into the constant address space (read-only global memory).  Table
// pointers read out of a DevTable are generic otherwise, and generic loads are
// FLAT instructions that count against both memory counters (every LDS wait
// then also waits for them) and never become scalar loads for wave-uniform
// addresses.  Twiddle tables and operands are immutable while a kernel runs.
template <class T>
using gptr = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T *p) {
    return (gptr<T>)p;
}

__device__ __forceinline__ Shoup ld(gptr<Shoup> p) { return Shoup{p->w, p->wq}; }
__device__ __forceinline__ ShoupF ld(gptr<ShoupF> p) { return ShoupF{p->w, p->wp}; }

// Lane-indexed loads of 16-byte table entries through a buffer resource: the
// table base sits in scalar registers, the wave-uniform part of the byte
// offset in soffset and the per-lane part in one VGPR, so a load needs no
// 64-bit address arithmetic of its own.  Word 3 is the CDNA raw-buffer
// format (32-bit data format, no swizzle); the record count only bounds
// offsets, which stay inside the table.  p must be the same on every lane:
// it is read from the first one, so that the compiler never has to loop over
// distinct resources.
template <class Tw>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(gptr<Tw> p) {
    const uint64_t a = (uint64_t)(const Tw *)p;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return __builtin_amdgcn_make_buffer_rsrc((void *)((hi << 32) | lo), (short)0, 0x7fffffff, 0x00020000);
}
template <class Tw>
__device__ __forceinline__ Tw ld_lane(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t uniform_bytes) {
    static_assert(sizeof(Tw) == 16, "16-byte table entries");
    return __builtin_bit_cast(Tw, __builtin_amdgcn_raw_buffer_load_b128(r, lane_bytes, uniform_bytes, 0));
}

// Per-modulus device view of one UintNttTable.
struct DevTable {
    uint64_t q, q2, q4, q8;   // q4 = 4q, q8 = 8q (IntCanonArith, q < 2^60)
    uint64_t inv_n, inv_n_q;  // N^-1 (Shoup)
    uint64_t sr, sr_q;        // N^-1 * inv_root_powers[N-1] (Shoup)
    uint64_t br0, br1;        // Barrett ratio floor(2^128/q)
    uint64_t bar_mu;          // q < 2^59 of b bits: floor(2^(b-1+64)/q) (one-word Barrett), else 0
    uint32_t bar_sh, pad0_;   // b - 1
    const Shoup *fwd;         // root_powers, N entries (bit-reversed order)
    const Shoup *inv;         // inv_root_powers, N entries (scrambled order)
    const Shoup *ord;         // ordinal_root_powers, 2N entries
    // binary64 view, valid when q < 2^50 (fwd_f != nullptr)
    double qd, qinv;
    double q_2p52;            // q + 2^52 (exact for q < 2^50)
    double inv_n_f, inv_n_p, sr_f, sr_p;
    double last_f, last_p;    // inv_root_powers[N-1] without N^-1 (for operands that carry it)
    const ShoupF *fwd_f;
    const ShoupF *inv_f;
    // the same in register order for single-pass launches with 2^(e+3)
    // coefficients per thread (e = 0, 1); see ct_round / gs_round
    const ShoupF *fwd_p[2];
    const ShoupF *inv_p[2];
    uint32_t log_n;
    uint32_t pad_;
};

// ---------------------------------------------------------------------------
// IntArith
// ---------------------------------------------------------------------------

struct IntArith {
    using T = uint64_t;
    using Tw = Shoup;

    __device__ __forceinline__ static gptr<Tw> fwd(const DevTable &t) { return as_global(t.fwd); }
    __device__ __forceinline__ static gptr<Tw> inv(const DevTable &t) { return as_global(t.inv); }
    static constexpr bool kRegisterOrder = false;  // integer tables keep the CPU order
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> fwd_reg(const DevTable &t) { return fwd(t); }
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> inv_reg(const DevTable &t) { return inv(t); }

    // Harvey forward butterfly, primitive.rs:225-228 (every stage reduces:
    // REDUCE is for IntCanonArith's schedule).
    template <bool REDUCE = true>
    __device__ __forceinline__ static void ct(T &a, T &b, const Tw &w, const DevTable &t) {
        uint64_t u = sub_if_ge(a, t.q2);
        uint64_t v = shoup_lazy(w.w, w.wq, b, t.q);
        a = u + v;
        b = u + t.q2 - v;
    }
    static constexpr bool kReduceAtRoundStart = false;
    // resident operands are kept unscaled (their Shoup pairs serve every mode)
    static constexpr bool kScaledOperand = false;
    __device__ __forceinline__ static void round_start(T &, const DevTable &) {}
    // Gentleman-Sande inverse butterfly, primitive.rs:262-265.
    __device__ __forceinline__ static void gs(T &a, T &b, const Tw &w, uint32_t, const DevTable &t) {
        uint64_t u = a, v = b;
        a = sub_if_ge(u + v, t.q2);
        b = shoup_lazy(w.w, w.wq, u + t.q2 - v, t.q);
    }
    // Last inverse stage with N^-1 folded in, primitive.rs:270-285.
    __device__ __forceinline__ static void gs_last(T &a, T &b, const DevTable &t) {
        uint64_t u = a, v = b;
        a = shoup_full(t.inv_n, t.inv_n_q, u + v, t.q);
        b = shoup_full(t.sr, t.sr_q, u + t.q2 - v, t.q);
    }
    // u64 in [0, 4q) -> lane value
    __device__ __forceinline__ static T load(uint64_t v, const DevTable &) { return v; }
    // the same, for values that go straight into forward rounds
    __device__ __forceinline__ static T load_fwd(uint64_t v, const DevTable &) { return v; }
    // lane value (forward output, [0,4q)) -> canonical u64
    __device__ __forceinline__ static uint64_t canon_fwd(T v, const DevTable &t) {
        return sub_if_ge(sub_if_ge(v, t.q2), t.q);
    }
    // lane value (inverse output, [0,q)) -> u64
    __device__ __forceinline__ static uint64_t out_inv(T v, const DevTable &) { return v; }
    // forward output * resident operand (canonical)
    __device__ __forceinline__ static T mul_operand(T x, const Shoup &w, const DevTable &t) {
        return shoup_full(w.w, w.wq, x, t.q);
    }
    // forward output * forward output, canonical (NttPolynomial::mul_assign, Barrett)
    __device__ __forceinline__ static T mul_pair(T x, T y, const DevTable &t) {
        return barrett_mul(canon_fwd(x, t), canon_fwd(y, t), t.q, t.br0, t.br1);
    }
    // Value ranges at round boundaries, checked in builds with PH_RANGE_CHECK
    // (the reference's debug_assert'ed input ranges, hexl/mod.rs:452-465 and
    // 546-558): forward below 4q, inverse below 2q, inverse outputs canonical.
    __device__ __forceinline__ static bool in_fwd_range(T v, const DevTable &t) { return v < t.q4; }
    __device__ __forceinline__ static bool in_inv_range(T v, const DevTable &t) { return v < t.q2; }
    __device__ __forceinline__ static bool in_out_range(T v, const DevTable &t) { return v < t.q; }
};

// IntArith with HexlNttTable's last inverse stage (the radix-2 path,
// primus_ntt/src/ntt/hexl/radix2.rs:358-374): the sum is reduced below 2q
// before its N^-1 product, and both products stay lazy in [0, 2q).  Every
// other butterfly of HEXL's radix-2 transforms is IntArith's (radix2.rs:40-50,
// 210-218), on the same tables.
struct HexlLazyArith : IntArith {
    __device__ __forceinline__ static void gs_last(T &a, T &b, const DevTable &t) {
        const uint64_t u = a, v = b;
        a = shoup_lazy(t.inv_n, t.inv_n_q, sub_if_ge(u + v, t.q2), t.q);
        b = shoup_lazy(t.sr, t.sr_q, u + t.q2 - v, t.q);
    }
    // outputs in [0, 2q) (output_mod_factor 2)
    __device__ __forceinline__ static bool in_out_range(T v, const DevTable &t) { return v < t.q2; }
};

// ---------------------------------------------------------------------------
// IntCanonArith (q < 2^60, canonical outputs only)
// ---------------------------------------------------------------------------

// The integer butterflies for operations whose callers see only canonical
// values, with the headroom of q < 2^60 (16q < 2^64) spent on cheaper
// steps.  Products use shoup_approx (two multiply-highs and a multiply-add for
// the quotient instead of four multiplies; results in [0, 4q)).  The forward
// reduces its unmultiplied inputs only on alternate stages (ct<REDUCE>): a
// stage that reduces (u = a - 8q if a >= 8q) takes inputs below 16q to
// outputs below 12q, a stage that does not takes inputs below 12q to outputs
// below 16q, and no two stages in a row skip it (ct_round reduces on its even
// local stages, so its first stage always does).  The inverse subtracts 4q
// from each sum (intermediates below 8q) and ends with exact Shoup products,
// canonical after one subtraction of q.
struct IntCanonArith {
    using T = uint64_t;
    using Tw = Shoup;

    __device__ __forceinline__ static gptr<Tw> fwd(const DevTable &t) { return as_global(t.fwd); }
    __device__ __forceinline__ static gptr<Tw> inv(const DevTable &t) { return as_global(t.inv); }
    static constexpr bool kRegisterOrder = false;
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> fwd_reg(const DevTable &t) { return fwd(t); }
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> inv_reg(const DevTable &t) { return inv(t); }

    // forward, see above: u < 12q, v < 4q, outputs below 16q
    template <bool REDUCE = true>
    __device__ __forceinline__ static void ct(T &a, T &b, const Tw &w, const DevTable &t) {
        const uint64_t u = REDUCE ? sub_if_ge(a, t.q8) : a;
        const uint64_t v = shoup_approx(w.w, w.wq, b, t.q);
        a = u + v;
        b = u + t.q4 - v;
    }
    static constexpr bool kReduceAtRoundStart = false;
    static constexpr bool kScaledOperand = false;
    __device__ __forceinline__ static void round_start(T &, const DevTable &) {}
    // inverse: inputs below 4q; the sum reduced below 4q, the difference
    // (positive, below 8q) multiplied to below 4q
    __device__ __forceinline__ static void gs(T &a, T &b, const Tw &w, uint32_t, const DevTable &t) {
        const uint64_t u = a, v = b;
        a = sub_if_ge(u + v, t.q4);
        b = shoup_approx(w.w, w.wq, u + t.q4 - v, t.q);
    }
    // last inverse stage, N^-1 folded in (primitive.rs:270-285), exact Shoup: canonical
    __device__ __forceinline__ static void gs_last(T &a, T &b, const DevTable &t) {
        const uint64_t u = a, v = b;
        a = shoup_full(t.inv_n, t.inv_n_q, u + v, t.q);
        b = shoup_full(t.sr, t.sr_q, u + t.q4 - v, t.q);
    }
    __device__ __forceinline__ static T load(uint64_t v, const DevTable &) { return v; }      // [0, 2q)
    __device__ __forceinline__ static T load_fwd(uint64_t v, const DevTable &) { return v; }  // [0, 4q)
    // forward output, below 16q
    __device__ __forceinline__ static uint64_t canon_fwd(T v, const DevTable &t) {
        return sub_if_ge(sub_if_ge(sub_if_ge(sub_if_ge(v, t.q8), t.q4), t.q2), t.q);
    }
    __device__ __forceinline__ static uint64_t out_inv(T v, const DevTable &) { return v; }
    // forward output times a resident operand: [0, 4q), an inverse input as it is
    __device__ __forceinline__ static T mul_operand(T x, const Shoup &w, const DevTable &t) {
        return shoup_approx(w.w, w.wq, x, t.q);
    }
    __device__ __forceinline__ static T mul_pair(T x, T y, const DevTable &t) {
        return barrett_mul(canon_fwd(x, t), canon_fwd(y, t), t.q, t.br0, t.br1);
    }
    // value ranges (as IntArith's): forward below 16q, inverse below 4q
    __device__ __forceinline__ static bool in_fwd_range(T v, const DevTable &t) { return v < t.q8 + t.q8; }
    __device__ __forceinline__ static bool in_inv_range(T v, const DevTable &t) { return v < t.q4; }
    __device__ __forceinline__ static bool in_out_range(T v, const DevTable &t) { return v < t.q; }
};

// ---------------------------------------------------------------------------
// F64Arith (q < 2^50)
// ---------------------------------------------------------------------------

struct F64Arith {
    using T = double;
    using Tw = ShoupF;
    static constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52

    __device__ __forceinline__ static gptr<Tw> fwd(const DevTable &t) { return as_global(t.fwd_f); }
    __device__ __forceinline__ static gptr<Tw> inv(const DevTable &t) { return as_global(t.inv_f); }
    // register-order tables of the single-pass shape with 2^loge coefficients per thread
    static constexpr bool kRegisterOrder = true;
    // (built for 8 and 16 coefficients per thread only: fwd_p / inv_p hold two)
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> fwd_reg(const DevTable &t) {
        static_assert(LOGE == 3 || LOGE == 4, "register-order tables exist for 8 and 16 coefficients per thread");
        return as_global(t.fwd_p[LOGE - 3]);
    }
    template <int LOGE>
    __device__ __forceinline__ static gptr<Tw> inv_reg(const DevTable &t) {
        static_assert(LOGE == 3 || LOGE == 4, "register-order tables exist for 8 and 16 coefficients per thread");
        return as_global(t.inv_p[LOGE - 3]);
    }

    // x*w mod q, centred.  Twiddles are stored centred (|w| <= q/2, |wp| < 1/2
    // with |wp - w/q| <= d = 2^-54 + 2^-64), so |x*wp| < 2^51 for |x| < 4q
    // (4q < 2^52) and the C-rounding is exact; |x*w/q - k| <= 1/2 + |x| d,
    // below 0.7503 for |x| < 4q, so |r| <= q (1/2 + |x| d) < 0.7503 q.
    // |h - k q| <= |r| + ulp(h)/2 < 2^53 keeps the last two steps exact.
    __device__ __forceinline__ static double mulmod(double x, double w, double wp, double q) {
        const double h = x * w;
        const double l = __fma_rn(x, w, -h);
        const double k = __fma_rn(x, wp, kRound) - kRound;
        const double t = __fma_rn(-k, q, h);
        return t + l;
    }
    // a - round(a/q)*q, |result| <= q/2 + 1, for |a| < 2^52.
    __device__ __forceinline__ static double reduce(double a, const DevTable &t) {
        const double k = __fma_rn(a, t.qinv, kRound) - kRound;
        return __fma_rn(-k, t.qd, a);
    }
    // Forward register round.  mulmod's output bound is q (1/2 + |x| d)
    // <= q (1/2 + B/16) for |x| <= B q (q d < 1/16), and its operand may be
    // anything below 4q.  Only the first stage's "a" inputs (the unmultiplied
    // half) are reduced as a round starts (|u| <= q/2 + 1); its "b" inputs go
    // straight into mulmod.  From inputs below 4q (a round's loaded values,
    // or the previous round's outputs):
    //   stage 1: v <= 0.75 q, outputs <= 1.25 q;  stage 2: v <= 0.58 q, <= 1.83 q;
    //   stage 3: v <= 0.62 q, <= 2.45 q;  stage 4: v <= 0.66 q, <= 3.11 q,
    // so every mulmod operand is below 4q and every round output below 3.2 q.
    static constexpr bool kReduceAtRoundStart = true;
    // the resident operand's binary64 copies are stored times N^-1, and the
    // inverse of a fused forward -> multiply -> inverse ends in gs_last_prescaled
    static constexpr bool kScaledOperand = true;
    __device__ __forceinline__ static void round_start(T &a, const DevTable &t) { a = reduce(a, t); }
    template <bool REDUCE = true>  // unused: F64Arith reduces at round starts
    __device__ __forceinline__ static void ct(T &a, T &b, const Tw &w, const DevTable &t) {
        const double v = mulmod(b, w.w, w.wp, t.qd);
        const double u = a;
        a = u + v;
        b = u - v;
    }
    // Inverse stage p: sums are reduced on odd p only.  Entering an even stage
    // every value is <= 0.744 q; its sums are <= 1.49 q and its differences,
    // multiplied, <= 0.744 q.  The odd stage after it sees differences <= 2.98 q
    // (multiplied to <= 0.744 q) and reduces its sums (<= q/2 + 1).
    __device__ __forceinline__ static void gs(T &a, T &b, const Tw &w, uint32_t p, const DevTable &t) {
        const double u = a, v = b;
        const double s = u + v;
        a = (p & 1u) ? reduce(s, t) : s;
        b = mulmod(u - v, w.w, w.wp, t.qd);
    }
    // Last stage, N^-1 folded in: inputs <= 2.98 q, outputs <= 0.744 q.
    __device__ __forceinline__ static void gs_last(T &a, T &b, const DevTable &t) {
        const double u = a, v = b;
        a = mulmod(u + v, t.inv_n_f, t.inv_n_p, t.qd);
        b = mulmod(u - v, t.sr_f, t.sr_p, t.qd);
    }
    // Last stage when N^-1 is already in the values (a resident operand
    // stored times N^-1): the sum is only reduced (<= q/2 + 1), the difference
    // takes the stage's plain twiddle.  The transform is linear, so the
    // outputs are those of gs_last on the unscaled values.
    __device__ __forceinline__ static void gs_last_prescaled(T &a, T &b, const DevTable &t) {
        const double u = a, v = b;
        a = reduce(u + v, t);
        b = mulmod(u - v, t.last_f, t.last_p, t.qd);
    }
    // u64 in [0, 4q) -> centred lane value, |x| <= q/2 + 1.  v < 2^52, so
    // the binary64 with exponent field 2^52 and v as its mantissa is 2^52 + v
    // exactly: one OR and one subtraction instead of a general conversion.
    __device__ __forceinline__ static T load(uint64_t v, const DevTable &t) { return reduce(load_fwd(v, t), t); }
    // u64 in [0, 4q) -> lane value in [0, 4q), unreduced: forward rounds reduce
    // every coefficient as they start, so a value loaded for them needs no
    // reduction of its own.
    __device__ __forceinline__ static T load_fwd(uint64_t v, const DevTable &) {
        const double biased = __longlong_as_double((long long)(v | 0x4330000000000000ull));
        return biased - 0x1p52;
    }
    // centred |x| < q -> [0, q): x + 2^52 (+ q when negative) lies in
    // [2^52, 2^53), where binary64 holds every integer and the low 52 bits of
    // the encoding are the integer minus 2^52.
    __device__ __forceinline__ static uint64_t to_canon(double x, const DevTable &t) {
        const double y = x + (x < 0.0 ? t.q_2p52 : 0x1p52);
        return (uint64_t)__double_as_longlong(y) & ((1ull << 52) - 1);
    }
    __device__ __forceinline__ static uint64_t canon_fwd(T v, const DevTable &t) {
        return to_canon(reduce(v, t), t);
    }
    __device__ __forceinline__ static uint64_t out_inv(T v, const DevTable &t) {
        return to_canon(v, t);  // |v| <= 0.744 q
    }
    // Forward output (|x| <= 3.2 q) times a resident operand stored like the
    // twiddles (centred w, w/q correctly rounded): the bounds of mulmod hold
    // as they are, and the product is <= 0.744 q.
    __device__ __forceinline__ static T mul_operand(T x, const ShoupF &w, const DevTable &t) {
        return mulmod(x, w.w, w.wp, t.qd);
    }
    // x*y mod q: both reduced to q/2 + 1, y/q formed on the fly (<= 0.57 q out).
    __device__ __forceinline__ static T mul_pair(T x, T y, const DevTable &t) {
        const double yr = reduce(y, t);
        return mulmod(reduce(x, t), yr, yr * t.qinv, t.qd);
    }
    // value ranges (as IntArith's), from the bounds above: forward |x| < 4q
    // (loaded values, round outputs below 3.2 q), inverse |x| <= 1.5 q, inverse
    // outputs |x| < 0.75 q before their conversion to [0, q)
    __device__ __forceinline__ static bool in_fwd_range(T v, const DevTable &t) { return fabs(v) < 4.0 * t.qd; }
    __device__ __forceinline__ static bool in_inv_range(T v, const DevTable &t) { return fabs(v) <= 1.5 * t.qd; }
    __device__ __forceinline__ static bool in_out_range(T v, const DevTable &t) { return fabs(v) < 0.75 * t.qd; }
};

// ---------------------------------------------------------------------------
// register rounds
// ---------------------------------------------------------------------------

// R forward stages on one register group x[0 .. 2^R), global stages [S, S+R),
// chunk c at stage S.  FENCE: schedule stage by stage (see below).
// PERM: tw is a register-order table (DevTable::fwd_p), in which the twiddle of
// chunk c*2^u + sub at stage S+u sits at 2^(S+u) + sub*2^S + c, so lanes with
// consecutive c read consecutive entries (requires c < 2^S: single pass).
template <class A, int R, bool FENCE = true, bool PERM = false, bool LANE = false>
__device__ __forceinline__ void ct_round(typename A::T *x, uint32_t S, uint32_t c,
                                         gptr<typename A::Tw> tw, const DevTable &t) {
    using Tw = typename A::Tw;
    if constexpr (A::kReduceAtRoundStart) {
        // the first stage's unmultiplied inputs: x[j], x[j + 2^(R-1)] pair up
#pragma unroll
        for (int j = 0; j < (1 << (R - 1)); ++j) A::round_start(x[j], t);
    }
    // LANE: c differs between lanes; the loads go through a buffer resource
    // with the c-term as the lane offset (ld_lane)
    const __amdgpu_buffer_rsrc_t rs = table_rsrc(tw);  // unused unless LANE
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int half = 1 << (R - u - 1);
        const uint32_t lane = PERM ? c : (c << u);
        const uint32_t base = (1u << (S + u)) + lane;
        const uint32_t step = PERM ? (1u << S) : 1u;
        Tw w[1 << (R - 1)];
#pragma unroll
        for (int sub = 0; sub < (1 << u); ++sub) {
            if constexpr (LANE)
                w[sub] = ld_lane<Tw>(rs, lane * (uint32_t)sizeof(Tw), ((1u << (S + u)) + sub * step) * (uint32_t)sizeof(Tw));
            else
                w[sub] = ld(tw + base + sub * step);
        }
        // A stage's vector loads all go out before its first butterfly; under
        // register pressure the scheduler would otherwise wait for each load
        // in turn.
        if constexpr (FENCE && LANE) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int sub = 0; sub < (1 << u); ++sub) {
#pragma unroll
            for (int jj = 0; jj < half; ++jj) {
                const int j0 = sub * 2 * half + jj;
                // even local stages reduce (IntCanonArith's schedule); u is a
                // constant once the loop is unrolled, so the branch folds away
                if (u & 1) A::template ct<false>(x[j0], x[j0 + half], w[sub], t);
                else A::template ct<true>(x[j0], x[j0 + half], w[sub], t);
            }
        }
        // one stage at a time: hoisting every stage's twiddles to the top of
        // the round keeps them all live at once and spills at 128 VGPRs.
        // Callers that interleave two independent rounds leave it off.
        if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
    }
}

// R inverse stages on one register group, global inverse stages [P, P+R)
// (gaps 2^P .. 2^(P+R-1)), chunk c = k >> (P+R).  LAST: the round ends with the
// transform's final stage (gap N/2).  PERM: itw is a register-order table
// (DevTable::inv_p): within stage P+u's segment the twiddle of chunk
// c*2^(R-u-1) + sub sits at sub*(N >> (P+R)) + c.
// SCALED: the values already carry N^-1 (F64Arith only), see gs_last_prescaled.
template <class A, int R, bool LAST, bool FENCE = true, bool PERM = false, bool LANE = false, bool SCALED = false>
__device__ __forceinline__ void gs_round(typename A::T *x, uint32_t P, uint32_t c, uint32_t n,
                                         gptr<typename A::Tw> itw, const DevTable &t) {
    using Tw = typename A::Tw;
    const __amdgpu_buffer_rsrc_t rs = table_rsrc(itw);  // LANE: as in ct_round
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int half = 1 << u;
        if (LAST && u == R - 1) {
#pragma unroll
            for (int jj = 0; jj < half; ++jj) {
                if constexpr (SCALED) A::gs_last_prescaled(x[jj], x[jj + half], t);
                else A::gs_last(x[jj], x[jj + half], t);
            }
        } else {
            const uint32_t lane = PERM ? c : (c << (R - u - 1));
            const uint32_t seg = 1u + n - (n >> (P + u));
            const uint32_t base = seg + lane;
            const uint32_t step = PERM ? (n >> (P + R)) : 1u;
            Tw w[1 << (R - 1)];
#pragma unroll
            for (int sub = 0; sub < (1 << (R - u - 1)); ++sub) {
                if constexpr (LANE)
                    w[sub] = ld_lane<Tw>(rs, lane * (uint32_t)sizeof(Tw), (seg + sub * step) * (uint32_t)sizeof(Tw));
                else
                    w[sub] = ld(itw + base + sub * step);
            }
            if constexpr (FENCE && LANE) __builtin_amdgcn_sched_barrier(0);  // as in ct_round
#pragma unroll
            for (int sub = 0; sub < (1 << (R - u - 1)); ++sub) {
#pragma unroll
                for (int jj = 0; jj < half; ++jj) {
                    const int j0 = sub * 2 * half + jj;
                    A::gs(x[j0], x[j0 + half], w[sub], P + u, t);
                }
            }
        }
        if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);  // as in ct_round
    }
}

// Two words for 16-byte accesses (global_load/store_dwordx4, ds_write_b128).
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// Coefficient rows are read once and written once per launch: non-temporal
// accesses stream them past the caches, which keep the twiddles, operands and
// keys every workgroup re-reads (and leave less dirty data to write back when
// the kernel ends).  Used where a wave's access covers whole contiguous rows;
// per-thread strided accesses keep the default policy, whose write-back cache
// merges them into full lines.
__device__ __forceinline__ uint64_t ld_stream(const uint64_t *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(uint64_t *p, uint64_t v) { __builtin_nontemporal_store(v, p); }

// A workgroup's contiguous run of `words` coefficients as a buffer resource:
// the base sits in scalar registers, each access is one 32-bit lane offset
// plus a constant (soffset), and no 64-bit address is formed or kept per
// access.  Accesses are non-temporal (aux 2: "nt"), as ld_stream/st_stream.
// `base` must be the same on every lane (it is read from the first).  The
// record count is the run's extent, so an offset past it reads zero and
// writes nothing.
struct RowBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ RowBuf(const uint64_t *base, uint32_t words) {
        const uint64_t a = (uint64_t)base;
        const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
        r = __builtin_amdgcn_make_buffer_rsrc((void *)((hi << 32) | lo), (short)0, (int)(words * 8u), 0x00020000);
    }
    __device__ __forceinline__ uint64_t ld(uint32_t lane_words, uint32_t const_words) const {
        return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, lane_words * 8u, const_words * 8u, 2));
    }
    __device__ __forceinline__ void st(uint32_t lane_words, uint32_t const_words, uint64_t v) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                              r, lane_words * 8u, const_words * 8u, 2);
    }
};

// LDS slot of local index k: one pad word per 16 keeps the stride-16 and
// stride-1 access patterns of the rounds free of bank conflicts.
__host__ __device__ __forceinline__ uint32_t lds_pos(uint32_t k) { return k + (k >> 4); }

}  // namespace ph
