// fhe_kernels.hip -- fused gfx950 kernels for the two batched callers of the
// NTT on the reference's hot path:
//
// * ext_product_kernel: RLWE x RGSW external product (k = 1) with the
//   approximate signed gadget decomposition, one workgroup per product:
//   decompose each input polynomial into l digit polynomials, forward-transform
//   each digit in LDS, multiply-accumulate it against the resident key rows,
//   inverse-transform the two accumulators.  Follows
//   CrtGlwe::mul_dcrt_ggsw_inplace (primus_lattice/src/glwe/crt.rs:202-229) ->
//   DcrtGlwe::add_dcrt_glev_mul_crt_poly_assign (primus_lattice/src/glwe/dcrt.rs:170-247).
// * rlwe_encrypt_kernel: persistent, uneven batch of RLWE encryptions
//   b = NTT(e + Delta*m) + a (.) NTT(s) over mixed ring degrees and limb
//   counts in one launch.  Follows NttRlweSecretKey::encrypt_inplace_with_embedding
//   (primus_fhe_core/src/secret_key/rlwe.rs:227-268).
#include <type_traits>

#include "fhe_launch.hpp"
#include "ntt_launch.hpp"
#include "ntt_rounds.hpp"

namespace ph {

// ---------------------------------------------------------------------------
// external product
// ---------------------------------------------------------------------------

// Digit `d` of a coefficient, as the residue the reference transforms:
// unsigned_decompose (primus_decompose/src/big_integer/common.rs:141-152) then
// wrapping_decompose_small_values (primus_rns/src/base.rs:201-236).  `carry`
// holds the carry into this digit and receives the carry out of it.
__device__ __forceinline__ uint64_t gadget_digit(uint64_t adj, uint32_t d, uint32_t &carry, const ExtArgs &g) {
    const uint32_t shift = g.drop_bits + d * g.log_basis;
    const uint64_t temp = ((adj >> shift) & g.basis_minus_one) + carry;
    carry = (temp & g.carry_mask) != 0;
    const uint64_t digit = temp & g.basis_minus_one;
    if (g.basis_is_two) return digit;
    return digit < g.half ? digit : g.q_minus_basis + digit;
}

// The same digit as a signed value for F64Arith: the digit below half the
// basis, digit - basis from half up.  It is congruent to the residue
// gadget_digit gives (q - B + digit), and |value| <= B/2 < q, so the forward
// transform's exact arithmetic takes it as it is.  `bits` is the coefficient
// shifted to the digit's window; 32-bit arithmetic, valid for bases up to 2^30
// (the digit plus its carry stays below 2^31).
__device__ __forceinline__ double signed_digit_f64(uint32_t bits, uint32_t mask, uint32_t carry_mask, uint32_t half,
                                                   bool two, uint32_t &carry) {
    const uint32_t temp = (bits & mask) + carry;
    carry = (temp & carry_mask) != 0;
    const uint32_t digit = temp & mask;
    const int32_t v = (two || digit < half) ? (int32_t)digit : (int32_t)digit - (int32_t)(mask + 1u);
    return (double)v;
}
constexpr uint32_t kSignedDigitMaxLogBasis = 30;

// acc += x * key (mod q), one term of DcrtPolynomial::add_mul_assign
// (primus_poly/src/dcrt/mod.rs:94-114); canonical for IntArith.  For F64Arith
// the exact centred products are summed unreduced: the caller reduces every
// accumulator after mac_reduce_every terms (ext_mac_reduce_interval) and at
// the end.
__device__ __forceinline__ void ext_mac(uint64_t &acc, uint64_t x, const Shoup &key, const DevTable &t) {
    acc = add_mod(acc, shoup_full(key.w, key.wq, x, t.q), t.q);
}
__device__ __forceinline__ void ext_mac(double &acc, double x, const ShoupF &key, const DevTable &t) {
    acc += F64Arith::mulmod(x, key.w, key.wp, t.qd);
}
// After a digit's multiply-accumulate: binary64 accumulators are reduced
// (|acc| <= q/2 + 1) once `every` key products have been added since the last
// reduction; one uniform branch for all of them.  Integer accumulators are
// canonical already.
template <class A, int E>
__device__ __forceinline__ void ext_mac_settle(typename A::T (&acc)[2][E], uint32_t &terms, uint32_t every,
                                               const DevTable &t) {
    if constexpr (std::is_same<A, F64Arith>::value) {
        if (every != 0 && ++terms >= every) {
            terms = 0;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                acc[0][j] = F64Arith::reduce(acc[0][j], t);
                acc[1][j] = F64Arith::reduce(acc[1][j], t);
            }
        }
    }
}
__device__ __forceinline__ gptr<Shoup> ext_key(const ExtArgs &g, const IntArith *) { return as_global(g.key); }
__device__ __forceinline__ gptr<ShoupF> ext_key(const ExtArgs &g, const F64Arith *) { return as_global(g.key_f); }
// Hides the table's twiddle pointers from loop-invariant code motion.
__device__ __forceinline__ void opaque(DevTable &t) {
    asm volatile("" : "+s"(t.fwd), "+s"(t.inv), "+s"(t.fwd_f), "+s"(t.inv_f));
    asm volatile("" : "+s"(t.fwd_p[0]), "+s"(t.fwd_p[1]), "+s"(t.inv_p[0]), "+s"(t.inv_p[1]));
}
__device__ __forceinline__ void ext_finish(uint64_t &, const DevTable &) {}
__device__ __forceinline__ void ext_finish(double &acc, const DevTable &t) { acc = F64Arith::reduce(acc, t); }

// Registers whose key entries (two rows each) the multiply-accumulate loads
// in one batch.  Binary64: four, which fit beside the accumulators at 4 waves
// per SIMD once the key products are summed without a per-term reduction, and
// keep twice the key loads of two in flight.  Integer arithmetic: two (four
// spill there, and 3 waves per SIMD with eight measured slower).
template <class A>
constexpr int ext_key_batch() {
    return std::is_same<A, F64Arith>::value ? 4 : 2;
}

template <class A, int LOGN>
__global__ __launch_bounds__(1 << (LOGN - kExtLogE), 4) void ext_product_kernel(ExtArgs g) {
    constexpr int LOGT = LOGN - kExtLogE;
    using G = Geo<LOGN, LOGT>;
    using T = typename A::T;
    static_assert(G::E == (1 << kExtLogE), "key permutation assumes this register layout");
    constexpr int R0 = G::fr(0);
    __shared__ T lds[G::LDS];
    const uint32_t t = threadIdx.x;
    const uint64_t prod = blockIdx.x;
    const DevTable tab = *g.tab;
    const uint32_t n = 1u << LOGN;
    const auto key = ext_key(g, (const A *)nullptr);
    T acc[2][G::E];
#pragma unroll
    for (int j = 0; j < G::E; ++j) acc[0][j] = acc[1][j] = (T)0;

    uint32_t terms = 0;  // key products since the last reduction (binary64)
    for (uint32_t p = 0; p < 2; ++p) {
        const uint64_t *src = g.in + (prod * 2 + p) * n;
        // init_value_carry_slice_inplace (primus_decompose/src/big_integer/basis.rs:307-331)
        uint64_t adj[G::E];
        uint32_t carry[G::E];
#pragma unroll
        for (int m = 0; m < (G::E >> R0); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R0); ++j) {
                uint64_t v = ld_stream(src + fwd_slot<LOGN, LOGT, 0, R0>(t, m, j));
                if (g.has_split && v >= g.split) v += g.pow2_sub_q;
                adj[(m << R0) + j] = v;
                carry[(m << R0) + j] = g.drop_bits ? (uint32_t)((v >> (g.drop_bits - 1)) & 1u) : 0u;
            }
        for (uint32_t d = 0; d < g.l; ++d) {
            T x[G::E];
            if (g.unsigned_digits) {
                // below the basis, so below 2q: a forward input as it is
                const uint32_t shift = g.drop_bits + d * g.log_basis;
#pragma unroll
                for (int i = 0; i < G::E; ++i) x[i] = A::load_fwd((adj[i] >> shift) & g.basis_minus_one, tab);
            } else if constexpr (std::is_same<A, F64Arith>::value) {
                if (g.log_basis <= kSignedDigitMaxLogBasis) {
                    const uint32_t shift = g.drop_bits + d * g.log_basis;
#pragma unroll
                    for (int i = 0; i < G::E; ++i)
                        x[i] = signed_digit_f64((uint32_t)(adj[i] >> shift), (uint32_t)g.basis_minus_one,
                                                (uint32_t)g.carry_mask, (uint32_t)g.half, g.basis_is_two != 0,
                                                carry[i]);
                } else {
#pragma unroll
                    for (int i = 0; i < G::E; ++i) x[i] = A::load_fwd(gadget_digit(adj[i], d, carry[i], g), tab);
                }
            } else {
#pragma unroll
                for (int i = 0; i < G::E; ++i) x[i] = A::load_fwd(gadget_digit(adj[i], d, carry[i], g), tab);
            }
            // The twiddles are the same for every digit; without this the
            // compiler hoists all their loads out of the digit loop and keeps
            // them live in registers for the whole kernel.
            // The same holds for the per-thread LDS slot addresses.
            DevTable td = tab;
            opaque(td);
            uint32_t tl = t;
            asm volatile("" : "+v"(tl));
            FwdRounds<A, LOGN, LOGT, 0, true, A::kRegisterOrder>::run(x, lds, tl, 0, 0, td);
            // thread t now holds slots t*E + j; key rows are stored permuted so
            // that (row, t*E + j) sits at row*N + j*T + t: coalesced loads,
            // one lane offset (t) and a uniform offset per register.  The
            // loads of KB registers' key entries are all issued before their
            // products: left to itself the scheduler waits for each load in
            // turn.
            using W = typename A::Tw;
            const auto krs = table_rsrc(key + ((uint64_t)(p * g.l + d) * 2) * n);
            constexpr int KB = ext_key_batch<A>() < G::E ? ext_key_batch<A>() : G::E;
            static_assert(G::E % KB == 0, "key batches tile the register group");
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j0 = 0; j0 < G::E; j0 += KB) {
                W k0[KB], k1[KB];
#pragma unroll
                for (int jj = 0; jj < KB; ++jj) {
                    k0[jj] = ld_lane<W>(krs, t * (uint32_t)sizeof(W), (j0 + jj) * G::T * (uint32_t)sizeof(W));
                    k1[jj] = ld_lane<W>(krs, t * (uint32_t)sizeof(W), ((j0 + jj) * G::T + n) * (uint32_t)sizeof(W));
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int jj = 0; jj < KB; ++jj) {
                    ext_mac(acc[0][j0 + jj], x[j0 + jj], k0[jj], tab);
                    ext_mac(acc[1][j0 + jj], x[j0 + jj], k1[jj], tab);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            ext_mac_settle<A, G::E>(acc, terms, g.mac_reduce_every, tab);
        }
    }
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        uint64_t *dst = g.out + (prod * 2 + o) * n;
        T y[G::E];
#pragma unroll
        for (int j = 0; j < G::E; ++j) {
            y[j] = acc[o][j];
            ext_finish(y[j], tab);
        }
        if (g.ntt_output) {
            // NTT-form result (bit-reversed slots), canonical
#pragma unroll
            for (int j = 0; j < G::E; ++j) dst[((uint64_t)t << G::LOGE) + j] = A::out_inv(y[j], tab);
        } else {
            InvRounds<A, LOGN, LOGT, 0, true, A::kRegisterOrder>::run(y, lds, t, n, 0, tab);
            InvRounds<A, LOGN, LOGT, G::NR - 1, true, A::kRegisterOrder>::store_global(dst, y, t, tab);
        }
    }
}

template <class A, int LOGN>
static hipError_t launch_ext_n(const ExtArgs &g, uint64_t count, hipStream_t s) {
    if (!grid_fits(count, 1u << (LOGN - kExtLogE))) return hipErrorInvalidConfiguration;
    ext_product_kernel<A, LOGN><<<dim3((uint32_t)count), dim3(1u << (LOGN - kExtLogE)), 0, s>>>(g);
    return hipGetLastError();
}

hipError_t launch_ext_product(const ExtArgs &g, uint32_t log_n, bool f64, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    switch (log_n) {
#define PH_EXT_CASE(LN)                                                                        \
    case LN:                                                                                   \
        return f64 ? launch_ext_n<F64Arith, LN>(g, count, s) : launch_ext_n<IntArith, LN>(g, count, s);
        PH_EXT_CASE(6)
        PH_EXT_CASE(7)
        PH_EXT_CASE(8)
        PH_EXT_CASE(9)
        PH_EXT_CASE(10)
        PH_EXT_CASE(11)
        PH_EXT_CASE(12)
#undef PH_EXT_CASE
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// External product over several limbs
// ---------------------------------------------------------------------------

// Word i of a kExtRnsWords-word value held in registers (i uniform or not).
__device__ __forceinline__ uint64_t word_at(const uint64_t (&v)[kExtRnsWords], uint32_t i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

__device__ __forceinline__ bool words_ge(const uint64_t (&v)[kExtRnsWords], const uint64_t (&w)[kExtRnsWords]) {
#pragma unroll
    for (int i = kExtRnsWords - 1; i >= 0; --i)
        if (v[i] != w[i]) return v[i] > w[i];
    return true;
}

__device__ __forceinline__ void words_sub(uint64_t (&v)[kExtRnsWords], const uint64_t (&w)[kExtRnsWords]) {
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < (int)kExtRnsWords; ++i) {
        const uint64_t d = v[i] - w[i];
        const uint64_t b1 = v[i] < w[i];
        v[i] = d - borrow;
        borrow = b1 | (d < borrow);
    }
}

__device__ __forceinline__ void words_add(uint64_t (&v)[kExtRnsWords], const uint64_t (&w)[kExtRnsWords]) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < (int)kExtRnsWords; ++i) {
        const uint64_t s1 = v[i] + w[i];
        const uint64_t c1 = s1 < w[i];
        v[i] = s1 + carry;
        carry = c1 | (v[i] < carry);
    }
}

// One thread per (product, input polynomial, coefficient):
// compose_inplace (primus_rns/src/base.rs:396-420) -- per limb the Shoup
// product r_i * (Q/q_i)^-1 mod q_i times Q/q_i added, Q subtracted on a carry
// or when the sum reaches Q --, init_value_carry (primus_decompose/src/
// big_integer/basis.rs:307-331), the digits with their carries
// (common.rs:141-152), and wrapping_decompose_small_values_inplace
// (base.rs:201-236) into every limb.
__global__ __launch_bounds__(256) void ext_rns_digits_kernel(ExtRnsArgs a) {
    const uint64_t n = 1ull << a.log_n;
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (gid >= a.count * 2 * n) return;
    const uint64_t poly = gid >> a.log_n;  // product * 2 + input polynomial
    const uint64_t k = gid & (n - 1);
    const uint64_t *src = a.in + poly * a.L * n + k;
    uint64_t v[kExtRnsWords] = {0, 0, 0, 0};
    uint64_t Q[kExtRnsWords];
#pragma unroll
    for (int w = 0; w < (int)kExtRnsWords; ++w) Q[w] = a.Q[w];
    for (uint32_t i = 0; i < a.L; ++i) {
        const uint64_t prod = shoup_full(a.inv[i], a.inv_q[i], src[(uint64_t)i * n], a.q[i]);
        uint64_t carry = 0;
#pragma unroll
        for (int w = 0; w < (int)kExtRnsWords; ++w) {
            // prod * punct + v + carry < 2^128: the high word takes both carries
            const uint64_t lo = prod * a.punct[i][w], hi = mulhi64(prod, a.punct[i][w]);
            const uint64_t s1 = v[w] + lo;
            uint64_t c = s1 < lo;
            const uint64_t s2 = s1 + carry;
            c += s2 < carry;
            v[w] = s2;
            carry = hi + c;
        }
        if (carry != 0 || words_ge(v, Q)) words_sub(v, Q);
    }
    if (a.has_split) {
        uint64_t sp[kExtRnsWords], ps[kExtRnsWords];
#pragma unroll
        for (int w = 0; w < (int)kExtRnsWords; ++w) {
            sp[w] = a.split[w];
            ps[w] = a.pow2_sub[w];
        }
        if (words_ge(v, sp)) words_add(v, ps);
    }
    uint64_t carry = 0;
    if (a.drop_bits > 0) carry = (word_at(v, (a.drop_bits - 1) >> 6) >> ((a.drop_bits - 1) & 63)) & 1;
    uint64_t *dst = a.digits + poly * a.l * a.L * n + k;
    for (uint32_t d = 0; d < a.l; ++d) {
        const uint32_t sh = a.drop_bits + d * a.log_basis;
        const uint32_t wi = sh >> 6, off = sh & 63;
        uint64_t bits = word_at(v, wi) >> off;
        if (off + a.log_basis > 64 && wi + 1 < kExtRnsWords) bits |= word_at(v, wi + 1) << (64 - off);
        const uint64_t temp = (bits & a.basis_minus_one) + carry;
        carry = (temp & a.carry_mask) != 0;
        const uint64_t digit = temp & a.basis_minus_one;
        const bool low = a.basis == 2 || digit < a.half;
        for (uint32_t i = 0; i < a.L; ++i) dst[((uint64_t)d * a.L + i) * n] = low ? digit : a.q[i] - a.basis + digit;
    }
}

// One thread per (product, limb, coefficient), both output polynomials c: sum
// over the input polynomials p and digits d of digit_hat * key[p][d][c] mod q_i
// (add_dcrt_glwe_mul_dcrt_polynomial_assign, primus_lattice/src/glwe/dcrt.rs:101-119),
// canonical; each digit is read once for both.
__global__ __launch_bounds__(256) void ext_rns_mac_kernel(ExtRnsMacArgs a) {
    const uint64_t n = 1ull << a.log_n;
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (gid >= a.count * a.L * n) return;
    const uint64_t k = gid & (n - 1);
    const uint64_t row = gid >> a.log_n;  // product * L + i
    const uint32_t i = (uint32_t)(row % a.L);
    const uint64_t prod = row / a.L;
    const uint64_t q = a.tabs[i].q;
    uint64_t acc0 = 0, acc1 = 0;
    for (uint32_t p = 0; p < 2; ++p)
        for (uint32_t d = 0; d < a.l; ++d) {
            const uint64_t x = a.digits[(((prod * 2 + p) * a.l + d) * a.L + i) * n + k];
            const uint64_t kr = (((uint64_t)p * a.l + d) * 2) * a.L + i;  // key row of c = 0
            const Shoup w0 = a.key[kr * n + k], w1 = a.key[(kr + a.L) * n + k];
            acc0 = add_mod(acc0, shoup_full(w0.w, w0.wq, x, q), q);
            acc1 = add_mod(acc1, shoup_full(w1.w, w1.wq, x, q), q);
        }
    a.out[((prod * 2) * a.L + i) * n + k] = acc0;
    a.out[((prod * 2 + 1) * a.L + i) * n + k] = acc1;
}

hipError_t launch_ext_rns_digits(const ExtRnsArgs &a, hipStream_t s) {
    const uint64_t threads = a.count * 2 * (1ull << a.log_n);
    if (threads == 0) return hipSuccess;
    if (!grid_fits((threads + 255) / 256, 256)) return hipErrorInvalidConfiguration;
    ext_rns_digits_kernel<<<dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_ext_rns_mac(const ExtRnsMacArgs &a, hipStream_t s) {
    const uint64_t threads = a.count * a.L * (1ull << a.log_n);
    if (threads == 0) return hipSuccess;
    if (!grid_fits((threads + 255) / 256, 256)) return hipErrorInvalidConfiguration;
    ext_rns_mac_kernel<<<dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// External product over several limbs, fused
// ---------------------------------------------------------------------------

// Multiply-accumulate of one transformed digit with a key entry, per
// arithmetic: canonical (IntArith), below 4q (IntCanonArith: products in
// [0, 4q), one subtraction of 4q per term), centred and summed unreduced
// between ext_mac_settle's reductions (F64Arith, as ext_mac).
template <class A>
struct ExtRnsOps;
template <>
struct ExtRnsOps<IntArith> {
    __device__ __forceinline__ static void mac(uint64_t &acc, uint64_t x, const Shoup &k, const DevTable &t) {
        acc = add_mod(acc, shoup_full(k.w, k.wq, x, t.q), t.q);
    }
    __device__ __forceinline__ static uint64_t finish(uint64_t acc, const DevTable &) { return acc; }
    __device__ __forceinline__ static uint64_t canon(uint64_t acc, const DevTable &) { return acc; }
};
template <>
struct ExtRnsOps<IntCanonArith> {
    __device__ __forceinline__ static void mac(uint64_t &acc, uint64_t x, const Shoup &k, const DevTable &t) {
        acc = sub_if_ge(acc + shoup_approx(k.w, k.wq, x, t.q), t.q4);
    }
    __device__ __forceinline__ static uint64_t finish(uint64_t acc, const DevTable &) { return acc; }  // < 4q
    __device__ __forceinline__ static uint64_t canon(uint64_t acc, const DevTable &t) {
        return sub_if_ge(sub_if_ge(acc, t.q2), t.q);
    }
};
template <>
struct ExtRnsOps<F64Arith> {
    __device__ __forceinline__ static void mac(double &acc, double x, const ShoupF &k, const DevTable &t) {
        ext_mac(acc, x, k, t);
    }
    __device__ __forceinline__ static double finish(double acc, const DevTable &t) { return F64Arith::reduce(acc, t); }
    __device__ __forceinline__ static uint64_t canon(double acc, const DevTable &t) { return F64Arith::to_canon(acc, t); }
};
__device__ __forceinline__ gptr<Shoup> ext_rns_key(const ExtRnsFusedArgs &a, const IntArith *) {
    return as_global(a.key);
}
__device__ __forceinline__ gptr<Shoup> ext_rns_key(const ExtRnsFusedArgs &a, const IntCanonArith *) {
    return as_global(a.key);
}
__device__ __forceinline__ gptr<ShoupF> ext_rns_key(const ExtRnsFusedArgs &a, const F64Arith *) {
    return as_global(a.key_f);
}

// Borrow out of the L-word subtraction x - y (y uniform): 0 when x >= y.
// 32-bit halves through the hardware borrow chain; the differences go unused.
template <int L>
__device__ __forceinline__ unsigned words_borrow(const uint64_t (&x)[L], const uint64_t *y) {
    unsigned c = 0;
#pragma unroll
    for (int w = 0; w < L; ++w) {
        (void)__builtin_subc((unsigned)x[w], (unsigned)y[w], c, &c);
        (void)__builtin_subc((unsigned)(x[w] >> 32), (unsigned)(y[w] >> 32), c, &c);
    }
    return c;
}

// One coefficient of one input polynomial composed into L words: the CRT
// value below Q of residues r_i < q_i (what compose_inplace,
// primus_rns/src/base.rs:396-420, sums) in Garner's mixed-radix form, one
// Shoup product per limb above the first plus the Horner steps of the
// partial value; then shifted by the split adjustment (init_value_carry,
// primus_decompose/src/big_integer/basis.rs:307-331: from split up the value
// stands for value - Q, taken mod 2^bits).  Returns its digit window: the 64
// bits from bit drop_bits - 1 (the carry into the first digit, when
// drop_bits > 0) or from bit 0, which hold every digit when
// l * log_basis + 1 <= 64.  src: the coefficient in limb 0, limbs N words
// apart.
template <int L>
__device__ __forceinline__ uint64_t ext_rns_window(const ExtRnsFusedArgs &a, const uint64_t *src, uint64_t n) {
    const ExtRnsArgs &g = a.g;
    const ExtRnsGarner &k = a.garner;
    uint64_t v[L];
    v[0] = src[0];
#pragma unroll
    for (int i = 1; i < L; ++i) {
        // (v_0 + v_1 q_0 + ... + v_{i-1} q_0..q_{i-2}) mod q_i, below q_i + q_0
        uint64_t part = v[i - 1];
#pragma unroll
        for (int j = i - 2; j >= 0; --j) part = shoup_full(k.qm[i][j], k.qm_q[i][j], part, g.q[i]) + v[j];
        v[i] = shoup_full(k.c[i], k.c_q[i], src[(uint64_t)i * n] + k.off[i] - part, g.q[i]);
    }
    // x = v_{L-1}; x = x q_j + v_j for j = L-2 .. 0 (x grows by one word a step)
    uint64_t x[L];
    x[0] = v[L - 1];
#pragma unroll
    for (int w = 1; w < L; ++w) x[w] = 0;
#pragma unroll
    for (int j = L - 2; j >= 0; --j) {
        uint64_t carry = v[j];
#pragma unroll
        for (int w = 0; w < L - 1 - j; ++w) {
            const uint64_t lo = x[w] * g.q[j] + carry;
            carry = mulhi64(x[w], g.q[j]) + (lo < carry);
            x[w] = lo;
        }
        x[L - 1 - j] = carry;
    }
    if (g.has_split && words_borrow<L>(x, g.split) == 0) {
        // x - Q mod 2^(64 L): the same bits below bit `bits` as x + pow2_sub
        unsigned c = 0;
#pragma unroll
        for (int w = 0; w < L; ++w) {
            const unsigned lo = __builtin_subc((unsigned)x[w], (unsigned)g.Q[w], c, &c);
            const unsigned hi = __builtin_subc((unsigned)(x[w] >> 32), (unsigned)(g.Q[w] >> 32), c, &c);
            x[w] = ((uint64_t)hi << 32) | lo;
        }
    }
    const uint32_t ws = g.drop_bits > 0 ? g.drop_bits - 1 : 0;
    const uint32_t wi = ws >> 6, off = ws & 63;
    uint64_t lo = x[0], hi = L > 1 ? x[1] : 0;
#pragma unroll
    for (int w = 1; w < L; ++w)
        if (wi == (uint32_t)w) {
            lo = x[w];
            hi = w + 1 < L ? x[(w + 1) % L] : 0;
        }
    return off ? (lo >> off) | (hi << (64 - off)) : lo;
}

// Digit d of a coefficient from its window, with the carry chain
// (unsigned_decompose, primus_decompose/src/big_integer/common.rs:141-152),
// wrapped into limb q (wrapping_decompose_small_values_inplace,
// primus_rns/src/base.rs:201-236).  first: 1 when the window's bit 0 is the
// carry into digit 0 (drop_bits > 0), else 0.
__device__ __forceinline__ uint64_t ext_rns_digit(const ExtRnsArgs &g, uint64_t window, uint32_t first, uint32_t d,
                                                  uint32_t &carry, uint64_t q) {
    const uint64_t bits = window >> (first + d * g.log_basis);
    const uint64_t temp = (bits & g.basis_minus_one) + carry;
    carry = (temp & g.carry_mask) != 0;
    const uint64_t digit = temp & g.basis_minus_one;
    return (g.basis == 2 || digit < g.half) ? digit : q - g.basis + digit;
}

// Two launch forms.
// * Per (product, limb) (COOP false): workgroup b handles product
//   8*(b / (8L)) + b % 8 and limb (b % (8L)) / 8.  The L workgroups of a
//   product sit on one XCD (workgroups go to the XCDs in turn) and start close
//   together, so the L - 1 re-reads of its input rows are served by that XCD's
//   L2; each composes every coefficient itself.
// * Per product (COOP true): one workgroup of L x 2^LOGT threads, thread
//   group h (2^LOGT threads, whole waves) running limb h's transforms in its
//   own slice of the LDS.  Each coefficient is composed once: all threads
//   compose the N coefficients of an input polynomial in natural order
//   (coalesced residue loads) and leave their digit windows in the LDS, which
//   the transforms have not used yet; each thread then reads the windows of
//   its first-round slots.  The limbs' transforms run in lockstep, so the
//   workgroup barriers of their LDS exchanges are reached by all threads.
// Requires l * log_basis + (drop_bits > 0) <= 64 (the digits fit one window).
// Four waves per SIMD (128 VGPRs: the two accumulators, the windows, a
// transform and the key batch); integer arithmetic above N = 2^10 needs more
// and runs three.
template <class A>
constexpr int ext_rns_waves(int log_n) {
    return std::is_same<A, F64Arith>::value || log_n <= 10 ? 4 : 3;
}
template <class A, int LOGN, int L, bool COOP>
__global__ __launch_bounds__((COOP ? L : 1) << (LOGN - kExtLogE), ext_rns_waves<A>(LOGN)) void ext_rns_fused_kernel(
    ExtRnsFusedArgs a) {
    constexpr int LOGT = LOGN - kExtLogE;
    using G = Geo<LOGN, LOGT>;
    using T = typename A::T;
    using Ops = ExtRnsOps<A>;
    static_assert(G::E == (1 << kExtLogE), "key permutation assumes this register layout");
    static_assert(!COOP || LOGT >= 6, "a limb's threads must be whole waves");
    constexpr int R0 = G::fr(0);
    __shared__ T lds_all[(COOP ? L : 1) * G::LDS];
    const ExtRnsArgs &g = a.g;
    uint32_t t, limb;
    uint64_t prod;
    if constexpr (COOP) {
        t = threadIdx.x & ((1u << LOGT) - 1);
        limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> LOGT);  // the same on a wave's lanes
        prod = blockIdx.x;  // the grid is the product count
    } else {
        t = threadIdx.x;
        const uint32_t group = 8u * g.L;
        const uint32_t r = blockIdx.x % group;
        prod = (uint64_t)(blockIdx.x / group) * 8u + (r & 7u);
        limb = r >> 3;
        if (prod >= g.count) return;  // the whole workgroup: no barrier reached
    }
    T *lds = lds_all + (COOP ? limb * G::LDS : 0);
    const DevTable tab = a.tabs[limb];
    const uint64_t n = 1ull << LOGN;
    const auto key = ext_rns_key(a, (const A *)nullptr);
    T acc[2][G::E];
#pragma unroll
    for (int j = 0; j < G::E; ++j) acc[0][j] = acc[1][j] = (T)0;

    uint32_t terms = 0;  // key products since the last reduction (binary64)
    for (uint32_t p = 0; p < 2; ++p) {
        // the thread's first-round coefficients, composed
        const uint64_t *src = g.in + (prod * 2 + p) * g.L * n;
        const uint32_t first = g.drop_bits > 0 ? 1u : 0u;
        uint64_t win[G::E];
        uint32_t carry[G::E];
        if constexpr (COOP) {
            // the previous polynomial's last exchange is read by every thread
            // before the windows overwrite the LDS, and the windows are all
            // written before any thread reads its own
            uint64_t *wl = reinterpret_cast<uint64_t *>(lds_all);
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < (uint32_t)n; k += (uint32_t)L << LOGT)
                wl[k] = ext_rns_window<L>(a, src + k, n);
            __syncthreads();
#pragma unroll
            for (int m = 0; m < (G::E >> R0); ++m)
#pragma unroll
                for (int j = 0; j < (1 << R0); ++j) {
                    const int e = (m << R0) + j;
                    win[e] = wl[fwd_slot<LOGN, LOGT, 0, R0>(t, m, j)];
                    carry[e] = (uint32_t)(win[e] & first);
                }
            // the first transform's first exchange stores after a workgroup
            // barrier (SYNC0), so these reads finish before the LDS is reused
        } else {
#pragma unroll
            for (int m = 0; m < (G::E >> R0); ++m)
#pragma unroll
                for (int j = 0; j < (1 << R0); ++j) {
                    const int e = (m << R0) + j;
                    win[e] = ext_rns_window<L>(a, src + fwd_slot<LOGN, LOGT, 0, R0>(t, m, j), n);
                    carry[e] = (uint32_t)(win[e] & first);
                }
        }
        for (uint32_t d = 0; d < g.l; ++d) {
            T x[G::E];
            if constexpr (std::is_same<A, F64Arith>::value) {
                if (g.log_basis <= kSignedDigitMaxLogBasis) {
                    const uint32_t shift = first + d * g.log_basis;
#pragma unroll
                    for (int e = 0; e < G::E; ++e)
                        x[e] = signed_digit_f64((uint32_t)(win[e] >> shift), (uint32_t)g.basis_minus_one,
                                                (uint32_t)g.carry_mask, (uint32_t)g.half, g.basis == 2, carry[e]);
                } else {
#pragma unroll
                    for (int e = 0; e < G::E; ++e)
                        x[e] = A::load_fwd(ext_rns_digit(g, win[e], first, d, carry[e], tab.q), tab);
                }
            } else {
#pragma unroll
                for (int e = 0; e < G::E; ++e)
                    x[e] = A::load_fwd(ext_rns_digit(g, win[e], first, d, carry[e], tab.q), tab);
            }
            // as in ext_product_kernel: keep the twiddle pointers and slot
            // addresses from being hoisted out of the digit loop
            DevTable td = tab;
            opaque(td);
            uint32_t tl = t;
            asm volatile("" : "+v"(tl));
            FwdRounds<A, LOGN, LOGT, 0, true, A::kRegisterOrder>::run(x, lds, tl, 0, 0, td);
            // thread t holds slots t*E + j; key row (p, d, c, limb) stores them at j*T + t
            using Wt = typename A::Tw;
            const uint64_t row0 = (((uint64_t)p * g.l + d) * 2) * g.L + limb;
            const auto k0 = table_rsrc(key + row0 * n);
            const auto k1 = table_rsrc(key + (row0 + g.L) * n);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j0 = 0; j0 < G::E; j0 += ext_key_batch<A>()) {
                Wt w0[ext_key_batch<A>()], w1[ext_key_batch<A>()];
#pragma unroll
                for (int jj = 0; jj < ext_key_batch<A>(); ++jj) {
                    w0[jj] = ld_lane<Wt>(k0, tl * (uint32_t)sizeof(Wt), (j0 + jj) * G::T * (uint32_t)sizeof(Wt));
                    w1[jj] = ld_lane<Wt>(k1, tl * (uint32_t)sizeof(Wt), (j0 + jj) * G::T * (uint32_t)sizeof(Wt));
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int jj = 0; jj < ext_key_batch<A>(); ++jj) {
                    Ops::mac(acc[0][j0 + jj], x[j0 + jj], w0[jj], td);
                    Ops::mac(acc[1][j0 + jj], x[j0 + jj], w1[jj], td);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            ext_mac_settle<A, G::E>(acc, terms, g.mac_reduce_every, td);
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        uint64_t *dst = a.out + ((prod * 2 + c) * g.L + limb) * n;
        T y[G::E];
#pragma unroll
        for (int j = 0; j < G::E; ++j) y[j] = Ops::finish(acc[c][j], tab);
        if (a.ntt_output) {
            // NTT-form result in the reference's slot order, canonical
            const RowBuf rb(dst, (uint32_t)n);
#pragma unroll
            for (int j = 0; j < G::E; ++j) rb.st(t << G::LOGE, (uint32_t)j, Ops::canon(y[j], tab));
        } else {
            InvRounds<A, LOGN, LOGT, 0, true, A::kRegisterOrder>::run(y, lds, t, (uint32_t)n, 0, tab);
            InvRounds<A, LOGN, LOGT, G::NR - 1, true, A::kRegisterOrder>::store_global(dst, y, t, tab);
        }
    }
}

template <class A, int LOGN, int L>
static hipError_t launch_ext_rns_fused_l(const ExtRnsFusedArgs &a, hipStream_t s) {
    constexpr int LOGT = LOGN - (int)kExtLogE;
    // one workgroup per product when the limbs' threads are whole waves and
    // fit one workgroup
    if constexpr (L >= 2 && LOGT >= 6 && (L << LOGT) <= 1024) {
        if (!grid_fits(a.g.count, L << LOGT)) return hipErrorInvalidConfiguration;
        ext_rns_fused_kernel<A, LOGN, L, true><<<dim3((uint32_t)a.g.count), dim3(L << LOGT), 0, s>>>(a);
    } else {
        const uint64_t groups = (a.g.count + 7) / 8;
        if (!grid_fits(groups * 8 * L, 1u << LOGT)) return hipErrorInvalidConfiguration;
        ext_rns_fused_kernel<A, LOGN, L, false><<<dim3((uint32_t)(groups * 8 * L)), dim3(1u << LOGT), 0, s>>>(a);
    }
    return hipGetLastError();
}

template <class A, int LOGN>
static hipError_t launch_ext_rns_fused_n(const ExtRnsFusedArgs &a, hipStream_t s) {
    switch (a.g.L) {
        case 1: return launch_ext_rns_fused_l<A, LOGN, 1>(a, s);
        case 2: return launch_ext_rns_fused_l<A, LOGN, 2>(a, s);
        case 3: return launch_ext_rns_fused_l<A, LOGN, 3>(a, s);
        case 4: return launch_ext_rns_fused_l<A, LOGN, 4>(a, s);
    }
    return hipErrorInvalidValue;
}

template <int LOGN>
static hipError_t launch_ext_rns_fused_a(const ExtRnsFusedArgs &a, int arith, hipStream_t s) {
    if (arith & kArithF64) return launch_ext_rns_fused_n<F64Arith, LOGN>(a, s);
    if (arith & kArithCanonInt) return launch_ext_rns_fused_n<IntCanonArith, LOGN>(a, s);
    return launch_ext_rns_fused_n<IntArith, LOGN>(a, s);
}

hipError_t launch_ext_rns_fused(const ExtRnsFusedArgs &a, int arith, hipStream_t s) {
    if (a.g.count == 0) return hipSuccess;
    switch (a.g.log_n) {
        case 6: return launch_ext_rns_fused_a<6>(a, arith, s);
        case 7: return launch_ext_rns_fused_a<7>(a, arith, s);
        case 8: return launch_ext_rns_fused_a<8>(a, arith, s);
        case 9: return launch_ext_rns_fused_a<9>(a, arith, s);
        case 10: return launch_ext_rns_fused_a<10>(a, arith, s);
        case 11: return launch_ext_rns_fused_a<11>(a, arith, s);
        case 12: return launch_ext_rns_fused_a<12>(a, arith, s);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// RLWE encryption
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t lift_small(int v, uint64_t q) { return v >= 0 ? (uint64_t)v : q - (uint64_t)(-v); }

// e + Delta*m (mod q) from the small values e and m: sample_gaussian_values_inplace
// then add_mul_factor_assign (primus_fhe_core/src/secret_key/rlwe.rs:245-260).
template <class A>
__device__ __forceinline__ uint64_t rlwe_x(const RlweItem &it, int8_t e, uint8_t m, const DevTable &t) {
    if constexpr (std::is_same<A, IntCanonArith>::value)
        // any representative below 16q is a forward input here: Delta*m in
        // [0, 4q) plus e lifted below q
        return lift_small(e, t.q) + shoup_approx(it.delta, it.delta_q, m, t.q);
    else
        return add_mod(lift_small(e, t.q), shoup_full(it.delta, it.delta_q, m, t.q), t.q);
}

// (a*s + x) mod q for a < q and s, x below 16q, q < 2^59 (b bits, b <= 59):
// Barrett with the one-word ratio mu = floor(2^(b-1+64) / q) on
// t = X >> (b-1) < 2^64 (X < 16 q^2 + 16q < 2^(2b+4)): the quotient estimate
// mulhi(t, mu) is at most two short, so two conditional subtractions make the
// remainder canonical.  The two-limb reduction (barrett_mul_add) needs twelve
// partial products where this needs six.
__device__ __forceinline__ uint64_t mac59(uint64_t a, uint64_t s, uint64_t x, const DevTable &t) {
    uint64_t lo = a * s, hi = mulhi64(a, s);
    lo += x;
    hi += lo < x;
    const uint32_t sh = t.bar_sh;  // b - 1, 1 <= sh <= 58
    const uint64_t tq = (lo >> sh) | (hi << (64 - sh));
    const uint64_t k = mulhi64(tq, t.bar_mu);
    const uint64_t r = sub_mul_lo(lo, k, 0 - t.q);
    return sub_if_ge(sub_if_ge(r, t.q2), t.q);
}

// b_hat = x_hat + a_hat (.) s_hat over one chunk of C = 2^LOGC slots starting
// at slot `base` of the row, NttPolynomial::add_mul_assign (rlwe.rs:267), the
// canonical (a*s + x) mod q.  After the forward rounds thread t holds slots
// t*E + j of s_hat and x_hat; both go through LDS so that the MAC runs in row
// order and a_hat and b_hat move as whole rows (a wave's load or store covers
// 64 consecutive words instead of 64 words E apart).  s_hat and x_hat are
// staged as the forward rounds leave them (below 8q): the Barrett reduction
// of a*s + x makes the sum canonical whatever their representatives, since
// a*s + x < 8q^2 + 8q stays far below 2^127, where its quotient estimate is
// at most one short.
template <class A, int LOGC, int LOGT>
__device__ __forceinline__ void rlwe_mac_rows(const RlweArgs &a, const RlweItem &it, uint64_t base, const uint64_t *x,
                                              const uint64_t *sx, uint64_t *lds, uint32_t tid, const DevTable &t) {
    using G = Geo<LOGC, LOGT>;
    uint64_t xv[G::E], av[G::E];
    __syncthreads();  // the rounds' last LDS reads are done
#pragma unroll
    for (int j = 0; j < G::E; ++j) lds[lds_pos((tid << G::LOGE) + j)] = x[j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < G::E; ++i) xv[i] = lds[lds_pos(tid + ((uint32_t)i << LOGT))];
#pragma unroll
    for (int i = 0; i < G::E; ++i) av[i] = ld_stream(a.a + it.row + base + tid + ((uint32_t)i << LOGT));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < G::E; ++j) lds[lds_pos((tid << G::LOGE) + j)] = sx[j];
    __syncthreads();
    // four products at a time: unfenced, the scheduler interleaves all E
    // Barrett reductions and their temporaries no longer fit 128 VGPRs
#pragma unroll
    for (int i0 = 0; i0 < G::E; i0 += 4) {
#pragma unroll
        for (int i = i0; i < i0 + 4 && i < G::E; ++i) {
            const uint32_t k = tid + ((uint32_t)i << LOGT);
            const uint64_t sv = lds[lds_pos(k)];
        st_stream(a.out + it.row + base + k,
                  t.bar_mu ? mac59(av[i], sv, xv[i], t) : barrett_mul_add(av[i], sv, xv[i], t.q, t.br0, t.br1));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// N <= 2^12: both transforms in LDS, 256 threads, E = N / 256.  A: IntArith,
// or IntCanonArith when every modulus is below 2^60 (only canonical values
// leave the kernel).
template <class A, int LOGN>
__device__ __forceinline__ void rlwe_single(const RlweArgs &a, const RlweItem &it, const DevTable &t, uint64_t *lds,
                                            uint32_t tid) {
    constexpr int LOGT = 8;
    using G = Geo<LOGN, LOGT>;
    constexpr int R0 = G::fr(0);
    uint64_t sx[G::E], x[G::E];
    // the small values' loads all go out before any of them is used
    int8_t sv[G::E], ev[G::E];
    uint8_t mv[G::E];
    const uint32_t lt = fwd_slot<LOGN, LOGT, 0, R0>(tid, 0, 0);
#pragma unroll
    for (int m = 0; m < (G::E >> R0); ++m)
#pragma unroll
        for (int j = 0; j < (1 << R0); ++j) {
            const uint64_t k = it.small + lt + fwd_slot<LOGN, LOGT, 0, R0>(0, m, j);
            sv[(m << R0) + j] = a.s[k];
            ev[(m << R0) + j] = a.e[k];
            mv[(m << R0) + j] = a.m[k];
        }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < G::E; ++i) {
        sx[i] = lift_small(sv[i], t.q);
        x[i] = rlwe_x<A>(it, ev[i], mv[i], t);
    }
    FwdRounds<A, LOGN, LOGT, 0, true>::run(sx, lds, tid, 0, 0, t);
    FwdRounds<A, LOGN, LOGT, 0, true>::run(x, lds, tid, 0, 0, t);
    rlwe_mac_rows<A, LOGN, LOGT>(a, it, 0, x, sx, lds, tid, t);
}

// Forward stage 0 of a column of the ternary secret: the multiplied inputs
// are -1, 0 or 1, so each product w*s is 0, w or q - w (w = root_powers[1])
// and no modular multiply is needed.  Outputs are below 2q, a valid input
// of the stages after it in either integer arithmetic.
template <int R>
__device__ __forceinline__ void ternary_stage0(uint64_t *v, const int8_t *sv, const DevTable &t) {
    constexpr int H = 1 << (R - 1);
    const uint64_t w = ld(as_global(t.fwd) + 1).w;
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const int8_t b = sv[j + H];
        const uint64_t p = b == 0 ? 0ull : (b > 0 ? w : t.q - w);
        const uint64_t a = v[j];  // canonical
        v[j] = a + p;
        v[j + H] = a + t.q - p;
    }
}

// N = 2^(12+R): column pass (stages 0..R-1) of s into scratch and of x into
// the output row, then 2^R chunk passes of 4096 in LDS with the MAC fused.
template <class A, int R>
__device__ __forceinline__ void rlwe_two_pass(const RlweArgs &a, const RlweItem &it, const DevTable &t, uint64_t *lds,
                                              uint32_t tid, uint64_t *scr) {
    constexpr int LOGC = 12, LOGT = 8;
    constexpr uint32_t C = 1u << LOGC;
    using G = Geo<LOGC, LOGT>;
    for (uint32_t cc = 0; cc < C / kRlweThreads; ++cc) {
        const uint32_t col = tid + cc * kRlweThreads;
        uint64_t v[1 << R], w[1 << R];
        int8_t sv[1 << R], ev[1 << R];
        uint8_t mv[1 << R];
#pragma unroll
        for (int i = 0; i < (1 << R); ++i) {
            const uint64_t k = it.small + (uint64_t)i * C + col;
            sv[i] = a.s[k];
            ev[i] = a.e[k];
            mv[i] = a.m[k];
        }
        __builtin_amdgcn_sched_barrier(0);  // as in rlwe_single
#pragma unroll
        for (int i = 0; i < (1 << R); ++i) {
            v[i] = lift_small(sv[i], t.q);
            w[i] = rlwe_x<A>(it, ev[i], mv[i], t);
        }
        ternary_stage0<R>(v, sv, t);
        if constexpr (R > 1) {
            ct_round<A, R - 1, false>(v, 1, 0, A::fwd(t), t);
            ct_round<A, R - 1, false>(v + (1 << (R - 1)), 1, 1, A::fwd(t), t);
        }
        ct_round<A, R, false>(w, 0, 0, A::fwd(t), t);
#pragma unroll
        for (int i = 0; i < (1 << R); ++i) {
            scr[(uint64_t)i * C + col] = v[i];
            a.out[it.row + (uint64_t)i * C + col] = w[i];
        }
    }
    __syncthreads();  // the column results are read by other threads of the workgroup
    for (uint32_t b = 0; b < (1u << R); ++b) {
        asm volatile("" : "+v"(tid));  // keep slot addresses out of loop-invariant code motion
        uint64_t sx[G::E], x[G::E];
        FwdRounds<A, LOGC, LOGT, 0, true>::load_global(scr + (uint64_t)b * C, sx, tid, t);
        FwdRounds<A, LOGC, LOGT, 0, true>::run(sx, lds, tid, R, b, t);
        FwdRounds<A, LOGC, LOGT, 0, true>::load_global(a.out + it.row + (uint64_t)b * C, x, tid, t);
        FwdRounds<A, LOGC, LOGT, 0, true>::run(x, lds, tid, R, b, t);
        rlwe_mac_rows<A, LOGC, LOGT>(a, it, (uint64_t)b * C, x, sx, lds, tid, t);
    }
}

// Two waves per SIMD: the 2^12 path holds both transforms (s and e + Delta m) in
// registers at once, and capping it lower spills.
template <class A>
__global__ __launch_bounds__(kRlweThreads, 2) void rlwe_encrypt_kernel(RlweArgs a) {
    __shared__ uint64_t lds[Geo<12, 8>::LDS];
    __shared__ uint32_t next;
    uint64_t *scr = a.scratch + (uint64_t)blockIdx.x * a.scratch_words;
    // Items are taken from the queue, largest ring degrees first, by whichever
    // workgroup is free: the batch's uneven items balance as they finish
    // rather than by a fixed stride.
    for (;;) {
        if (threadIdx.x == 0) next = __hip_atomic_fetch_add(a.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t i = next;
        if (i >= a.n_items) break;
        const RlweItem it = a.items[i];
        const DevTable t = a.tabs[it.tab];
        // Slot addresses depend on the thread alone; hidden from loop-invariant
        // code motion they are recomputed per item instead of being held in
        // registers (for every ring degree at once) across the whole loop.
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        switch (it.log_n) {
            case 10: rlwe_single<A, 10>(a, it, t, lds, tid); break;
            case 11: rlwe_single<A, 11>(a, it, t, lds, tid); break;
            case 12: rlwe_single<A, 12>(a, it, t, lds, tid); break;
            case 13: rlwe_two_pass<A, 1>(a, it, t, lds, tid, scr); break;
            case 14: rlwe_two_pass<A, 2>(a, it, t, lds, tid, scr); break;
            case 15: rlwe_two_pass<A, 3>(a, it, t, lds, tid, scr); break;
            case 16: rlwe_two_pass<A, 4>(a, it, t, lds, tid, scr); break;
            default: break;
        }
        __syncthreads();  // LDS, scratch and `next` are reused by the next item
    }
    // Every workgroup counts itself out after its last take from the queue
    // (release); the last one (acquire: all takes are done) zeroes the queue
    // for the next launch over the same list.
    if (threadIdx.x == 0) {
        const uint32_t done = __hip_atomic_fetch_add(a.queue + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done + 1 == gridDim.x) {
            __hip_atomic_store(a.queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.queue + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Inputs of RLWE encryption drawn on the device, with the distributions of
// the uneven encryption batch: s ternary uniform, e the rounded Gaussian of
// sigma 3.2 cut at 19, m uniform in [0, 256), a_hat uniform mod q_i.  Each
// element's values come from the Philox block of counter (element, request,
// purpose, 0).
__global__ __launch_bounds__(256) void rlwe_sample_kernel(SampleArgs a) {
    const SampleChunk ch = a.chunks[blockIdx.x];
    uint64_t gauss[kGaussThresholds];
    if (ch.purpose == kSampleError) {
#pragma unroll
        for (int i = 0; i < kGaussThresholds; ++i) gauss[i] = a.gauss[i];
    }
    for (uint32_t i = threadIdx.x; i < ch.count; i += 256) {
        uint32_t c[4] = {ch.k0 + i, ch.request, ch.purpose, 0u};
        philox4x32_10(c, a.seed_lo, a.seed_hi);
        const uint64_t lo = (uint64_t)c[0] | ((uint64_t)c[1] << 32);
        const uint64_t hi = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
        switch (ch.purpose) {
            case kSampleSecret: a.s[ch.dst + i] = (int8_t)sample_ternary(lo); break;
            case kSampleError: a.e[ch.dst + i] = (int8_t)sample_gauss(lo, gauss); break;
            case kSampleMessage: a.m[ch.dst + i] = (uint8_t)(c[0] & 255u); break;
            default: a.a[ch.dst + i] = sample_mod(lo, hi, ch.q); break;
        }
    }
}

hipError_t launch_rlwe_sample(const SampleArgs &a, hipStream_t st) {
    if (a.n_chunks == 0) return hipSuccess;
    if (!grid_fits(a.n_chunks, 256)) return hipErrorInvalidConfiguration;
    rlwe_sample_kernel<<<dim3(a.n_chunks), dim3(256), 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_rlwe_encrypt(const RlweArgs &a, uint32_t grid, hipStream_t s) {
    if (a.n_items == 0 || grid == 0) return hipSuccess;
    if (a.canon_int) rlwe_encrypt_kernel<IntCanonArith><<<dim3(grid), dim3(kRlweThreads), 0, s>>>(a);
    else rlwe_encrypt_kernel<IntArith><<<dim3(grid), dim3(kRlweThreads), 0, s>>>(a);
    return hipGetLastError();
}

// Range-check report of this translation unit's kernels (PH_RANGE_CHECK
// builds): {violations, first site, first value's bits, its modulus},
// cleared afterwards when asked.  Blocking copies on the null stream.
hipError_t range_report_fhe(uint64_t *out, bool clear) {
#if PH_RANGE_CHECK
    hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_range_report), 4 * sizeof(uint64_t));
    if (e == hipSuccess && clear) {
        const uint64_t zero[4] = {0, 0, 0, 0};
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_range_report), zero, sizeof(zero));
    }
    return e;
#else
    (void)out, (void)clear;
    return hipErrorNotSupported;
#endif
}

}  // namespace ph
