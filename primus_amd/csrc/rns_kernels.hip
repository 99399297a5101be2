// rns_kernels.hip -- RNS base conversion of CRT polynomials on gfx950.
//
// One thread per coefficient position: it reads the coefficient's residues in
// every input limb (limb-major rows, so a wave's loads and stores cover whole
// rows), scales each by [(Q/q_i)^-1]_{q_i}, and forms every output residue as
// a dot product with the matrix (Q/q_i) mod p_j, reduced exactly (Barrett on
// the 128-bit sums).  The exact variant also rounds sum_i y_i/q_i in binary64,
// operation for operation as the reference does, and subtracts that multiple
// of Q.  Work per coefficient is in_count * out_count multiply-adds.
#include "arith.hpp"
#include "ntt_device.hpp"
#include "rns_launch.hpp"

namespace ph {

// IN / OUT > 0 fix the input and output limb counts at compile time (the limb
// loops then unroll exactly, and the matrix offsets fold); 0 takes them from c.
template <bool EXACT, uint32_t IN, uint32_t OUT>
__global__ __launch_bounds__(256) void base_convert_kernel(BaseConvDev c, const uint64_t *in, uint64_t *out,
                                                           uint32_t log_n, uint64_t polys) {
    static_assert(IN <= kMaxRnsModuli && OUT <= kMaxRnsModuli, "at most kMaxRnsModuli limbs");
    constexpr uint32_t kIn = IN ? IN : kMaxRnsModuli;
    const uint32_t in_count = IN ? IN : c.in_count;
    const uint32_t out_count = OUT ? OUT : c.out_count;
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= (polys << log_n)) return;
    const uint64_t n = 1ull << log_n;
    const uint64_t poly = g >> log_n, k = g & (n - 1);
    const uint64_t *src = in + poly * in_count * n + k;
    const auto q = as_global(c.q), q_r0 = as_global(c.q_r0), q_r1 = as_global(c.q_r1);
    const auto inv = as_global(c.inv), inv_q = as_global(c.inv_q);
    const auto qd = as_global(c.qd);
    const auto m = as_global(c.m);
    uint64_t y[kIn];
    double v = -0.0;
#pragma unroll
    for (uint32_t i = 0; i < kIn; ++i)
        if (i < in_count) y[i] = ld_stream(src + (uint64_t)i * n);  // every row load issued first
    // y_i = [x_i (Q/q_i)^-1]_{q_i}: fill_fast_convert_array_buffer
    // (primus_rns/src/converter.rs:103-146), a factor of one reducing x instead
#pragma unroll
    for (uint32_t i = 0; i < kIn; ++i) {
        if (i < in_count) {
            const uint64_t x = y[i];
            const uint64_t w = inv[i];
            y[i] = w == 1 ? barrett_reduce128(x, 0, q[i], q_r0[i], q_r1[i]) : shoup_full(w, inv_q[i], x, q[i]);
            // exact_convert_array (converter.rs:255-288): v_i = y_i / q_i in
            // binary64, summed in limb order (-0.0 start, as Rust's f64 sum)
            if constexpr (EXACT) v += (double)y[i] / qd[i];
        }
    }
    uint64_t *dst = out + poly * out_count * n + k;
    for (uint32_t j = 0; j < out_count; ++j) {
        const uint64_t p = c.p[j], r0 = c.p_r0[j], r1 = c.p_r1[j];
        // reduce_dot_product: canonical sum_i y_i * ((Q/q_i) mod p_j).  The
        // products are summed in 128 bits, lazy_terms at a time, and each
        // partial sum (below p * 2^64) is reduced once.
        uint64_t acc = 0, lo = 0, hi = 0;
        uint32_t terms = 0;
#pragma unroll
        for (uint32_t i = 0; i < kIn; ++i) {
            if (i < in_count) {
                const uint64_t mm = m[(uint64_t)j * in_count + i];
                const uint64_t plo = y[i] * mm, phi = mulhi64(y[i], mm);
                lo += plo;
                hi += phi + (lo < plo);
                if (++terms == c.lazy_terms) {
                    acc = add_mod(acc, barrett_reduce128_wide(lo, hi, p, r0, r1), p);
                    lo = hi = 0;
                    terms = 0;
                }
            }
        }
        if (terms) acc = add_mod(acc, barrett_reduce128_wide(lo, hi, p, r0, r1), p);
        if constexpr (EXACT) {
            // converter.rs:293-315: v = (sum + 0.5) truncated; out = acc - v (Q mod p)
            const uint64_t vr = (uint64_t)(v + 0.5);
            const uint64_t vq = barrett_mul(barrett_reduce128(vr, 0, p, r0, r1), c.q_mod_p[j], p, r0, r1);
            acc = acc >= vq ? acc - vq : acc + p - vq;
        }
        st_stream(dst + (uint64_t)j * n, acc);
    }
}

// a - round(a/m) m, |result| <= m/2 + 1, for |a| < 2^52 (F64Arith::reduce)
__device__ __forceinline__ double reduce_f(double a, double m, double minv) {
    const double k = __fma_rn(a, minv, F64Arith::kRound) - F64Arith::kRound;
    return __fma_rn(-k, m, a);
}
// centred |x| < m (m < 2^50) -> [0, m) as u64 (F64Arith::to_canon)
__device__ __forceinline__ uint64_t canon_f(double x, double m) {
    const double y = x + (x < 0.0 ? m + 0x1p52 : 0x1p52);
    return (uint64_t)__double_as_longlong(y) & ((1ull << 52) - 1);
}

// The same conversion in F64Arith's exact binary64 integer arithmetic, for
// bases of moduli below 2^50.  y_i is canonical, as in the integer kernel;
// each dot product adds centred products (each within 0.57 p_j, since
// y_i < 2^50) and reduces every third term, so every partial sum stays below
// 2^52.
template <bool EXACT>
__global__ __launch_bounds__(256) void base_convert_f64_kernel(BaseConvDev c, const uint64_t *in, uint64_t *out,
                                                               uint32_t log_n, uint64_t polys) {
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= (polys << log_n)) return;
    const uint64_t n = 1ull << log_n;
    const uint64_t poly = g >> log_n, k = g & (n - 1);
    const uint64_t *src = in + poly * c.in_count * n + k;
    const auto qd = as_global(c.qd);
    const auto inv_f = as_global(c.inv_f), m_f = as_global(c.m_f);
    uint64_t x[kMaxRnsModuli];
#pragma unroll
    for (uint32_t i = 0; i < kMaxRnsModuli; ++i)
        if (i < c.in_count) x[i] = ld_stream(src + (uint64_t)i * n);
    double y[kMaxRnsModuli];
    double v = -0.0;
#pragma unroll
    for (uint32_t i = 0; i < kMaxRnsModuli; ++i) {
        if (i < c.in_count) {
            uint64_t xi = x[i];
            // residues at or above 2^52 (no canonical input is) are reduced
            // exactly first; below it the binary64 encoding is exact
            if (xi >= (1ull << 52)) xi = barrett_reduce128(xi, 0, c.q[i], c.q_r0[i], c.q_r1[i]);
            const double xd = __longlong_as_double((long long)(xi | 0x4330000000000000ull)) - 0x1p52;
            const ShoupF w = ld(inv_f + i);
            // a factor of one reduces x (converter.rs:103-146): the same product.
            // The dot products take y_i canonical: the conversion's result
            // depends on the representative (another one adds a multiple of Q).
            const double yr = F64Arith::mulmod(xd, w.w, w.wp, qd[i]);
            y[i] = yr < 0.0 ? yr + qd[i] : yr;  // exact, in [0, q_i)
            if constexpr (EXACT) v += y[i] / qd[i];
        }
    }
    uint64_t *dst = out + poly * c.out_count * n + k;
    for (uint32_t j = 0; j < c.out_count; ++j) {
        const double p = c.pd[j], pinv = c.pinv[j];
        double acc = 0.0;
#pragma unroll
        for (uint32_t i = 0; i < kMaxRnsModuli; ++i) {
            if (i < c.in_count) {
                const ShoupF w = ld(m_f + ((uint64_t)j * c.in_count + i));
                acc += F64Arith::mulmod(y[i], w.w, w.wp, p);
                if (i % 3 == 2) acc = reduce_f(acc, p, pinv);
            }
        }
        if constexpr (EXACT) {
            // converter.rs:293-315: v = (sum + 0.5) truncated; out = acc - v (Q mod p)
            const double vr = (double)(uint64_t)(v + 0.5);
            const ShoupF w = ld(as_global(c.qmp_f) + j);
            acc -= F64Arith::mulmod(vr, w.w, w.wp, p);
        }
        st_stream(dst + (uint64_t)j * n, canon_f(reduce_f(acc, p, pinv), p));
    }
}

hipError_t launch_base_convert(const BaseConvDev &c, const uint64_t *in, uint64_t *out, uint32_t log_n,
                               uint64_t polys, bool exact, hipStream_t s) {
    const uint64_t threads = polys << log_n;
    if (threads == 0) return hipSuccess;
    if (!grid_fits((threads + 255) / 256, 256)) return hipErrorInvalidConfiguration;
    const dim3 grid((uint32_t)((threads + 255) / 256)), block(256);
    if (c.f64) {
        if (exact) base_convert_f64_kernel<true><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
        else base_convert_f64_kernel<false><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
    } else {
        // the shapes of a 4-limb ciphertext modulus: fast 4 -> 4 (modulus
        // switching, key switching's extension) and exact 4 -> 1
        if (exact && c.in_count == 4 && c.out_count == 1)
            base_convert_kernel<true, 4, 1><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
        else if (exact)
            base_convert_kernel<true, 0, 0><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
        else if (c.in_count == 4 && c.out_count == 4)
            base_convert_kernel<false, 4, 4><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
        else
            base_convert_kernel<false, 0, 0><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
    }
    return hipGetLastError();
}

}  // namespace ph
