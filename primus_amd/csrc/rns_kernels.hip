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

template <bool EXACT>
__global__ __launch_bounds__(256) void base_convert_kernel(BaseConvDev c, const uint64_t *in, uint64_t *out,
                                                           uint32_t log_n, uint64_t polys) {
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= (polys << log_n)) return;
    const uint64_t n = 1ull << log_n;
    const uint64_t poly = g >> log_n, k = g & (n - 1);
    const uint64_t *src = in + poly * c.in_count * n + k;
    const auto q = as_global(c.q), q_r0 = as_global(c.q_r0), q_r1 = as_global(c.q_r1);
    const auto inv = as_global(c.inv), inv_q = as_global(c.inv_q);
    const auto qd = as_global(c.qd);
    const auto m = as_global(c.m);
    uint64_t y[kMaxRnsModuli];
    double v = -0.0;
    // y_i = [x_i (Q/q_i)^-1]_{q_i}: fill_fast_convert_array_buffer
    // (primus_rns/src/converter.rs:103-146), a factor of one reducing x instead
#pragma unroll
    for (uint32_t i = 0; i < kMaxRnsModuli; ++i) {
        if (i < c.in_count) {
            const uint64_t x = ld_stream(src + (uint64_t)i * n);
            const uint64_t w = inv[i];
            y[i] = w == 1 ? barrett_reduce128(x, 0, q[i], q_r0[i], q_r1[i]) : shoup_full(w, inv_q[i], x, q[i]);
            // exact_convert_array (converter.rs:255-288): v_i = y_i / q_i in
            // binary64, summed in limb order (-0.0 start, as Rust's f64 sum)
            if constexpr (EXACT) v += (double)y[i] / qd[i];
        }
    }
    uint64_t *dst = out + poly * c.out_count * n + k;
    for (uint32_t j = 0; j < c.out_count; ++j) {
        const uint64_t p = c.p[j], r0 = c.p_r0[j], r1 = c.p_r1[j];
        // reduce_dot_product: canonical sum_i y_i * ((Q/q_i) mod p_j)
        uint64_t acc = 0;
#pragma unroll
        for (uint32_t i = 0; i < kMaxRnsModuli; ++i)
            if (i < c.in_count) acc = barrett_mul_add(y[i], m[(uint64_t)j * c.in_count + i], acc, p, r0, r1);
        if constexpr (EXACT) {
            // converter.rs:293-315: v = (sum + 0.5) truncated; out = acc - v (Q mod p)
            const uint64_t vr = (uint64_t)(v + 0.5);
            const uint64_t vq = barrett_mul(barrett_reduce128(vr, 0, p, r0, r1), c.q_mod_p[j], p, r0, r1);
            acc = acc >= vq ? acc - vq : acc + p - vq;
        }
        st_stream(dst + (uint64_t)j * n, acc);
    }
}

hipError_t launch_base_convert(const BaseConvDev &c, const uint64_t *in, uint64_t *out, uint32_t log_n,
                               uint64_t polys, bool exact, hipStream_t s) {
    const uint64_t threads = polys << log_n;
    if (threads == 0) return hipSuccess;
    const dim3 grid((uint32_t)((threads + 255) / 256)), block(256);
    if (exact) base_convert_kernel<true><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
    else base_convert_kernel<false><<<grid, block, 0, s>>>(c, in, out, log_n, polys);
    return hipGetLastError();
}

}  // namespace ph
