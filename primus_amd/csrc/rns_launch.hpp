// rns_launch.hpp -- host-side view of the RNS base-conversion kernels (rns_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ph {

struct ShoupF;

constexpr uint32_t kMaxRnsModuli = 32;

// BaseConverter (primus_rns/src/converter.rs:11-46) on the device: every
// pointer is into one allocation owned by the converter.
struct BaseConvDev {
    uint32_t in_count, out_count;
    const uint64_t *q, *q_r0, *q_r1;  // input moduli and their Barrett ratios floor(2^128/q)
    const uint64_t *inv, *inv_q;      // [(Q/q_i)^-1]_{q_i} with its Shoup quotient
    const double *qd;                 // q_i as binary64 (rounded as (double)q_i)
    const uint64_t *p, *p_r0, *p_r1;  // output moduli and their Barrett ratios
    const uint64_t *m;                // [out][in]: (Q/q_i) mod p_j
    const uint64_t *q_mod_p;          // [out]: Q mod p_j
    uint32_t lazy_terms;              // products summed in 128 bits per reduction: floor(2^64 / max q_i)
    // The binary64 form (F64Arith, ntt_device.hpp), present when every input
    // and output modulus is below 2^50; f64 != 0 selects it.
    uint32_t f64;
    const ShoupF *inv_f;  // [(Q/q_i)^-1]_{q_i}, centred pairs {w, w/q}
    const ShoupF *m_f;    // [out][in]: (Q/q_i) mod p_j, centred pairs
    const ShoupF *qmp_f;  // [out]: Q mod p_j, centred pairs
    const double *pd;     // [out]: p_j
    const double *pinv;   // [out]: 1/p_j
};

// `polys` CRT polynomials of 2^log_n coefficients: in [polys][in_count][N],
// out [polys][out_count][N].  exact: exact_convert_array (out_count == 1).
hipError_t launch_base_convert(const BaseConvDev &c, const uint64_t *in, uint64_t *out, uint32_t log_n,
                               uint64_t polys, bool exact, hipStream_t s);

}  // namespace ph
