// capi_internal.hpp -- entry points shared inside the library (not part of
// the C ABI): forms of the public calls for callers that stage their own
// memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/primus_hip.h"

namespace ph {

// Bytes of host staging and of device memory the RLWE item list of these
// requests takes (one item per request and limb), and the device scratch
// its launch needs; both are 0 on an invalid request list.
struct RlweStageSize {
    uint64_t item_bytes;
    uint64_t scratch_bytes;
};
RlweStageSize rlwe_stage_size(const ph_rlwe_ctx *ctx, const ph_rlwe_request *requests, uint32_t n_requests);

// ph_rlwe_encrypt_device without allocation or host synchronisation: the
// item list is written to host_items (page-locked, stage.item_bytes or more)
// and copied from there to d_items on the stream, and the launch uses
// d_scratch (stage.scratch_bytes or more).  host_items must stay untouched
// until the stream has passed this call.
int rlwe_encrypt_staged(const ph_rlwe_ctx *ctx, const ph_rlwe_request *requests, uint32_t n_requests,
                        const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e, const uint8_t *d_m,
                        uint64_t *d_out, hipStream_t stream, void *host_items, void *d_items, void *d_scratch,
                        const RlweStageSize &capacity);

// Device scratch ph_dcrt_negacyclic_mul_rows_device allocates for `count`
// limb-polynomials (0 for the single-pass ring degrees), and that call with
// the scratch supplied: no allocation inside.
uint64_t negacyclic_mul_scratch_bytes(const ph_dcrt_table *table, uint64_t count);
int negacyclic_mul_rows_staged(const ph_dcrt_table *table, uint64_t *d_a, const uint64_t *d_b,
                               const uint8_t *d_limb_of, uint64_t count, void *d_scratch, hipStream_t stream);

}  // namespace ph
