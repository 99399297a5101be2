// copy_kernels.hip -- many host<->device copies in one launch.
//
// The serving loop moves each request's buffers between page-locked host
// memory and the batch arena.  Issued as one DMA copy per buffer, a batch of
// hundreds of requests spends more time in the runtime's per-copy work than
// on the link; here one workgroup per span reads and writes across PCIe
// directly (page-locked memory is mapped into the device's address space),
// 16 bytes per lane where the span allows it.
#include "copy_launch.hpp"

namespace ph {

__global__ __launch_bounds__(256) void copy_spans_kernel(const CopySpan *spans, uint32_t count) {
    if (blockIdx.x >= count) return;
    const CopySpan sp = spans[blockIdx.x];
    const uint32_t t = threadIdx.x;
    if (((sp.src | sp.dst | sp.bytes) & 15) == 0) {
        const uint4 *s = (const uint4 *)sp.src;
        uint4 *d = (uint4 *)sp.dst;
        const uint64_t words = sp.bytes >> 4;
        for (uint64_t i = t; i < words; i += 256) d[i] = s[i];
    } else if (((sp.src | sp.dst | sp.bytes) & 7) == 0) {
        const uint64_t *s = (const uint64_t *)sp.src;
        uint64_t *d = (uint64_t *)sp.dst;
        const uint64_t words = sp.bytes >> 3;
        for (uint64_t i = t; i < words; i += 256) d[i] = s[i];
    } else {
        const unsigned char *s = (const unsigned char *)sp.src;
        unsigned char *d = (unsigned char *)sp.dst;
        for (uint64_t i = t; i < sp.bytes; i += 256) d[i] = s[i];
    }
}

hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    copy_spans_kernel<<<dim3(count), dim3(256), 0, s>>>(spans, count);
    return hipGetLastError();
}

}  // namespace ph
