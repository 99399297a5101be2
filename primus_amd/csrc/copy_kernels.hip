// copy_kernels.hip -- many host<->device copies in one launch.
//
// The serving loop moves each request's buffers between page-locked host
// memory and the batch arena.  Issued as one DMA copy per buffer, a batch of
// hundreds of requests spends more time in the runtime's per-copy work than
// on the link; here a few workgroups read and write across PCIe directly
// (page-locked memory is mapped into the device's address space), one span at
// a time each, 16 bytes per lane where the span allows it.
//
// The grid is kept small on purpose.  A read across PCIe takes microseconds,
// and every request a copy keeps outstanding waits in the memory system's
// queues, in front of the requests of kernels running on other streams.  One
// workgroup per span (hundreds per batch) kept megabytes in flight and held
// the other streams' kernels back for the whole copy; a few dozen workgroups
// with two loads per lane in flight keep 8 KiB each, enough to fill the link.
#include "copy_launch.hpp"

namespace ph {

__global__ __launch_bounds__(256) void copy_spans_kernel(const CopySpan *spans, uint32_t count) {
    const uint32_t t = threadIdx.x;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const CopySpan sp = spans[k];
        if (((sp.src | sp.dst | sp.bytes) & 15) == 0) {
            const uint4 *s = (const uint4 *)sp.src;
            uint4 *d = (uint4 *)sp.dst;
            const uint64_t words = sp.bytes >> 4;
            uint64_t i = t;
            for (; i + 256 < words; i += 512) {
                const uint4 a = s[i], b = s[i + 256];
                d[i] = a;
                d[i + 256] = b;
            }
            if (i < words) d[i] = s[i];
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
}

hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, uint32_t groups, hipStream_t s) {
    if (count == 0) return hipSuccess;
    if (groups == 0) return hipErrorInvalidValue;
    copy_spans_kernel<<<dim3(count < groups ? count : groups), dim3(256), 0, s>>>(spans, count);
    return hipGetLastError();
}

}  // namespace ph
