// copy_launch.hpp -- host-side view of the span-copy kernel (copy_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ph {

// One contiguous copy between device memory and page-locked host memory
// (hipHostMalloc: the same address on host and device).
struct CopySpan {
    uint64_t src, dst;  // addresses, both dereferenceable by the device
    uint64_t bytes;
};

// Largest span the host builds: longer copies are split so that the
// workgroups share them evenly (a span is one workgroup's unit of work).
constexpr uint64_t kCopySpanMax = 256u << 10;

// Workgroups of one span-copy launch (fewer when there are fewer spans):
// 64 keep one direction of the host link at its full rate (57 GB/s measured
// on MI355X, the same as a DMA copy) and both directions near 44 GB/s each.
constexpr uint32_t kCopyGroups = 64;

// spans: `count` spans in memory the device can read (page-locked host memory).
hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, hipStream_t s);

}  // namespace ph
