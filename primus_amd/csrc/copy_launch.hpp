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

// Largest span the host builds: longer copies are split so that many
// workgroups share them.
constexpr uint64_t kCopySpanMax = 1ull << 20;

// spans: `count` spans in memory the device can read (page-locked host memory).
hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, hipStream_t s);

}  // namespace ph
