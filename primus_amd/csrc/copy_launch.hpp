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
#ifndef PH_COPY_SPAN_MAX
#define PH_COPY_SPAN_MAX (256u << 10)
#endif
constexpr uint64_t kCopySpanMax = PH_COPY_SPAN_MAX;

// Workgroups of one span-copy launch (fewer when there are fewer spans).
#ifndef PH_COPY_GROUPS
#define PH_COPY_GROUPS 64
#endif
constexpr uint32_t kCopyGroups = PH_COPY_GROUPS;

// spans: `count` spans in memory the device can read (page-locked host memory).
hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, hipStream_t s);

}  // namespace ph
