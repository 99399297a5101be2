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

// Workgroups of one span-copy launch (fewer when there are fewer spans).
// Inputs: 32.  64 keep one direction of the host link at its full rate alone
// (57 GB/s measured on MI355X, the same as a DMA copy), but in the serving mix,
// where compute launches and the result copies run at the same time, 32 gave
// 20.4k requests/s saturated against 18.4k for 64 (16-24: 19.8-20.2k, 48:
// 18.7k, 128: 18.2k): fewer reads across the link outstanding, and less
// interference with the kernels of the other streams.  Results: 16, enough for
// the mix's result traffic (about 0.6 of its input bytes) with fewer stores
// outstanding while other streams compute (8 and 32 measured lower).
constexpr uint32_t kCopyGroupsIn = 32;
constexpr uint32_t kCopyGroupsOut = 16;

// spans: `count` spans in memory the device can read (page-locked host
// memory), copied by at most `groups` workgroups.
hipError_t launch_copy_spans(const CopySpan *spans, uint32_t count, uint32_t groups, hipStream_t s);

}  // namespace ph
