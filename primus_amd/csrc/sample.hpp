// sample.hpp -- counter-based sampling of RLWE encryption inputs on gfx950.
//
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Every sampled value is a
// function of (seed, request, purpose, element) alone, so any thread can
// draw any element in any order and the host oracle reproduces it exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ph {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// The 4 x 32-bit block of counter c under key (k0, k1).
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c[0];
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
}

// What each block is drawn for: counter (element, request, purpose, 0).
enum : uint32_t {
    kSampleSecret = 0,   // ternary s
    kSampleError = 1,    // rounded Gaussian e
    kSampleMessage = 2,  // m in [0, 256)
    kSampleA = 3,        // a_hat limb j uses kSampleA + j
};

// Discrete Gaussian by inversion of its cumulative table: e = -kGaussCut plus
// the number of thresholds at or below the 64-bit uniform u.
constexpr int kGaussCut = 19;  // 6 sigma at sigma = 3.2, rounded down
constexpr int kGaussThresholds = 2 * kGaussCut;

__host__ __device__ __forceinline__ int sample_ternary(uint64_t u) {
    // floor(u * 3 / 2^64) in {0, 1, 2}, minus one
    const uint64_t hi = (uint64_t)(((unsigned __int128)u * 3u) >> 64);
    return (int)hi - 1;
}

__host__ __device__ __forceinline__ int sample_gauss(uint64_t u, const uint64_t *thresholds) {
    int e = -kGaussCut;
#pragma unroll
    for (int i = 0; i < kGaussThresholds; ++i) e += u >= thresholds[i] ? 1 : 0;
    return e;
}

// floor(v * q / 2^128) for the 128-bit uniform v = (hi:lo): uniform in [0, q)
// up to q / 2^128.
__host__ __device__ __forceinline__ uint64_t sample_mod(uint64_t lo, uint64_t hi, uint64_t q) {
    const unsigned __int128 x = (unsigned __int128)hi * q;
    const uint64_t y = (uint64_t)(((unsigned __int128)lo * q) >> 64);
    const uint64_t xl = (uint64_t)x;
    const uint64_t s = xl + y;
    return (uint64_t)(x >> 64) + (s < xl ? 1u : 0u);
}

}  // namespace ph
