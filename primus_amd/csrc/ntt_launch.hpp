// ntt_launch.hpp -- host-side view of the NTT kernels (ntt_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntt_device.hpp"

namespace ph {

constexpr uint32_t kMinLogN = 1;
constexpr uint32_t kMaxLogN = 17;
constexpr uint32_t kMaxSinglePassLogN = 13;
constexpr uint64_t kF64MaxModulus = 1ull << 50;
constexpr uint64_t kCanonIntMaxModulus = 1ull << 60;

// Arithmetic a launch may use, as bits: kArithF64 (every modulus < 2^50:
// F64Arith for canonical-output operations), kArithCanonInt (every modulus
// < 2^60: IntCanonArith for canonical-output operations).  Lazy-output
// operations always run IntArith, whose lazy values are the reference's.
constexpr int kArithF64 = 2;
constexpr int kArithCanonInt = 1;

// Kernel arguments, passed by value.
struct KArgs {
    uint64_t *data;          // in/out limb-polynomials, [count][N]
    const uint64_t *data2;   // second operand (poly-multiply) / column-pass source
    const Shoup *operand;    // Shoup-prepared NTT-domain operand, [L][N]
    const ShoupF *operand_f[2]; // the same operand as centred binary64 pairs (q < 2^50), in the
                                // register order of 8 / 16 coefficients per thread
    const DevTable *tabs;    // L device tables; limb-polynomial i uses tabs[i % L]
    const uint8_t *limb_of;  // when set (poly-multiply only): limb-polynomial i uses
                             // tabs[limb_of[i]], for batches of uneven limb counts
    uint32_t L;
    uint32_t blocks_per_poly;
    uint64_t count;          // limb-polynomials of the launch (persistent kernels)
};

enum class Op : int {
    FwdLazy,    // lazy_transform_slice:          [0,4q) -> [0,4q)
    FwdCanon,   // transform_slice:               [0,4q) -> [0,q)
    Inv,        // (lazy_)inverse_transform_slice: [0,2q) -> [0,q)
    FwdMulInv,  // forward, Shoup multiply by a resident operand, inverse
    PolyMul,    // negacyclic product: forward both, Barrett multiply, inverse
    InvHexlLazy,  // HexlNttTable::compute_inverse(.., 2): [0,2q) -> [0,2q), HEXL's last stage
};
constexpr int kOpCount = 6;

bool needs_scratch(Op op, uint32_t log_n);
// Range-check reports (ntt_rounds.hpp, PH_RANGE_CHECK builds) of the kernels in
// ntt_kernels.hip and fhe_kernels.hip; hipErrorNotSupported in default builds.
hipError_t range_report_ntt(uint64_t *out, bool clear);
// Clock-probe log of chunk_kernel waves (ntt_kernels.hip, PH_CLOCK_PROBE
// builds): per hardware wave slot (8 XCCs x 8 shader engines x 16 CUs x 4
// SIMDs x 16 waves) a ring of its last kClockProbeDepth entries of 4 words
// and a count; hipErrorNotSupported in default builds.
constexpr uint32_t kClockProbeSlots = 8u * 8u * 16u * 4u * 16u;
constexpr uint32_t kClockProbeDepth = 32;
hipError_t clock_probe_report(uint64_t *out, uint32_t *counts, bool clear);
hipError_t range_report_fhe(uint64_t *out, bool clear);
// Launch shape of the transform kernels, as bit fields (0 everywhere = the
// default shape):
//   bits 0-3  log2 of the coefficients per thread of the chunk kernels, 3 or 4
//             (0: 4), i.e. radix-8 or radix-16 register rounds;
//   bit 4     kShapeStream: the persistent streaming kernel (forward-first
//             operations, 2^10 <= N <= 2^13);
//   bits 5-6  with kShapeStream, log2 of the limb-polynomials each workgroup
//             walks: 0 sizes the grid to one resident generation, k = 1..3 to
//             ceil(count / 2^k) workgroups (at most one generation);
//   bits 8-11 two-pass schedule (N >= 2^14): log2 of the chunk the chunk pass
//             holds in LDS, 11..13 (0: 12), with N / chunk <= 32.
constexpr int kShapeStream = 16;
constexpr int kShapeLpbShift = 5;
constexpr int kShapeTileShift = 8;
constexpr int kShapeMask = 0xF7F;
inline int shape_loge(int shape) { return (shape & 15) ? (shape & 15) : 4; }
inline int shape_lpb_log(int shape) { return (shape >> kShapeLpbShift) & 3; }
inline int shape_tile_log(int shape) { return ((shape >> kShapeTileShift) & 15) ? ((shape >> kShapeTileShift) & 15) : 12; }
inline bool valid_shape(int shape) {
    const int e = shape & 15, tile = (shape >> kShapeTileShift) & 15;
    if (shape < 0 || (shape & ~kShapeMask) != 0) return false;
    if (e != 0 && e != 3 && e != 4) return false;
    if (shape_lpb_log(shape) != 0 && !(shape & kShapeStream)) return false;
    return tile == 0 || (tile >= 11 && tile <= 13);
}
// A shape's fields that apply at ring degree 2^log_n are within range there:
// a two-pass chunk only for N >= 2^14, and at most 32 columns per chunk.
inline bool valid_shape_for(int shape, uint32_t log_n) {
    if (!valid_shape(shape)) return false;
    const int tile = (shape >> kShapeTileShift) & 15;
    if (tile == 0) return true;
    return log_n > kMaxSinglePassLogN && (int)log_n - tile >= 1 && (int)log_n - tile <= 5;
}
// Every shape the autotuner tries at ring degree 2^log_n for an operation;
// returns how many it wrote to out (at most max).
int tuning_candidates(uint32_t log_n, Op op, int *out, int max);

// arith: kArithF64 / kArithCanonInt bits.  F64Arith where it applies
// (kArithF64, canonical outputs, 2^6 <= N <= 2^13), else IntCanonArith for
// canonical single-pass outputs (kArithCanonInt), else IntArith.
hipError_t launch_transform(Op op, const KArgs &a, uint32_t log_n, uint64_t count, void *scratch, int arith,
                            int shape, hipStream_t s);
hipError_t launch_monomial(const DevTable *tabs, uint32_t log_n, uint64_t *out, uint64_t coeff,
                           uint64_t degree, int kind, hipStream_t s);
// NTT-domain automorphism x -> x^k (k odd, < 2N) of `rows` rows of 2^log_n words.
hipError_t launch_automorphism(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t k, uint64_t rows,
                               hipStream_t s);
// Scalar multiply of limb-major data by one factor per limb (MultiplyFactor).
constexpr uint32_t kMulFactorMaxLimbs = 64;
struct MulFactorArgs {
    uint64_t *data;
    uint64_t total;  // words: count << log_n
    uint64_t factor[kMulFactorMaxLimbs], barrett[kMulFactorMaxLimbs], q[kMulFactorMaxLimbs];
    uint32_t L, log_n, lazy, pad_;
};
hipError_t launch_mul_factor(const MulFactorArgs &a, uint32_t bit_shift, hipStream_t s);
hipError_t launch_pointwise(const DevTable *tabs, uint32_t L, uint32_t log_n, uint64_t total,
                            const uint64_t *a, const uint64_t *b, uint64_t *out, int mac,
                            hipStream_t s);

}  // namespace ph
