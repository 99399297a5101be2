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
// Launch shape of the single-pass kernels (N <= 2^13): log2 of the
// coefficients per thread (3 or 4) in the low bits, kShapeStream set for the
// persistent streaming kernel (forward-first operations, 2^10 <= N <= 2^13).
// 0 is the default: 16 coefficients per thread, one workgroup per polynomial.
constexpr int kShapeStream = 16;
inline bool valid_shape(int shape) { return shape == 0 || shape == 3 || shape == 4 || shape == 19 || shape == 20; }

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
