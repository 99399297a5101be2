// ntt_kernels.hip -- gfx950 kernels of the negacyclic NTT and the fused
// transform pipelines built on it.
//
// Two kernel shapes cover ring degrees 2^1 .. 2^17:
//
// * chunk_kernel<A, LOGC, LOGT, MODE>: one workgroup of 2^LOGT threads owns a
//   contiguous chunk of 2^LOGC coefficients in LDS (plus one pad word per 16)
//   and runs the chunk's stages as radix-2^LOGE register rounds, E = 2^LOGE
//   coefficients per thread.  For N <= 2^13 the chunk is the whole
//   polynomial (single pass, one HBM round trip); the forward's last round and
//   the inverse's first round have the same register layout, so a fused
//   forward -> pointwise -> inverse needs no LDS exchange in the middle.
// * column_kernel<LOGR, MODE>: for N >= 2^14 the first log2(N/4096) forward
//   stages (the last ones of the inverse) act on columns k = i*4096 + j;
//   each thread runs one column's 2^LOGR coefficients in registers, and the
//   4096-wide rows of a wave are read and written coalesced.  Together with a
//   chunk pass of 4096 this is the two-pass (four-step) schedule.
//
// A is the arithmetic policy (IntArith or F64Arith, ntt_device.hpp).  MODE
// bits select what a chunk does between its load and its store.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "ntt_device.hpp"
#include "ntt_launch.hpp"
#include "ntt_rounds.hpp"

namespace ph {

enum : int {
    M_FWD = 1,        // forward stages of this chunk
    M_INV = 2,        // inverse stages of this chunk
    M_MUL_SHOUP = 4,  // multiply by a resident Shoup operand between them
    M_MUL_POLY = 8,   // multiply by the transform of a second polynomial
    M_CANON = 16,     // forward output reduced to [0, q)
    M_INV_LAST = 32,  // the chunk holds the inverse's final stage (single pass)
};

__device__ __forceinline__ const Shoup *operand_of(const KArgs &a, const IntArith *, int) { return a.operand; }
__device__ __forceinline__ const Shoup *operand_of(const KArgs &a, const IntCanonArith *, int) { return a.operand; }
__device__ __forceinline__ const ShoupF *operand_of(const KArgs &a, const F64Arith *, int loge) {
    return a.operand_f[loge - 3];
}

// Waves per SIMD the launch bounds ask the register allocator for: as many as
// the LDS footprint lets resident, at most 4 (128 VGPRs) with 16 coefficients
// per thread and at most 8 (64 VGPRs) with 8 or fewer, halved when a second
// operand's transform is held in registers too.  Coefficient rows go through
// buffer resources (RowBuf), so no 64-bit address is kept per access and the
// radix-8 rounds fit 64 VGPRs without spilling.
template <int LOGC, int LOGT, int MODE>
struct Occupancy {
    static constexpr int lds_bytes = (Geo<LOGC, LOGT>::LDS) * 8;
    static constexpr int wgs_lds = (160 * 1024) / lds_bytes;
    static constexpr int wgs = wgs_lds < 1 ? 1 : (wgs_lds > 32 ? 32 : wgs_lds);
    static constexpr int waves_wg = (1 << LOGT) < 64 ? 1 : (1 << LOGT) / 64;
    static constexpr int per_simd_raw = (wgs * waves_wg) / 4;
    static constexpr int cap = (LOGC - LOGT) <= 3 ? 8 : 4;
    static constexpr int per_simd = per_simd_raw < 1 ? 1 : (per_simd_raw > cap ? cap : per_simd_raw);
    static constexpr int value = (MODE & M_MUL_POLY) ? (per_simd / 2 < 1 ? 1 : per_simd / 2) : per_simd;
};

// What a chunk does after its forward rounds: the optional multiply by the
// resident operand, then either the forward's store or the inverse rounds
// and their store.  After the last forward round thread t holds local slots
// t*E .. t*E+E-1.  hook() runs once on the way (InvRounds::run), or first
// when there are no inverse rounds.
template <class A, int LOGC, int LOGT, int MODE, class H = NoHook>
__device__ __forceinline__ void after_forward(const KArgs &a, typename A::T *x, typename A::T *lds, uint32_t t,
                                              uint32_t b, uint32_t limb, uint32_t n, const DevTable &tab,
                                              uint64_t *data, const H &hook = H()) {
    using G = Geo<LOGC, LOGT>;
    constexpr int LE = G::LOGE;
    // Keep the scheduler from hoisting the next phase's loads into the last
    // forward round: their registers would be live across it and spill.
    __builtin_amdgcn_sched_barrier(0);
    if constexpr ((MODE & M_MUL_SHOUP) != 0) {
        if constexpr (A::kRegisterOrder) {
            // single pass: the operand is stored in register order, entry
            // (t, j) at j*T + t, so each of the E loads of a wave is one row
            const auto rs = table_rsrc(as_global(operand_of(a, (const A *)nullptr, LE)) + (uint64_t)limb * n);
            using W = typename A::Tw;
            // Every load issued before the first product: left to itself the
            // scheduler pairs each load with its product and waits for each.
            // The streaming kernel (a hook) has no registers to spare for it.
            if constexpr (std::is_same<H, NoHook>::value) {
                W w[G::E];
#pragma unroll
                for (int j = 0; j < G::E; ++j)
                    w[j] = ld_lane<W>(rs, t * (uint32_t)sizeof(W), j * G::T * (uint32_t)sizeof(W));
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < G::E; ++j) x[j] = A::mul_operand(x[j], w[j], tab);
            } else {
#pragma unroll
                for (int j = 0; j < G::E; ++j)
                    x[j] = A::mul_operand(x[j], ld_lane<W>(rs, t * (uint32_t)sizeof(W), j * G::T * (uint32_t)sizeof(W)),
                                          tab);
            }
        } else {
            const auto op = as_global(operand_of(a, (const A *)nullptr, LE)) + (uint64_t)limb * n +
                            (uint64_t)b * G::C + ((uint64_t)t << LE);
#pragma unroll
            for (int j = 0; j < G::E; ++j) x[j] = A::mul_operand(x[j], ld(op + j), tab);
        }
    }
    if constexpr ((MODE & M_INV) == 0) {
        hook();
        uint64_t *dst = data + ((uint64_t)t << LE);
#pragma unroll
        for (int j = 0; j < G::E; ++j) {
            if constexpr ((MODE & M_CANON) != 0) dst[j] = A::canon_fwd(x[j], tab);
            else dst[j] = (uint64_t)x[j];  // IntArith only: lazy [0,4q)
        }
    } else {
        constexpr bool LAST = (MODE & M_INV_LAST) != 0;
        // N^-1 is in the operand's binary64 copies: the last stage only reduces sums
        constexpr bool SCALED = (MODE & M_MUL_SHOUP) != 0 && LAST && A::kScaledOperand;
        __builtin_amdgcn_sched_barrier(0);
        // With a hook (the streaming kernel, which holds the next
        // polynomial's words through the last rounds) the inverse forms its
        // addresses afresh rather than keeping the forward's alive.
        uint32_t ti = t;
        if constexpr (!std::is_same<H, NoHook>::value) asm volatile("" : "+v"(ti));
        // the inverse's first exchange stores the slots each thread loaded
        // for the forward's last round: no barrier before it (SYNC0 false)
        InvRounds<A, LOGC, LOGT, 0, LAST, A::kRegisterOrder, false, SCALED>::run(x, lds, ti, n, b, tab, hook);
        InvRounds<A, LOGC, LOGT, G::NR - 1, LAST, A::kRegisterOrder>::store_global(data, x, ti, tab);
    }
}

// PH_CLOCK_PROBE=1 builds (a measurement variant of the library) log every
// wave of every chunk_kernel launch as it ends: the shader-clock counter
// (s_memtime), the constant 100 MHz counter (s_memrealtime), its placement
// (HW_ID, XCC_ID) and its workgroup.  Successive launches' entries of one
// hardware wave slot give the engine clock the kernel ran at (the ratio of
// the two counters' advances times 100 MHz), and one launch's entries the
// spread of its workgroups' end times (clock_probe_report).
// * Only the end is stamped, after the wave's last load: a counter read or a
//   store earlier in the kernel counts as a possible write to the tables it
//   reads afterwards, which turns the scalar twiddle loads into vector loads
//   and changes the kernel being measured.
// * Each hardware wave slot (XCC, shader engine, CU, SIMD, wave slot) keeps a
//   ring of its last kClockProbeDepth entries and a count.  The waves of one
//   slot run one after another, so a plain load and store of the count
//   suffice; one counter shared by every wave would serialise their ends.
// Default builds compile the log out.
#ifndef PH_CLOCK_PROBE
#define PH_CLOCK_PROBE 0
#endif
#if PH_CLOCK_PROBE
// entry: {shader clock, 100 MHz clock, HW_ID | XCC_ID << 32, blockIdx.x}
static __device__ unsigned long long g_clock_log[(uint64_t)kClockProbeSlots * kClockProbeDepth * 4];
static __device__ unsigned int g_clock_count[kClockProbeSlots];

__device__ __forceinline__ void clock_probe_end() {
    const uint64_t clk = clock64(), wall = wall_clock64();
    const uint32_t hw = (uint32_t)__builtin_amdgcn_s_getreg(4 | (31 << 11));    // HW_ID
    const uint32_t xcc = (uint32_t)__builtin_amdgcn_s_getreg(20 | (31 << 11));  // XCC_ID
    const uint32_t wave = hw & 15u, simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u, se = (hw >> 13) & 7u;
    const uint32_t slot = ((((xcc & 7u) * 8u + se) * 16u + cu) * 4u + simd) * 16u + wave;
    if (__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)) == 0) {  // the wave's first active lane
        const uint32_t c = g_clock_count[slot];
        g_clock_count[slot] = c + 1;
        unsigned long long *p = g_clock_log + ((uint64_t)slot * kClockProbeDepth + c % kClockProbeDepth) * 4u;
        p[0] = clk;
        p[1] = wall;
        p[2] = hw | ((uint64_t)xcc << 32);
        p[3] = blockIdx.x;
    }
}
#define PH_PROBE_END() clock_probe_end()
#else
#define PH_PROBE_END() ((void)0)
#endif

template <class A, int LOGC, int LOGT, int MODE>
__global__ __launch_bounds__(1 << LOGT, (Occupancy<LOGC, LOGT, MODE>::value))
__attribute__((amdgpu_waves_per_eu(1, Occupancy<LOGC, LOGT, MODE>::value))) void chunk_kernel(KArgs a) {
    using G = Geo<LOGC, LOGT>;
    using T = typename A::T;
    constexpr int LE = G::LOGE;
    __shared__ T lds[G::LDS];
    const uint32_t t = threadIdx.x;
    const uint64_t blk = blockIdx.x;
    const uint64_t poly = blk / a.blocks_per_poly;
    const uint32_t b = (uint32_t)(blk - poly * a.blocks_per_poly);
    uint32_t limb;
    if constexpr ((MODE & M_MUL_POLY) != 0) {
        limb = a.limb_of ? a.limb_of[poly] : (uint32_t)(poly % a.L);
        if (limb >= a.L) {  // a row with no table is left as it is (whole workgroup)
            PH_PROBE_END();
            return;
        }
    } else {
        limb = (uint32_t)(poly % a.L);
    }
    const DevTable tab = a.tabs[limb];
    const uint32_t n = 1u << tab.log_n;
    const uint32_t st0 = tab.log_n - LOGC;
    const uint64_t off = poly * n + (uint64_t)b * G::C;
    uint64_t *data = a.data + off;
    T x[G::E];

    if constexpr ((MODE & M_FWD) != 0) {
        if constexpr ((MODE & M_MUL_POLY) != 0) {
            // transform the second operand first and keep it in registers
            T y[G::E];
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::load_global(a.data2 + off, y, t, tab);
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder, false>::run(y, lds, t, st0, b, tab);
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::load_global(data, x, t, tab);
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder, true>::run(x, lds, t, st0, b, tab);
            // NttPolynomial::mul_assign (primus_poly/src/ntt/mul.rs:87-95)
#pragma unroll
            for (int j = 0; j < G::E; ++j) x[j] = A::mul_pair(x[j], y[j], tab);
        } else {
            // the first use of the LDS in this workgroup: no barrier before its first exchange
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::load_global(data, x, t, tab);
            FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder, false>::run(x, lds, t, st0, b, tab);
        }
        after_forward<A, LOGC, LOGT, MODE>(a, x, lds, t, b, limb, n, tab, data);
    } else {
        const uint64_t *src = data + ((uint64_t)t << LE);
#pragma unroll
        for (int j = 0; j < G::E; ++j) x[j] = A::load(src[j], tab);
        constexpr bool LAST = (MODE & M_INV_LAST) != 0;
        __builtin_amdgcn_sched_barrier(0);
        InvRounds<A, LOGC, LOGT, 0, LAST, A::kRegisterOrder, false>::run(x, lds, t, n, b, tab);
        InvRounds<A, LOGC, LOGT, G::NR - 1, LAST, A::kRegisterOrder>::store_global(data, x, t, tab);
    }
    PH_PROBE_END();
}

// Waves per SIMD of the streaming kernel: as the chunk kernel's, but two at
// 16 coefficients per thread, where the next polynomial's coefficients beside
// the current one's spill at 128 VGPRs.
template <int LOGC, int LOGT, int MODE>
struct StreamOccupancy {
    static constexpr int base0 = Occupancy<LOGC, LOGT, MODE>::value;
    static constexpr int base = base0 > 4 ? 4 : base0;  // the prefetched words need the registers
    static constexpr int value = (LOGC - LOGT) >= 4 ? (base > 2 ? 2 : base) : base;
};

// Persistent, software-pipelined form of the single-pass chunk kernel for
// modes that start with forward stages and have no second operand.  Each
// workgroup walks limb-polynomials blockIdx.x, +gridDim.x, ... (the grid is
// at most one resident generation, so every workgroup has at least one).
// The loads of its next polynomial are issued as the inverse reaches its
// rounds with scalar twiddle loads, so no twiddle wait of the current
// polynomial waits for them (vector-memory counters complete in order); they
// land while those rounds and the stores run, and the stores in turn drain
// during the next polynomial's first rounds.  Outputs are those of
// chunk_kernel with the same template arguments.
template <class A, int LOGC, int LOGT, int MODE>
__global__ __launch_bounds__(1 << LOGT, (StreamOccupancy<LOGC, LOGT, MODE>::value)) void stream_kernel(KArgs a) {
    static_assert((MODE & M_FWD) != 0 && (MODE & M_MUL_POLY) == 0, "streaming covers forward-first modes");
    using G = Geo<LOGC, LOGT>;
    using T = typename A::T;
    __shared__ T lds[G::LDS];
    const uint32_t t0 = threadIdx.x;
    uint64_t poly = blockIdx.x;
    uint64_t raw[G::E];
    FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::load_raw(a.data + poly * G::C, raw, t0);
    while (true) {
        // an opaque copy of the thread index: the addresses derived from it
        // are formed afresh in each iteration instead of held across the loop
        uint32_t t = t0;
        asm volatile("" : "+v"(t));
        const uint32_t limb = (uint32_t)(poly % a.L);
        const DevTable tab = a.tabs[limb];
        T x[G::E];
#pragma unroll
        for (int j = 0; j < G::E; ++j) x[j] = A::load_fwd(raw[j], tab);
        const uint64_t next = poly + gridDim.x;
        const bool more = next < a.count;  // uniform over the workgroup
        const auto prefetch = [&]() {
            if (more) FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::load_raw(a.data + next * G::C, raw, t);
        };
        FwdRounds<A, LOGC, LOGT, 0, true, A::kRegisterOrder>::run(x, lds, t, 0, 0, tab);
        after_forward<A, LOGC, LOGT, MODE>(a, x, lds, t, 0, limb, G::C, tab, a.data + poly * G::C, prefetch);
        if (!more) break;
        poly = next;
    }
}

// Column pass of the two-pass schedule (IntArith, or HexlLazyArith for HEXL's
// lazy inverse): forward stages [0, LOGR) or inverse stages [logN-LOGR, logN)
// on columns k = i*C + j, C = N >> LOGR.
template <int LOGR, int MODE, class CA = IntArith>
__global__ __launch_bounds__(256) void column_kernel(KArgs a) {
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint32_t logc = a.tabs[0].log_n - LOGR;
    const uint64_t poly = gid >> logc;
    const uint32_t j = (uint32_t)(gid & ((1u << logc) - 1u));
    const uint32_t limb = a.limb_of ? a.limb_of[poly] : (uint32_t)(poly % a.L);
    if (limb >= a.L) return;  // a row with no table is left as it is (no barriers here)
    const DevTable tab = a.tabs[limb];
    const uint32_t n = 1u << tab.log_n;
    const uint64_t off = poly * n + j;
    const uint64_t *src = a.data2 + off;
    uint64_t *dst = a.data + off;
    uint64_t x[1 << LOGR];
#pragma unroll
    for (int i = 0; i < (1 << LOGR); ++i) x[i] = src[(uint64_t)i << logc];
    if constexpr ((MODE & M_FWD) != 0) {
        ct_round<IntArith, LOGR>(x, 0, 0, IntArith::fwd(tab), tab);
    } else {
        gs_round<CA, LOGR, true>(x, logc, 0, n, CA::inv(tab), tab);
    }
#pragma unroll
    for (int i = 0; i < (1 << LOGR); ++i) dst[(uint64_t)i << logc] = x[i];
}

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------

// transform_monomial family, primus_ntt/src/ntt/primitive.rs:298-397.
// kind 0: coeff * X^degree, 1: X^degree, 2: -X^degree.
__global__ void monomial_kernel(const DevTable *tabs, uint64_t *out, uint64_t coeff,
                                uint64_t degree, int kind) {
    const DevTable t = tabs[0];
    const uint32_t n = 1u << t.log_n;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t mask = 2ull * n - 1ull;
    const uint64_t r = __brev(k) >> (32 - t.log_n);
    const uint64_t idx = ((2ull * r + 1ull) * degree) & mask;
    uint64_t v;
    if (kind == 1) {
        v = degree == 0 ? 1ull : t.ord[idx].w;
    } else if (kind == 2) {
        v = degree == 0 ? t.q - 1 : t.ord[idx ^ n].w;
    } else {
        if (coeff == 0) v = 0;
        else if (degree == 0) v = coeff;
        else if (coeff == 1) v = t.ord[idx].w;
        else if (coeff == t.q - 1) v = t.ord[idx ^ n].w;
        else v = shoup_full(t.ord[idx].w, t.ord[idx].wq, coeff, t.q);
    }
    out[k] = v;
}

// Source slot of output slot d under x -> x^k in bit-reversed NTT order:
// natural slot i = rev(d) holds the evaluation at psi^(2i+1); sigma_k(f) there
// is f at psi^(k(2i+1) mod 2N), natural slot (j-1)/2 (generate_ntt_permutation,
// primus_fhe_core/src/automorphism/glwe/dcrt.rs:43-61).  k is odd and < 2N.
__device__ __forceinline__ uint32_t auto_source(uint32_t d, uint32_t log_n, uint32_t k) {
    const uint32_t i = __builtin_bitreverse32(d) >> (32 - log_n);
    const uint32_t j = (uint32_t)(((uint64_t)k * (2u * i + 1u)) & ((2ull << log_n) - 1));
    return __builtin_bitreverse32((j - 1u) >> 1) >> (32 - log_n);
}

// out[d] = in[src(d)] on every limb-polynomial (ntt_poly_auto_inplace,
// dcrt.rs:97-112), rows of up to 2^12 words.  A row is staged in LDS, so HBM
// sees only whole-row reads and writes, 16 bytes per lane when both pointers
// allow it (WIDE), else 8.  The grid is persistent: a workgroup walks rows
// blockIdx.x, + gridDim.x, ..., and loads its next row into registers before
// it stores the current one, so its reads and writes overlap.  On 1024 rows of
// 2^12 words the wide form runs at 0.74 of the HBM peak with 512 workgroups
// (11.3 us; a plain 16-byte copy of the same bytes takes 11.9 us), against
// 0.58 for one row per workgroup with 8-byte accesses.  LOGN > 0 fixes the
// degree at compile time (its index arithmetic and bounds fold away: 12.2 ->
// 11.3 us at 2^12); LOGN = 0 takes it from log_n_arg.
// Needs 1 <= log_n <= 12 and 8 << log_n bytes of dynamic LDS.
template <bool WIDE, uint32_t LOGN>
__global__ __launch_bounds__(256) void automorphism_rows_kernel(const uint64_t *in, uint64_t *out,
                                                                uint32_t log_n_arg, uint32_t k, uint64_t rows) {
    static_assert(LOGN <= 12, "rows of up to 2^12 words");
    const uint32_t log_n = LOGN ? LOGN : log_n_arg;
    using V = std::conditional_t<WIDE, u64x2, uint64_t>;
    constexpr uint32_t kWords = WIDE ? 2 : 1;
    constexpr uint32_t kPer = (1u << 12) / kWords / 256;  // elements per thread at N = 2^12
    extern __shared__ uint64_t lds[];
    const uint32_t elems = (1u << log_n) / kWords;
    const uint32_t t = threadIdx.x;
    uint64_t row = blockIdx.x;
    if (row >= rows) return;
    V r[kPer];
    {
        const V *src = reinterpret_cast<const V *>(in + (row << log_n));
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i)
            if (i * 256 + t < elems) r[i] = __builtin_nontemporal_load(src + i * 256 + t);
    }
    for (;;) {
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i)
            if (i * 256 + t < elems) *reinterpret_cast<V *>(lds + kWords * (i * 256 + t)) = r[i];
        __syncthreads();
        const uint64_t next = row + gridDim.x;
        if (next < rows) {
            const V *src = reinterpret_cast<const V *>(in + (next << log_n));
#pragma unroll
            for (uint32_t i = 0; i < kPer; ++i)
                if (i * 256 + t < elems) r[i] = __builtin_nontemporal_load(src + i * 256 + t);
        }
        V *dst = reinterpret_cast<V *>(out + (row << log_n));
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            const uint32_t p = i * 256 + t;
            if (p < elems) {
                V v;
                if constexpr (WIDE) {
                    v.x = lds[auto_source(2 * p, log_n, k)];
                    v.y = lds[auto_source(2 * p + 1, log_n, k)];
                } else {
                    v = lds[auto_source(p, log_n, k)];
                }
                __builtin_nontemporal_store(v, dst + p);
            }
        }
        // every lane has read the row from LDS before the next one lands
        __syncthreads();
        if (next >= rows) break;
        row = next;
    }
}

// Rows longer than 2^12 words: 256 consecutive output slots per workgroup,
// gathered from global memory, where the row's lines stay in L2 between their
// reads.
__global__ __launch_bounds__(256) void automorphism_kernel(const uint64_t *in, uint64_t *out, uint32_t log_n,
                                                           uint32_t k, uint64_t rows) {
    const uint32_t n = 1u << log_n;
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= (rows << log_n)) return;
    const uint64_t row = g >> log_n;
    const uint32_t d = (uint32_t)(g & (n - 1));
    st_stream(out + g, in[(row << log_n) + auto_source(d, log_n, k)]);
}

// x <- c_i * x mod q_i by MultiplyFactor (primus_factor/src/multiply.rs:50-70)
// at BIT_SHIFT 32, 52 or 64, for limb-major [count][N] data, limb = poly % L:
// the quotient estimate is (x * factor) >> BIT_SHIFT, the result
// c*x - estimate*q (mod 2^64) in [0, 2q) (lazy) or [0, q).  Shift 64 is
// ShoupFactor's factor_mul_modulo (NttPolynomial::mul_factor_assign,
// primus_poly/src/ntt/mul.rs:63-68); shifts 32 and 52 are HEXL's factor forms
// (inputs and q below 2^BIT_SHIFT; at 32 the product is formed in 64 bits, as
// the reference does).
template <int BIT_SHIFT>
__global__ __launch_bounds__(256) void mul_factor_kernel(MulFactorArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= a.total) return;
    const uint32_t limb = (uint32_t)((i >> a.log_n) % a.L);
    const uint64_t c = a.factor[limb], bf = a.barrett[limb], q = a.q[limb];
    const uint64_t x = a.data[i];
    uint64_t hw;
    if constexpr (BIT_SHIFT == 32) hw = (bf * x) >> 32;
    else if constexpr (BIT_SHIFT == 52) hw = (mulhi64(bf, x) << 12) | ((bf * x) >> 52);
    else hw = mulhi64(bf, x);
    uint64_t r = c * x - q * hw;
    if (!a.lazy) r = sub_if_ge(r, q);
    a.data[i] = r;
}

hipError_t launch_mul_factor(const MulFactorArgs &a, uint32_t bit_shift, hipStream_t s) {
    if (a.total == 0) return hipSuccess;
    if (!grid_fits((a.total + 255) / 256, 256)) return hipErrorInvalidConfiguration;
    const dim3 grid((uint32_t)((a.total + 255) / 256)), block(256);
    switch (bit_shift) {
        case 32: mul_factor_kernel<32><<<grid, block, 0, s>>>(a); break;
        case 52: mul_factor_kernel<52><<<grid, block, 0, s>>>(a); break;
        case 64: mul_factor_kernel<64><<<grid, block, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// out = a * b (mod q) or out = acc + a * b (mod q), limb-major [count][N],
// limb = poly % L.  primus_poly/src/ntt/mul.rs:87-119, primus_poly/src/ntt/mod.rs:97-108.
__global__ void pointwise_kernel(const DevTable *tabs, uint32_t L, uint32_t log_n, uint64_t total,
                                 const uint64_t *a, const uint64_t *b, uint64_t *out, int mac) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t limb = (uint32_t)((i >> log_n) % L);
    const DevTable t = tabs[limb];
    uint64_t v = mac ? barrett_mul_add(a[i], b[i], out[i], t.q, t.br0, t.br1)
                     : barrett_mul(a[i], b[i], t.q, t.br0, t.br1);
    out[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <class A, int LOGC, int LOGT, int MODE>
static hipError_t launch_chunk(const KArgs &a, uint64_t count, hipStream_t s) {
    const uint64_t blocks = count * a.blocks_per_poly;
    if (blocks == 0) return hipSuccess;
    if (!grid_fits(blocks, 1u << LOGT)) return hipErrorInvalidConfiguration;
    chunk_kernel<A, LOGC, LOGT, MODE><<<dim3((uint32_t)blocks), dim3(1u << LOGT), 0, s>>>(a);
    return hipGetLastError();
}

// One resident generation of stream_kernel workgroups on the current device.
template <class A, int LOGC, int LOGT, int MODE>
static uint32_t stream_slots() {
    static std::atomic<uint32_t> cached{0};
    uint32_t v = cached.load(std::memory_order_relaxed);
    if (v) return v;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stream_kernel<A, LOGC, LOGT, MODE>, 1 << LOGT, 0) !=
            hipSuccess)
        return 0;
    v = (uint32_t)std::max(1, cus * std::max(per_cu, 1));
    cached.store(v, std::memory_order_relaxed);
    return v;
}

// lpb_log: 0 sizes the grid to one resident generation (or the batch, if
// smaller); k > 0 to ceil(count / 2^k) workgroups, each walking about 2^k
// limb-polynomials, at most one generation.
template <class A, int LOGC, int LOGT, int MODE>
static hipError_t launch_stream(const KArgs &a0, uint64_t count, int lpb_log, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const uint32_t slots = stream_slots<A, LOGC, LOGT, MODE>();
    if (slots == 0) return hipErrorInvalidDevice;
    KArgs a = a0;
    a.count = count;
    const uint64_t want = lpb_log > 0 ? (count + (1ull << lpb_log) - 1) >> lpb_log : count;
    const uint64_t grid = std::min<uint64_t>(want, slots);
    stream_kernel<A, LOGC, LOGT, MODE><<<dim3((uint32_t)grid), dim3(1u << LOGT), 0, s>>>(a);
    return hipGetLastError();
}

template <int LOGR, int MODE, class CA = IntArith>
static hipError_t launch_column(const KArgs &a, uint32_t log_n, uint64_t count, hipStream_t s) {
    const uint64_t threads = count << (log_n - LOGR);
    if (threads == 0) return hipSuccess;
    if (!grid_fits(threads / 256u, 256)) return hipErrorInvalidConfiguration;
    column_kernel<LOGR, MODE, CA><<<dim3((uint32_t)(threads / 256u)), dim3(256), 0, s>>>(a);
    return hipGetLastError();
}

// Chunk or streaming launch of one forward-first mode.
template <class A, int LOGN, int LOGT, int MODE>
static hipError_t launch_fwd_first(const KArgs &a, uint64_t count, int shape, hipStream_t s) {
    if constexpr (LOGN >= 10) {
        if (shape & kShapeStream) return launch_stream<A, LOGN, LOGT, MODE>(a, count, shape_lpb_log(shape), s);
    }
    return launch_chunk<A, LOGN, LOGT, MODE>(a, count, s);
}

// Single-pass launch for log_n = LOGN <= 13, E = 2^LOGE coefficients per
// thread, the mode chosen at run time; kShapeStream in `shape` selects the
// persistent streaming kernel where the mode allows it.
template <int LOGN, int LOGE>
static hipError_t single_pass_e(Op op, KArgs a, uint64_t count, int arith, int stream, hipStream_t s) {
    constexpr int LOGT = LOGN > LOGE ? LOGN - LOGE : 0;
    a.blocks_per_poly = 1;
    if constexpr (LOGN >= 6) {
        if (arith & kArithF64) {
            switch (op) {
                case Op::FwdLazy:
                case Op::InvHexlLazy: break;  // lazy values are defined by the integer butterflies
                case Op::FwdCanon:
                    return launch_fwd_first<F64Arith, LOGN, LOGT, M_FWD | M_CANON>(a, count, stream, s);
                case Op::Inv: return launch_chunk<F64Arith, LOGN, LOGT, M_INV | M_INV_LAST>(a, count, s);
                case Op::FwdMulInv:
                    return launch_fwd_first<F64Arith, LOGN, LOGT, M_FWD | M_MUL_SHOUP | M_INV | M_INV_LAST>(
                        a, count, stream, s);
                case Op::PolyMul:
                    return launch_chunk<F64Arith, LOGN, LOGT, M_FWD | M_MUL_POLY | M_INV | M_INV_LAST>(a, count, s);
            }
        }
    }
    if (arith & kArithCanonInt) {
        using C = IntCanonArith;
        switch (op) {
            case Op::FwdLazy:
            case Op::InvHexlLazy: break;  // lazy values are defined by IntArith's butterflies
            case Op::FwdCanon: return launch_fwd_first<C, LOGN, LOGT, M_FWD | M_CANON>(a, count, stream, s);
            case Op::Inv: return launch_chunk<C, LOGN, LOGT, M_INV | M_INV_LAST>(a, count, s);
            case Op::FwdMulInv:
                return launch_fwd_first<C, LOGN, LOGT, M_FWD | M_MUL_SHOUP | M_INV | M_INV_LAST>(a, count, stream, s);
            case Op::PolyMul: return launch_chunk<C, LOGN, LOGT, M_FWD | M_MUL_POLY | M_INV | M_INV_LAST>(a, count, s);
        }
    }
    switch (op) {
        case Op::FwdLazy: return launch_fwd_first<IntArith, LOGN, LOGT, M_FWD>(a, count, stream, s);
        case Op::FwdCanon: return launch_fwd_first<IntArith, LOGN, LOGT, M_FWD | M_CANON>(a, count, stream, s);
        case Op::Inv: return launch_chunk<IntArith, LOGN, LOGT, M_INV | M_INV_LAST>(a, count, s);
        case Op::InvHexlLazy: return launch_chunk<HexlLazyArith, LOGN, LOGT, M_INV | M_INV_LAST>(a, count, s);
        case Op::FwdMulInv:
            return launch_fwd_first<IntArith, LOGN, LOGT, M_FWD | M_MUL_SHOUP | M_INV | M_INV_LAST>(a, count,
                                                                                                  stream, s);
        case Op::PolyMul:
            return launch_chunk<IntArith, LOGN, LOGT, M_FWD | M_MUL_POLY | M_INV | M_INV_LAST>(a, count, s);
    }
    return hipErrorInvalidValue;
}

// Coefficients per thread: 16 by default; 8 (twice the waves, one more LDS
// round) for 2^10 <= N <= 2^13 when the tuning asks for it.
template <int LOGN>
static hipError_t single_pass(Op op, KArgs a, uint64_t count, int arith, int shape, hipStream_t s) {
    const int stream = shape & (kShapeStream | (3 << kShapeLpbShift));
    if constexpr (LOGN >= 10) {
        if (shape_loge(shape) == 3) return single_pass_e<LOGN, 3>(op, a, count, arith, stream, s);
    }
    return single_pass_e<LOGN, 4>(op, a, count, arith, stream, s);
}

// Two-pass schedule for log_n = LOGC + LOGR: columns of 2^LOGR, then chunks of
// 2^LOGC in LDS with 2^LOGE coefficients per thread.
template <int LOGC, int LOGR, int LOGE>
static hipError_t two_pass(Op op, KArgs a, uint64_t count, void *scratch, hipStream_t s) {
    constexpr int LOGT = LOGC - LOGE;
    using A = IntArith;
    const uint32_t log_n = LOGC + LOGR;
    a.blocks_per_poly = 1u << LOGR;
    hipError_t e = hipSuccess;
    KArgs col = a;
    switch (op) {
        case Op::FwdLazy:
        case Op::FwdCanon:
            col.data2 = a.data;
            e = launch_column<LOGR, M_FWD>(col, log_n, count, s);
            if (e != hipSuccess) return e;
            if (op == Op::FwdLazy) return launch_chunk<A, LOGC, LOGT, M_FWD>(a, count, s);
            return launch_chunk<A, LOGC, LOGT, M_FWD | M_CANON>(a, count, s);
        case Op::Inv:
            e = launch_chunk<A, LOGC, LOGT, M_INV>(a, count, s);
            if (e != hipSuccess) return e;
            col.data2 = a.data;
            return launch_column<LOGR, M_INV>(col, log_n, count, s);
        case Op::InvHexlLazy:
            e = launch_chunk<A, LOGC, LOGT, M_INV>(a, count, s);
            if (e != hipSuccess) return e;
            col.data2 = a.data;
            return launch_column<LOGR, M_INV, HexlLazyArith>(col, log_n, count, s);
        case Op::FwdMulInv:
            col.data2 = a.data;
            e = launch_column<LOGR, M_FWD>(col, log_n, count, s);
            if (e != hipSuccess) return e;
            e = launch_chunk<A, LOGC, LOGT, M_FWD | M_MUL_SHOUP | M_INV>(a, count, s);
            if (e != hipSuccess) return e;
            return launch_column<LOGR, M_INV>(col, log_n, count, s);
        case Op::PolyMul: {
            // the second operand's column pass goes to scratch, leaving the caller's b intact
            KArgs colb = a;
            colb.data = (uint64_t *)scratch;
            colb.data2 = a.data2;
            e = launch_column<LOGR, M_FWD>(colb, log_n, count, s);
            if (e != hipSuccess) return e;
            col.data2 = a.data;
            e = launch_column<LOGR, M_FWD>(col, log_n, count, s);
            if (e != hipSuccess) return e;
            KArgs ch = a;
            ch.data2 = (const uint64_t *)scratch;
            e = launch_chunk<A, LOGC, LOGT, M_FWD | M_MUL_POLY | M_INV>(ch, count, s);
            if (e != hipSuccess) return e;
            return launch_column<LOGR, M_INV>(col, log_n, count, s);
        }
    }
    return hipErrorInvalidValue;
}

// The two-pass launch of ring degree 2^LOGN in the tile and radix the shape
// names (valid_shape_for); any other shape runs the default tile of 4096.
template <int LOGN>
static hipError_t two_pass_shape(Op op, KArgs a, uint64_t count, void *scratch, int shape, hipStream_t s) {
    const int tile = valid_shape_for(shape, LOGN) ? shape_tile_log(shape) : 12;
    const bool e8 = shape_loge(shape) == 3;
    if constexpr (LOGN - 11 <= 5) {
        if (tile == 11) return e8 ? two_pass<11, LOGN - 11, 3>(op, a, count, scratch, s)
                                  : two_pass<11, LOGN - 11, 4>(op, a, count, scratch, s);
    }
    if (tile == 13) return e8 ? two_pass<13, LOGN - 13, 3>(op, a, count, scratch, s)
                              : two_pass<13, LOGN - 13, 4>(op, a, count, scratch, s);
    return e8 ? two_pass<12, LOGN - 12, 3>(op, a, count, scratch, s)
              : two_pass<12, LOGN - 12, 4>(op, a, count, scratch, s);
}

bool needs_scratch(Op op, uint32_t log_n) { return op == Op::PolyMul && log_n > kMaxSinglePassLogN; }

hipError_t launch_transform(Op op, const KArgs &a, uint32_t log_n, uint64_t count, void *scratch, int arith,
                            int shape, hipStream_t s) {
    switch (log_n) {
        case 1: return single_pass<1>(op, a, count, arith, shape, s);
        case 2: return single_pass<2>(op, a, count, arith, shape, s);
        case 3: return single_pass<3>(op, a, count, arith, shape, s);
        case 4: return single_pass<4>(op, a, count, arith, shape, s);
        case 5: return single_pass<5>(op, a, count, arith, shape, s);
        case 6: return single_pass<6>(op, a, count, arith, shape, s);
        case 7: return single_pass<7>(op, a, count, arith, shape, s);
        case 8: return single_pass<8>(op, a, count, arith, shape, s);
        case 9: return single_pass<9>(op, a, count, arith, shape, s);
        case 10: return single_pass<10>(op, a, count, arith, shape, s);
        case 11: return single_pass<11>(op, a, count, arith, shape, s);
        case 12: return single_pass<12>(op, a, count, arith, shape, s);
        case 13: return single_pass<13>(op, a, count, arith, shape, s);
        case 14: return two_pass_shape<14>(op, a, count, scratch, shape, s);
        case 15: return two_pass_shape<15>(op, a, count, scratch, shape, s);
        case 16: return two_pass_shape<16>(op, a, count, scratch, shape, s);
        case 17: return two_pass_shape<17>(op, a, count, scratch, shape, s);
    }
    return hipErrorInvalidValue;
}

int tuning_candidates(uint32_t log_n, Op op, int *out, int max) {
    int n = 0;
    const auto add = [&](int shape) {
        if (n < max) out[n] = shape;
        ++n;
    };
    if (log_n > kMaxSinglePassLogN) {
        // LDS tile x radix of the chunk pass
        for (int tile = 11; tile <= 13; ++tile)
            for (int e : {4, 3})
                if (valid_shape_for(e | (tile << kShapeTileShift), log_n)) add(e | (tile << kShapeTileShift));
    } else if (log_n >= 10) {
        const bool fwd_first = op == Op::FwdLazy || op == Op::FwdCanon || op == Op::FwdMulInv;
        for (int e : {4, 3}) {
            add(e);
            if (fwd_first)
                for (int lpb = 0; lpb <= 3; ++lpb) add(e | kShapeStream | (lpb << kShapeLpbShift));
        }
    } else {
        add(4);
    }
    return std::min(n, max);
}

hipError_t launch_monomial(const DevTable *tabs, uint32_t log_n, uint64_t *out, uint64_t coeff,
                           uint64_t degree, int kind, hipStream_t s) {
    const uint32_t n = 1u << log_n;
    const uint32_t tpb = n < 256 ? n : 256;
    monomial_kernel<<<dim3((n + tpb - 1) / tpb), dim3(tpb), 0, s>>>(tabs, out, coeff, degree, kind);
    return hipGetLastError();
}

hipError_t launch_automorphism(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t k, uint64_t rows,
                               hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (log_n == 0 || log_n > 17) return hipErrorInvalidValue;
    if (log_n <= 12) {
        // about 2^21 words of rows in flight at a time: 512 workgroups at
        // N = 2^12 (two per CU), more for shorter rows
        const uint64_t groups = std::min<uint64_t>(rows, 512ull << (12 - log_n));
        const dim3 grid((uint32_t)groups), block(256);
        const uint32_t lds = 8u << log_n;
        if ((((uintptr_t)in | (uintptr_t)out) & 15) == 0) {
            switch (log_n) {
            case 12: automorphism_rows_kernel<true, 12><<<grid, block, lds, s>>>(in, out, log_n, k, rows); break;
            case 11: automorphism_rows_kernel<true, 11><<<grid, block, lds, s>>>(in, out, log_n, k, rows); break;
            case 10: automorphism_rows_kernel<true, 10><<<grid, block, lds, s>>>(in, out, log_n, k, rows); break;
            default: automorphism_rows_kernel<true, 0><<<grid, block, lds, s>>>(in, out, log_n, k, rows);
            }
        } else {
            automorphism_rows_kernel<false, 0><<<grid, block, lds, s>>>(in, out, log_n, k, rows);
        }
        return hipGetLastError();
    }
    const uint64_t blocks = ((rows << log_n) + 255) / 256;
    if (!grid_fits(blocks, 256)) return hipErrorInvalidConfiguration;
    automorphism_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(in, out, log_n, k, rows);
    return hipGetLastError();
}

hipError_t launch_pointwise(const DevTable *tabs, uint32_t L, uint32_t log_n, uint64_t total,
                            const uint64_t *a, const uint64_t *b, uint64_t *out, int mac,
                            hipStream_t s) {
    if (total == 0) return hipSuccess;
    const uint64_t blocks = (total + 255) / 256;
    if (!grid_fits(blocks, 256)) return hipErrorInvalidConfiguration;
    pointwise_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(tabs, L, log_n, total, a, b, out, mac);
    return hipGetLastError();
}

// Range-check report of this translation unit's kernels (PH_RANGE_CHECK
// builds): {violations, first site, first value's bits, its modulus},
// cleared afterwards when asked.  Blocking copies on the null stream.
hipError_t range_report_ntt(uint64_t *out, bool clear) {
#if PH_RANGE_CHECK
    hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_range_report), 4 * sizeof(uint64_t));
    if (e == hipSuccess && clear) {
        const uint64_t zero[4] = {0, 0, 0, 0};
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_range_report), zero, sizeof(zero));
    }
    return e;
#else
    (void)out, (void)clear;
    return hipErrorNotSupported;
#endif
}

// Clock-probe log (PH_CLOCK_PROBE builds): every slot's ring of entries
// (kClockProbeSlots x kClockProbeDepth x 4 words) into out and its count of
// waves logged into counts (either may be null); zeroed afterwards when
// asked.  Blocking copies on the null stream.
hipError_t clock_probe_report(uint64_t *out, uint32_t *counts, bool clear) {
#if PH_CLOCK_PROBE
    hipError_t e = hipSuccess;
    if (out) e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_clock_log), sizeof(g_clock_log));
    if (e == hipSuccess && counts) e = hipMemcpyFromSymbol(counts, HIP_SYMBOL(g_clock_count), sizeof(g_clock_count));
    if (e == hipSuccess && clear) {
        static const unsigned int zero[kClockProbeSlots] = {};
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_clock_count), zero, sizeof(zero));
    }
    return e;
#else
    (void)out, (void)counts, (void)clear;
    return hipErrorNotSupported;
#endif
}

}  // namespace ph
