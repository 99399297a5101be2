// ntt_rounds.hpp -- register-round schedule of the chunk kernels: geometry,
// slot numbering and the forward/inverse round sequences with their LDS
// exchanges (see ntt_device.hpp for the butterflies and the arithmetic).
#pragma once

#include "ntt_device.hpp"

// PH_RANGE_CHECK=1 builds (a debug variant of the library) check every value
// at the register rounds' boundaries against its arithmetic's contract
// (A::in_fwd_range / in_inv_range / in_out_range) and count violations in a
// device array of the translation unit: the kernels neither trap nor change
// their results.  Default builds compile the checks out.
#ifndef PH_RANGE_CHECK
#define PH_RANGE_CHECK 0
#endif

namespace ph {

#if PH_RANGE_CHECK
// {violations, first site, first value's bits, its modulus}; one copy per
// translation unit (no relocatable device code), read by its range_report_*.
static __device__ unsigned long long g_range_report[4];
#endif

// Check sites.
enum RangeSite : uint32_t {
    kRangeFwdIn = 1,    // values entering forward rounds
    kRangeFwdRound = 2, // values leaving a forward round
    kRangeInvIn = 3,    // values entering inverse rounds
    kRangeInvRound = 4, // values leaving an inverse round
    kRangeInvOut = 5,   // inverse outputs
};

template <class A>
__device__ __forceinline__ void range_check(const typename A::T *x, int count, uint32_t site, const DevTable &t) {
#if PH_RANGE_CHECK
    for (int j = 0; j < count; ++j) {
        const bool ok = site == kRangeFwdIn || site == kRangeFwdRound ? A::in_fwd_range(x[j], t)
                        : site == kRangeInvOut                       ? A::in_out_range(x[j], t)
                                                                     : A::in_inv_range(x[j], t);
        if (!ok) {
            const unsigned long long prev = atomicAdd(&g_range_report[0], 1ull);
            if (prev == 0) {
                uint64_t bits;
                __builtin_memcpy(&bits, &x[j], sizeof(bits));
                g_range_report[1] = site;
                g_range_report[2] = bits;
                g_range_report[3] = t.q;
            }
        }
    }
#else
    (void)x, (void)count, (void)site, (void)t;
#endif
}

template <int LOGC, int LOGT>
struct Geo {
    static constexpr int C = 1 << LOGC;
    static constexpr int T = 1 << LOGT;
    static constexpr int LOGE = LOGC - LOGT;
    static constexpr int E = 1 << LOGE;
    static constexpr int REM = LOGC % LOGE;
    static constexpr int NR = LOGC / LOGE + (REM ? 1 : 0);
    static constexpr int LDS = C + (C >> 4);
    // forward rounds: a short round first (if any), then full rounds
    static constexpr int fr(int i) { return (REM != 0 && i == 0) ? REM : LOGE; }
    static constexpr int fs(int i) { return i == 0 ? 0 : (REM != 0 ? REM + (i - 1) * LOGE : i * LOGE); }
    // inverse rounds: full rounds, then the short one
    static constexpr int ir(int i) { return (REM != 0 && i == NR - 1) ? REM : LOGE; }
    static constexpr int is(int i) { return i * LOGE; }
    // First inverse round whose twiddle indices are equal on all lanes of a
    // wave (gaps of 64 and up, or of a whole smaller workgroup): it and the
    // rounds after it read their twiddles with scalar loads.  NR if none.
    static constexpr int first_scalar_inv() {
        for (int i = 0; i < NR; ++i)
            if (is(i) >= (LOGT < 6 ? LOGT : 6)) return i;
        return NR;
    }
    static_assert(LOGE >= 1, "a thread needs at least two coefficients");
};

// Chunk index of a round whose groups span 2^R slots at stride 2^B (B = OB
// for the forward, P for the inverse) as c = gi >> B.  With B = 3 and R = 3 a
// 16-lane quarter of a wave covers two chunks of 8 consecutive slots 64 slots
// apart, and lds_pos puts the second run 68 words after the first: four of
// its 64-bit accesses then share banks with the first run's (32 four-byte
// banks, 16 lanes per cycle).  Swapping bits 0 and 1 of c pairs chunks two
// apart instead (136 words, a disjoint half of the banks).  Bits 0 and 1 of c
// are bits B and B+1 of the thread index (LOGT >= B + 2), so the slot stays a
// thread part plus a register part and every round's map stays a bijection.
template <int LOGT, int B, int R>
__host__ __device__ __forceinline__ constexpr uint32_t chunk_perm(uint32_t c) {
    if constexpr (B == 3 && R == 3 && LOGT >= B + 2) return (c & ~3u) | ((c & 1u) << 1) | ((c >> 1) & 1u);
    else return c;
}

// Local slot of register (m, j) in the forward round covering local stages [S, S+R).
template <int LOGC, int LOGT, int S, int R>
__host__ __device__ __forceinline__ constexpr uint32_t fwd_slot(uint32_t t, int m, int j) {
    constexpr int OB = LOGC - S - R;
    const uint32_t gi = t + ((uint32_t)m << LOGT);
    const uint32_t c = chunk_perm<LOGT, OB, R>(gi >> OB);
    const uint32_t o = gi & ((1u << OB) - 1u);
    return (c << (LOGC - S)) + ((uint32_t)j << OB) + o;
}

// Local slot of register (m, j) in the inverse round covering local stages [P, P+R).
template <int LOGC, int LOGT, int P, int R>
__host__ __device__ __forceinline__ constexpr uint32_t inv_slot(uint32_t t, int m, int j) {
    const uint32_t gi = t + ((uint32_t)m << LOGT);
    const uint32_t c = chunk_perm<LOGT, P, R>(gi >> P);
    const uint32_t o = gi & ((1u << P) - 1u);
    return (c << (P + R)) + ((uint32_t)j << P) + o;
}

// True when every wave of the workgroup covers the same LDS slots with the
// layout `slot_a` as with `slot_b` (both bijections of (thread, m, j) onto
// the chunk): each wave's slots under both are then one aligned block of
// 64 * E, checked by its extent.  An exchange between two such layouts only
// moves data inside each wave.
template <int LOGC, int LOGT, class FA, class FB>
constexpr bool same_wave_blocks(FA slot_a, FB slot_b) {
    constexpr int T = 1 << LOGT, E = 1 << (LOGC - LOGT);
    if (T <= 64) return true;
    for (int w = 0; w < T / 64; ++w) {
        uint32_t lo_a = ~0u, hi_a = 0, lo_b = ~0u, hi_b = 0;
        for (int lane = 0; lane < 64; ++lane)
            for (int i = 0; i < E; ++i) {
                const uint32_t t = (uint32_t)(w * 64 + lane);
                const uint32_t a = slot_a(t, i), b = slot_b(t, i);
                lo_a = a < lo_a ? a : lo_a;
                hi_a = a > hi_a ? a : hi_a;
                lo_b = b < lo_b ? b : lo_b;
                hi_b = b > hi_b ? b : hi_b;
            }
        if (lo_a != lo_b || hi_a - lo_a + 1 != 64u * E || hi_b - lo_b + 1 != 64u * E) return false;
    }
    return true;
}

// Orders one wave's LDS stores before its later LDS loads.  A wave's LDS
// instructions execute in issue order, so only the compiler has to be kept
// from moving a load above a store of another lane's slot.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barriers of one LDS exchange (store in the previous round's
// layout, load in this round's).  Before the store: only when the exchange is
// the first of a transform and the LDS was used before it (SYNC0); a later
// exchange stores exactly the slots the same thread loaded in the exchange
// before it, which no other thread reads.  Between store and load: a wave
// barrier when the exchange stays inside each wave, else a workgroup barrier.
template <bool PRE, bool LOCAL>
struct Exchange {
    __device__ __forceinline__ static void before_store() {
        if constexpr (PRE) __syncthreads();
    }
    __device__ __forceinline__ static void before_load() {
        if constexpr (LOCAL) wave_lds_sync();
        else __syncthreads();
    }
};

// FENCE: schedule each round stage by stage (ct_round); off for kernels that
// keep two transforms in flight and need the scheduler's freedom.  PERM: read
// the register-order twiddle tables of this launch shape (single pass only).
// SYNC0: the LDS may hold another transform's data that other threads still
// read when this one starts (see Exchange).
template <class A, int LOGC, int LOGT, int I, bool FENCE = true, bool PERM = false, bool SYNC0 = true>
struct FwdRounds {
    using G = Geo<LOGC, LOGT>;
    using T = typename A::T;
    static constexpr int S = G::fs(I), R = G::fr(I);
    // the exchange into this round (I > 0) stays inside each wave
    static constexpr bool kLocal =
        I == 0 ? true
               : same_wave_blocks<LOGC, LOGT>(
                     [](uint32_t t, int i) {
                         constexpr int R0 = G::fr(I == 0 ? 0 : I - 1), S0 = G::fs(I == 0 ? 0 : I - 1);
                         return fwd_slot<LOGC, LOGT, S0, R0>(t, i >> R0, i & ((1 << R0) - 1));
                     },
                     [](uint32_t t, int i) { return fwd_slot<LOGC, LOGT, S, R>(t, i >> R, i & ((1 << R) - 1)); });

    // LDS exchanges.  A slot is its thread part plus its register part, and
    // the two never carry into each other's low four bits (the thread part's
    // low bits are t mod 2^OB, the register part's a multiple of 2^OB below
    // 16), so lds_pos of the sum is the sum of the lds_pos: one thread base
    // per round and a constant offset per access.
    __device__ __forceinline__ static void store(T *lds, const T *x, uint32_t t) {
        T *base = lds + lds_pos(fwd_slot<LOGC, LOGT, S, R>(t, 0, 0));
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) base[lds_pos(fwd_slot<LOGC, LOGT, S, R>(0, m, j))] = x[(m << R) + j];
    }
    __device__ __forceinline__ static void load(const T *lds, T *x, uint32_t t) {
        const T *base = lds + lds_pos(fwd_slot<LOGC, LOGT, S, R>(t, 0, 0));
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) x[(m << R) + j] = base[lds_pos(fwd_slot<LOGC, LOGT, S, R>(0, m, j))];
    }
    // Global accesses split each slot into its thread part, fwd_slot(t, 0, 0),
    // and its register part, fwd_slot(0, m, j), a constant (the slot is their
    // sum): every access is then a uniform base plus one 32-bit lane offset.
    // g: the chunk's first word, the same for the whole workgroup (RowBuf).
    __device__ __forceinline__ static void load_global(const uint64_t *__restrict__ g, T *x, uint32_t t,
                                                       const DevTable &tab) {
        const RowBuf rb(g, G::C);
        const uint32_t lt = fwd_slot<LOGC, LOGT, S, R>(t, 0, 0);
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j)
                x[(m << R) + j] = A::load_fwd(rb.ld(lt, fwd_slot<LOGC, LOGT, S, R>(0, m, j)), tab);
    }
    // The words load_global reads, unconverted, in the same register order
    // (converted with A::load_fwd before the rounds).
    __device__ __forceinline__ static void load_raw(const uint64_t *__restrict__ g, uint64_t *r, uint32_t t) {
        const RowBuf rb(g, G::C);
        const uint32_t lt = fwd_slot<LOGC, LOGT, S, R>(t, 0, 0);
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) r[(m << R) + j] = rb.ld(lt, fwd_slot<LOGC, LOGT, S, R>(0, m, j));
    }

    // Runs rounds I .. NR-1.  st0: global stage of local stage 0; b: chunk index
    // of this workgroup inside the polynomial.
    __device__ __forceinline__ static void run(T *x, T *lds, uint32_t t, uint32_t st0, uint32_t b,
                                               const DevTable &tab) {
        if constexpr (I > 0) {
            using X = Exchange<I == 1 && SYNC0, kLocal>;
            X::before_store();
            FwdRounds<A, LOGC, LOGT, I - 1, FENCE, PERM, SYNC0>::store(lds, x, t);
            X::before_load();
            load(lds, x, t);
        } else {
            range_check<A>(x, G::E, kRangeFwdIn, tab);
        }
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m) {
            constexpr int OB = LOGC - S - R;
            uint32_t cl;
            if constexpr (OB >= LOGT) cl = (uint32_t)m >> (OB - LOGT);  // workgroup-uniform
            else cl = chunk_perm<LOGT, OB, R>((t + ((uint32_t)m << LOGT)) >> OB);
            // a wave's 64 lanes share cl: scalar twiddle loads
            if constexpr (OB >= 6 && OB < LOGT) cl = __builtin_amdgcn_readfirstlane(cl);
            constexpr bool LANE = OB < 6 && OB < LOGT;  // cl differs between lanes
            if constexpr (PERM)
                ct_round<A, R, FENCE, true, LANE>(x + (m << R), st0 + S, (b << S) + cl, A::template fwd_reg<G::LOGE>(tab),
                                                  tab);
            else
                ct_round<A, R, FENCE, false, LANE>(x + (m << R), st0 + S, (b << S) + cl, A::fwd(tab), tab);
        }
        range_check<A>(x, G::E, kRangeFwdRound, tab);
        if constexpr (I + 1 < G::NR) FwdRounds<A, LOGC, LOGT, I + 1, FENCE, PERM, SYNC0>::run(x, lds, t, st0, b, tab);
    }
};

// The default hook of InvRounds::run: nothing.
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// SYNC0 as in FwdRounds; SCALED: the values carry N^-1 already (gs_round).
template <class A, int LOGC, int LOGT, int I, bool LAST, bool PERM = false, bool SYNC0 = true, bool SCALED = false>
struct InvRounds {
    using G = Geo<LOGC, LOGT>;
    using T = typename A::T;
    static constexpr int P = G::is(I), R = G::ir(I);
    static constexpr bool kLocal =
        I == 0 ? true
               : same_wave_blocks<LOGC, LOGT>(
                     [](uint32_t t, int i) {
                         constexpr int R0 = G::ir(I == 0 ? 0 : I - 1), P0 = G::is(I == 0 ? 0 : I - 1);
                         return inv_slot<LOGC, LOGT, P0, R0>(t, i >> R0, i & ((1 << R0) - 1));
                     },
                     [](uint32_t t, int i) { return inv_slot<LOGC, LOGT, P, R>(t, i >> R, i & ((1 << R) - 1)); });

    // thread base plus constant offsets, as in FwdRounds::store
    __device__ __forceinline__ static void store(T *lds, const T *x, uint32_t t) {
        T *base = lds + lds_pos(inv_slot<LOGC, LOGT, P, R>(t, 0, 0));
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) base[lds_pos(inv_slot<LOGC, LOGT, P, R>(0, m, j))] = x[(m << R) + j];
    }
    __device__ __forceinline__ static void load(const T *lds, T *x, uint32_t t) {
        const T *base = lds + lds_pos(inv_slot<LOGC, LOGT, P, R>(t, 0, 0));
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) x[(m << R) + j] = base[lds_pos(inv_slot<LOGC, LOGT, P, R>(0, m, j))];
    }
    // thread part plus constant register part, as FwdRounds::load_global
    // g: the chunk's first word, the same for the whole workgroup (RowBuf).
    __device__ __forceinline__ static void store_global(uint64_t *__restrict__ g, const T *x, uint32_t t,
                                                        const DevTable &tab) {
        const RowBuf rb(g, G::C);
        const uint32_t lt = inv_slot<LOGC, LOGT, P, R>(t, 0, 0);
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m)
#pragma unroll
            for (int j = 0; j < (1 << R); ++j)
                rb.st(lt, inv_slot<LOGC, LOGT, P, R>(0, m, j), A::out_inv(x[(m << R) + j], tab));
    }

    // Runs rounds I .. NR-1.  hook() is called once, as round H starts
    // (H: the first round with scalar twiddle loads, else the last round):
    // vector loads it issues are then not waited for by later twiddle loads.
    template <class H = NoHook>
    __device__ __forceinline__ static void run(T *x, T *lds, uint32_t t, uint32_t n, uint32_t b,
                                               const DevTable &tab, const H &hook = H()) {
        if constexpr (I > 0) {
            using X = Exchange<I == 1 && SYNC0, kLocal>;
            X::before_store();
            InvRounds<A, LOGC, LOGT, I - 1, LAST, PERM, SYNC0, SCALED>::store(lds, x, t);
            X::before_load();
            load(lds, x, t);
        } else {
            range_check<A>(x, G::E, kRangeInvIn, tab);
        }
        constexpr int HR = G::first_scalar_inv() < G::NR ? G::first_scalar_inv() : G::NR - 1;
        if constexpr (I == HR) {
            __builtin_amdgcn_sched_barrier(0);
            hook();
            __builtin_amdgcn_sched_barrier(0);
        }
        constexpr bool last_here = LAST && (I + 1 == G::NR);
#pragma unroll
        for (int m = 0; m < (G::E >> R); ++m) {
            uint32_t cl;
            if constexpr (P >= LOGT) cl = (uint32_t)m >> (P - LOGT);  // workgroup-uniform
            else cl = chunk_perm<LOGT, P, R>((t + ((uint32_t)m << LOGT)) >> P);
            if constexpr (P >= 6 && P < LOGT) cl = __builtin_amdgcn_readfirstlane(cl);  // as in FwdRounds
            constexpr bool LANE = P < 6 && P < LOGT;
            if constexpr (PERM)
                gs_round<A, R, last_here, true, true, LANE, SCALED>(x + (m << R), P, (b << (LOGC - P - R)) + cl, n,
                                                                    A::template inv_reg<G::LOGE>(tab), tab);
            else
                gs_round<A, R, last_here, true, false, LANE, SCALED>(x + (m << R), P, (b << (LOGC - P - R)) + cl,
                                                                     n, A::inv(tab), tab);
        }
        range_check<A>(x, G::E, last_here ? kRangeInvOut : kRangeInvRound, tab);
        if constexpr (I + 1 < G::NR)
            InvRounds<A, LOGC, LOGT, I + 1, LAST, PERM, SYNC0, SCALED>::run(x, lds, t, n, b, tab, hook);
    }
};

}  // namespace ph
