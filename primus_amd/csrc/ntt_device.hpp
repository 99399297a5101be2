// ntt_device.hpp -- register/LDS building blocks of the negacyclic NTT on gfx950.
//
// The transform is the reference's Harvey-lazy Cooley-Tukey forward
// (normal -> bit-reversed, primus_ntt/src/ntt/primitive.rs:211-232) and
// Gentleman-Sande inverse (bit-reversed -> normal, :246-286), butterfly for
// butterfly: every butterfly pairs the same two slots with the same twiddle
// and computes the same lazy value, so lazy outputs are bit-identical to the
// CPU table and canonical outputs follow.
//
// Only the schedule differs.  A thread holds a group of 2^R slots in VGPRs and
// runs R consecutive stages on them (radix-2^R); groups are re-formed between
// register rounds through LDS.  Slot numbering inside a round:
//
//   forward round, global stages [S, S+R):  k = (c << (logN-S)) + (j << (logN-S-R)) + o
//   inverse round, global stages [P, P+R):  k = (c << (P+R)) + (j << P) + o
//
// with j < 2^R the register index, c the chunk the group lives in at the
// round's first stage and o the offset inside it.  The twiddle of a
// butterfly is the one the CPU loop would fetch for that chunk:
//   forward stage s, chunk c_s:  root_powers[2^s + c_s]
//   inverse stage p, chunk c_p:  inv_root_powers[1 + N - (N >> p) + c_p]
//   (the inverse's last stage, gap N/2, folds N^-1 as in primitive.rs:270-285).
#pragma once

#include "arith.hpp"

namespace ph {

struct alignas(16) Shoup {
    uint64_t w;   // value
    uint64_t wq;  // floor(w * 2^64 / q)
};

// Per-modulus device view of one UintNttTable.
struct DevTable {
    uint64_t q, q2;
    uint64_t inv_n, inv_n_q;  // N^-1 (Shoup)
    uint64_t sr, sr_q;        // N^-1 * inv_root_powers[N-1] (Shoup)
    uint64_t br0, br1;        // Barrett ratio floor(2^128/q)
    const Shoup *fwd;         // root_powers, N entries (bit-reversed order)
    const Shoup *inv;         // inv_root_powers, N entries (scrambled order)
    const Shoup *ord;         // ordinal_root_powers, 2N entries
    uint32_t log_n;
    uint32_t pad_;
};

// Harvey forward butterfly, primitive.rs:225-228.
__device__ __forceinline__ void ct_bf(uint64_t &a, uint64_t &b, uint64_t w, uint64_t wq, uint64_t q,
                                      uint64_t q2) {
    uint64_t u = sub_if_ge(a, q2);
    uint64_t v = shoup_lazy(w, wq, b, q);
    a = u + v;
    b = u + q2 - v;
}

// Gentleman-Sande inverse butterfly, primitive.rs:262-265.
__device__ __forceinline__ void gs_bf(uint64_t &a, uint64_t &b, uint64_t w, uint64_t wq, uint64_t q,
                                      uint64_t q2) {
    uint64_t u = a, v = b;
    a = sub_if_ge(u + v, q2);
    b = shoup_lazy(w, wq, u + q2 - v, q);
}

// Last inverse stage with N^-1 folded in, primitive.rs:270-285.
__device__ __forceinline__ void gs_last_bf(uint64_t &a, uint64_t &b, const DevTable &t) {
    uint64_t u = a, v = b;
    a = shoup_full(t.inv_n, t.inv_n_q, u + v, t.q);
    b = shoup_full(t.sr, t.sr_q, u + t.q2 - v, t.q);
}

// R forward stages on one register group x[0 .. 2^R), global stages [S, S+R),
// chunk c at stage S.
template <int R>
__device__ __forceinline__ void ct_round(uint64_t *x, uint32_t S, uint32_t c,
                                         const Shoup *__restrict__ tw, uint64_t q, uint64_t q2) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int half = 1 << (R - u - 1);
        const uint32_t base = (1u << (S + u)) + (c << u);
#pragma unroll
        for (int sub = 0; sub < (1 << u); ++sub) {
            const Shoup w = tw[base + sub];
#pragma unroll
            for (int jj = 0; jj < half; ++jj) {
                const int j0 = sub * 2 * half + jj;
                ct_bf(x[j0], x[j0 + half], w.w, w.wq, q, q2);
            }
        }
    }
}

// R inverse stages on one register group, global inverse stages [P, P+R)
// (gaps 2^P .. 2^(P+R-1)), chunk c at stage P+R-1's grouping (k >> (P+R)).
// LAST: the round ends with the transform's final stage (gap N/2).
template <int R, bool LAST>
__device__ __forceinline__ void gs_round(uint64_t *x, uint32_t P, uint32_t c, uint32_t n,
                                         const Shoup *__restrict__ itw, const DevTable &t) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int half = 1 << u;
        if (LAST && u == R - 1) {
#pragma unroll
            for (int jj = 0; jj < half; ++jj) gs_last_bf(x[jj], x[jj + half], t);
        } else {
            const uint32_t base = 1u + n - (n >> (P + u)) + (c << (R - u - 1));
#pragma unroll
            for (int sub = 0; sub < (1 << (R - u - 1)); ++sub) {
                const Shoup w = itw[base + sub];
#pragma unroll
                for (int jj = 0; jj < half; ++jj) {
                    const int j0 = sub * 2 * half + jj;
                    gs_bf(x[j0], x[j0 + half], w.w, w.wq, t.q, t.q2);
                }
            }
        }
    }
}

// LDS slot of local index k: one pad word per 16 keeps the stride-16 and
// stride-1 access patterns of the rounds free of bank conflicts.
__device__ __forceinline__ uint32_t lds_pos(uint32_t k) { return k + (k >> 4); }

}  // namespace ph
