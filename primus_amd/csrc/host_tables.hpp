// host_tables.hpp -- table construction on the host: exact 128-bit modular
// arithmetic, the primitive root, the Shoup twiddle tables of
// UintNttTable::new and their binary64 and register-order forms.
//
// Plain C++ (no HIP): the library builds its device tables from it, and a
// host-only sanitizer build checks it without a GPU.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/primus_hip.h"

namespace ph {

typedef unsigned __int128 u128;

inline uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

inline uint64_t h_powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = h_mulmod(r, b, q);
        b = h_mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

inline uint64_t h_shoup_quotient(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

// floor(2^128 / q) as two words: BarrettModulus's ratio
// (primus_modulus/src/barrett/mod.rs:44-45, 55-75)
inline void h_barrett_ratio(uint64_t q, uint64_t *r0, uint64_t *r1) {
    u128 all = ~(u128)0;
    u128 r = all / q;
    if (all % q == q - 1) r += 1;
    *r0 = (uint64_t)r;
    *r1 = (uint64_t)(r >> 64);
}

struct alignas(16) Shoup {
    uint64_t w;   // value
    uint64_t wq;  // floor(w * 2^64 / q)
};

struct alignas(16) ShoupF {
    double w;   // value (exact integer)
    double wp;  // w / q rounded to binary64
};

inline uint64_t rev_bits(uint64_t x, uint32_t bits) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1ull) << (bits - 1 - i);
    return r;
}

// w^(2^(log_degree-1)) == q - 1, primus_ntt/src/root.rs:27-44.
inline bool is_primitive_root(uint64_t w, uint32_t log_degree, uint64_t q) {
    if (w == 0 || log_degree == 0) return false;
    uint64_t x = w;
    for (uint32_t i = 0; i + 1 < log_degree; ++i) x = h_mulmod(x, x, q);
    return x == q - 1;
}

// try_minimal_primitive_root (primus_ntt/src/root.rs:46-111): the smallest
// primitive 2^log_degree-th root.  Candidates r = 2, 3, ... replace the
// reference's random draws; the minimum over odd powers does not depend on
// which primitive root is found first.
inline int minimal_primitive_root(uint32_t log_degree, uint64_t q, uint64_t *root) {
    uint64_t qm1 = q - 1;
    uint64_t quotient = qm1 >> log_degree;
    if ((quotient << log_degree) != qm1) return PH_ERR_NO_PRIMITIVE_ROOT;
    uint64_t w = 0;
    bool found = false;
    for (uint64_t r = 2; r <= qm1 && r < (2ull + (1ull << 20)); ++r) {
        w = h_powmod(r, quotient, q);
        if (is_primitive_root(w, log_degree, q)) {
            found = true;
            break;
        }
    }
    if (!found) return PH_ERR_NO_PRIMITIVE_ROOT;
    uint64_t best = w, g2 = h_mulmod(w, w, q), cur = w;
    const uint64_t degree = 1ull << log_degree;
    for (uint64_t i = 0; i < degree; ++i) {
        if (cur < best) best = cur;
        cur = h_mulmod(g2, cur, q);
    }
    *root = best;
    return PH_OK;
}

inline uint64_t inv_mod(uint64_t a, uint64_t q) {
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        __int128 k = r / nr, tmp;
        tmp = t - k * nt; t = nt; nt = tmp;
        tmp = r - k * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

inline Shoup mk(uint64_t w, uint64_t q) { return Shoup{w, h_shoup_quotient(w, q)}; }

// The register rounds of a single-pass launch with 2^loge coefficients per
// thread (Geo in ntt_rounds.hpp): forward round i covers stages
// [fs(i), fs(i) + fr(i)), inverse round i stages [is(i), is(i) + ir(i)).
struct RoundPlan {
    int loge, rem, nr;
    RoundPlan(uint32_t log_n, int loge_req) {
        loge = (int)log_n > loge_req ? loge_req : (int)log_n;
        rem = (int)log_n % loge;
        nr = (int)log_n / loge + (rem ? 1 : 0);
    }
    int fr(int i) const { return (rem && i == 0) ? rem : loge; }
    int fs(int i) const { return i == 0 ? 0 : (rem ? rem + (i - 1) * loge : i * loge); }
    int ir(int i) const { return (rem && i == nr - 1) ? rem : loge; }
    int is(int i) const { return i * loge; }
};

// Twiddle tables in register order for one launch shape.  A round starting at
// stage S gives the group of chunk c, at its stage S+u, the twiddles of chunks
// c*2^u + sub (sub < 2^u); the CPU order stores them together per c, so
// neighbouring lanes (neighbouring c) read addresses 2^u entries apart and a
// wave touches up to 64 lines per load.  Stored instead at
// 2^(S+u) + sub*2^S + c, each load of a wave reads consecutive entries.  The
// inverse does the same within each stage's segment, with the chunks counted
// at the round's last stage (n >> (P+R) of them).  Entries are the same
// values; only positions change.
inline void register_order(const std::vector<ShoupF> &fwd, const std::vector<ShoupF> &inv, uint32_t log_n, int loge,
                     std::vector<ShoupF> &pf, std::vector<ShoupF> &pi) {
    const uint64_t n = 1ull << log_n;
    const RoundPlan g(log_n, loge);
    pf.assign(n, fwd[0]);
    pi.assign(n, inv[0]);
    for (int i = 0; i < g.nr; ++i) {
        const int S = g.fs(i), R = g.fr(i);
        for (int u = 0; u < R; ++u) {
            const uint64_t s = (uint64_t)(S + u);
            for (uint64_t c = 0; c < (1ull << S); ++c)
                for (uint64_t sub = 0; sub < (1ull << u); ++sub)
                    pf[(1ull << s) + (sub << S) + c] = fwd[(1ull << s) + (c << u) + sub];
        }
    }
    for (int i = 0; i < g.nr; ++i) {
        const int P = g.is(i), R = g.ir(i);
        const uint64_t C = n >> (P + R);
        for (int u = 0; u < R; ++u) {
            const uint64_t s = (uint64_t)(P + u);
            if (s + 1 == log_n) continue;  // the last stage folds N^-1 and uses no table entry
            const uint64_t base = 1 + n - (n >> s);
            const uint64_t subs = 1ull << (R - u - 1);
            for (uint64_t c = 0; c < C; ++c)
                for (uint64_t sub = 0; sub < subs; ++sub) pi[base + sub * C + c] = inv[base + c * subs + sub];
        }
    }
}

// A resident [L][N] operand in the register order of single-pass launches
// with 2^loge coefficients per thread: thread t of a workgroup holds local
// slots t*E .. t*E+E-1 after the forward rounds and multiplies slot t*E + j
// by the entry stored at j*T + t, so the j-th operand load of a wave reads
// consecutive entries.
inline std::vector<ShoupF> operand_register_order(const std::vector<ShoupF> &f, uint32_t L, uint32_t log_n, int loge) {
    const uint64_t n = 1ull << log_n;
    const uint64_t E = 1ull << std::min<uint32_t>((uint32_t)loge, log_n), T = n / E;
    std::vector<ShoupF> out(f.size());
    for (uint64_t l = 0; l < L; ++l)
        for (uint64_t t = 0; t < T; ++t)
            for (uint64_t j = 0; j < E; ++j) out[l * n + j * T + t] = f[l * n + t * E + j];
    return out;
}

// binary64 twiddle, centred: w' = w or w - q (|w'| <= q/2) and w'/q from the
// Shoup quotient, |wp - w'/q| <= 2^-54 + 2^-64 (the subtraction of 1 is exact).
inline ShoupF mkf(const Shoup &s, uint64_t q) {
    const double frac = (double)s.wq * 0x1p-64;  // w/q
    if (s.w > q / 2) return ShoupF{(double)s.w - (double)q, frac - 1.0};
    return ShoupF{(double)s.w, frac};
}


// The host half of UintNttTable::new (primus_ntt/src/ntt/primitive.rs:114-185)
// with the constants the device kernels add to it.
struct HostTable {
    uint64_t root = 0, inv_root = 0;
    uint64_t inv_n = 0, inv_n_q = 0;  // N^-1 mod q and its Shoup quotient
    uint64_t sr = 0, sr_q = 0;        // N^-1 * inv_root_powers[N-1], the last inverse stage's root
    uint64_t br0 = 0, br1 = 0;        // floor(2^128 / q)
    uint64_t bar_mu = 0;              // floor(2^(b-1+64) / q) for q < 2^59 (b = bit length of q), else 0
    uint32_t bar_sh = 0;              // b - 1
    std::vector<Shoup> ord, fwd, inv; // ordinal_root_powers, root_powers, inv_root_powers
};

// Status PH_OK, PH_ERR_NO_PRIMITIVE_ROOT, PH_ERR_DEGREE_TOO_LARGE or
// PH_ERR_NO_MEMORY; the caller has checked 1 < q < 2^62 and 1 <= log_n.
inline int host_table(uint32_t log_n, uint64_t q, HostTable &h) {
    int st = minimal_primitive_root(log_n + 1, q, &h.root);
    if (st != PH_OK) return st;
    const uint64_t n = 1ull << log_n;
    if (n >= q) return PH_ERR_DEGREE_TOO_LARGE;
    try {
        h.ord.resize(2 * n);
        h.fwd.resize(n);
        h.inv.resize(n);
    } catch (const std::bad_alloc &) {
        return PH_ERR_NO_MEMORY;
    }
    // ordinal_root_powers[k] = root^k (primitive.rs:128-137)
    uint64_t p = 1 % q;
    for (uint64_t k = 0; k < 2 * n; ++k) {
        h.ord[k] = mk(p, q);
        p = h_mulmod(p, h.root, q);
    }
    h.inv_root = h.ord[2 * n - 1].w;
    // root_powers[rev(k)] = root^k (:145-149)
    for (uint64_t k = 0; k < n; ++k) h.fwd[rev_bits(k, log_n)] = h.ord[k];
    // inv_root_powers[0] = 1, inv_root_powers[rev(k)+1] = root^-(k+1) (:151-159)
    h.inv[0] = h.ord[0];
    for (uint64_t k = 0; k + 1 < n; ++k) h.inv[rev_bits(k, log_n) + 1] = h.ord[2 * n - 1 - k];
    h.inv_n = inv_mod(n % q, q);
    h.inv_n_q = h_shoup_quotient(h.inv_n, q);
    // scaled root of the last inverse stage: N^-1 * inv_root_powers[N-1] (:272-277)
    h.sr = (uint64_t)(((u128)h.inv[n - 1].w * h.inv_n) % q);
    h.sr_q = h_shoup_quotient(h.sr, q);
    h_barrett_ratio(q, &h.br0, &h.br1);
    if (q < (1ull << 59) && q > 2) {
        const uint32_t b = 64 - __builtin_clzll(q);
        h.bar_sh = b - 1;
        h.bar_mu = (uint64_t)((((u128)1) << (b - 1 + 64)) / q);
    }
    return PH_OK;
}

}  // namespace ph
