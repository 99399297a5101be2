// Host-side check of the register-round slot maps of ntt_rounds.hpp for every
// chunk geometry the kernels instantiate (2^1 <= C <= 2^13, 1 <= E <= C/2):
//   * each round's slot map (thread t, register group m, register j) is a
//     bijection onto the chunk;
//   * a slot is its thread part plus its register part, and so is its LDS
//     position: lds_pos(slot(t,m,j)) == lds_pos(slot(t,0,0)) + lds_pos(slot(0,m,j)).
// The kernels address global rows and the LDS exchanges with one thread base
// per round plus constant offsets, which is valid only under the second
// property.
// It also counts LDS bank conflicts of every round's 64-bit accesses under
// the model the hardware's counters agree with (32 four-byte banks, a wave's
// access served 16 lanes per cycle): the single-pass shapes of 8 and 16
// coefficients per thread (2^6 <= C <= 2^13, full waves) must have none.
// Built for the host only; prints "ok <checks>" and exits 0.
#include <cstdio>
#include <vector>

#include "../../primus_amd/csrc/ntt_rounds.hpp"

using namespace ph;

static long checks = 0, failures = 0;

static long conflict_cycles = 0, conflict_failures = 0;

// Extra cycles of one round's accesses: for every wave, register and 16-lane
// quarter, the largest number of lanes meeting in one bank, less one.
template <class F>
static long bank_conflicts(int logt, int E, int R, F slot) {
    long extra = 0;
    for (uint32_t w = 0; w < (1u << logt) / 64; ++w)
        for (int m = 0; m < (E >> R); ++m)
            for (int j = 0; j < (1 << R); ++j)
                for (uint32_t q = 0; q < 64; q += 16) {
                    int use[32] = {0}, worst = 0;
                    for (uint32_t l = q; l < q + 16; ++l) {
                        const uint32_t p = lds_pos(slot(64 * w + l, m, j));
                        for (uint32_t d = 0; d < 2; ++d) {
                            const int b = (int)((2 * p + d) % 32);
                            worst = ++use[b] > worst ? use[b] : worst;
                        }
                    }
                    extra += worst - 1;
                }
    return extra;
}

template <class F>
static void check_round(int logc, int logt, int I, char kind, int E, int R, F slot) {
    if (logt >= 6 && logc >= 6 && (E == 8 || E == 16)) {
        const long c = bank_conflicts(logt, E, R, slot);
        conflict_cycles += c;
        if (c && conflict_failures++ < 10)
            std::printf("BANK CONFLICTS C=2^%d T=2^%d round %c%d: %ld\n", logc, logt, kind, I, c);
    }
    std::vector<int> seen(1u << logc, 0);
    for (uint32_t t = 0; t < (1u << logt); ++t) {
        const uint32_t base = lds_pos(slot(t, 0, 0));
        for (int m = 0; m < (E >> R); ++m)
            for (int j = 0; j < (1 << R); ++j) {
                const uint32_t s = slot(t, m, j);
                ++checks;
                if (s >= (1u << logc) || seen[s]++ ||
                    lds_pos(s) != base + lds_pos(slot(0, m, j)) || s != slot(t, 0, 0) + slot(0, m, j)) {
                    if (failures++ < 10)
                        std::printf("FAIL C=2^%d T=2^%d round %c%d t=%u m=%d j=%d\n", logc, logt, kind, I, t, m, j);
                }
            }
    }
}

template <int LOGC, int LOGT, int I>
static void rounds() {
    using G = Geo<LOGC, LOGT>;
    if constexpr (I < G::NR) {
        constexpr int S = G::fs(I), R = G::fr(I);
        check_round(LOGC, LOGT, I, 'f', G::E, R,
                    [](uint32_t t, int m, int j) { return fwd_slot<LOGC, LOGT, S, R>(t, m, j); });
        constexpr int P = G::is(I), RI = G::ir(I);
        check_round(LOGC, LOGT, I, 'i', G::E, RI,
                    [](uint32_t t, int m, int j) { return inv_slot<LOGC, LOGT, P, RI>(t, m, j); });
        rounds<LOGC, LOGT, I + 1>();
    }
}

template <int LOGC, int LOGT>
static void geometries() {
    if constexpr (LOGT < LOGC) {
        rounds<LOGC, LOGT, 0>();
        geometries<LOGC, LOGT + 1>();
    }
}

template <int LOGC>
static void all() {
    if constexpr (LOGC <= 13) {
        geometries<LOGC, 0>();
        all<LOGC + 1>();
    }
}

int main() {
    all<1>();
    if (failures || conflict_failures) {
        std::printf("%ld of %ld checks failed; %ld rounds with bank conflicts\n", failures, checks, conflict_failures);
        return 1;
    }
    std::printf("ok %ld\n", checks);
    return 0;
}
