// Philox4x32-10 blocks from rocRAND's engine (an implementation independent
// of this repository's), for the oracle's known-answer test.  Built for the
// host only; prints one line per (counter, key): the 4 output words in hex.
#include <cstdio>

#include <rocrand/rocrand_philox4x32_10.h>

struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
    const unsigned cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {7u, 3u, 1u, 0u, 0x12345678u, 0x9abcdef0u},
        {4095u, 1023u, 14u, 0u, 0xdeadbeefu, 0x00c0ffeeu},
    };
    Engine eng;
    for (const auto &c : cases) {
        const uint4 r = eng.ten_rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{c[4], c[5]});
        std::printf("%08x %08x %08x %08x %08x %08x -> %08x %08x %08x %08x\n", c[0], c[1], c[2], c[3], c[4], c[5],
                    r.x, r.y, r.z, r.w);
    }
    return 0;
}
