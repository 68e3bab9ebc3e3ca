// Writes Philox4x32-10 known-answer vectors from rocRAND's independent
// implementation (rocrand_philox4x32_10.h, host path).  Usage:
//   hipcc -o /tmp/gen_philox_kat tests/golden/gen_philox_kat.cpp && /tmp/gen_philox_kat > tests/golden/philox_kat.json
#include <cstdio>
#include <cstdint>
#include <rocrand/rocrand_philox4x32_10.h>

struct kat_engine : rocrand_device::philox4x32_10_engine {
    static uint4 run(uint4 c, uint2 k) { return kat_engine().ten_rounds(c, k); }
};

int main() {
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {7u, 0xffffffffu, 100u, (3u << 16) | (1u << 10) | (4u << 7) | 2u, 0x12345678u, 0x9abcdef0u},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    uint32_t lcg = 12345u;
    for (int i = 0; i < n + 16; i++) {
        uint32_t c[6];
        if (i < n) { for (int j = 0; j < 6; j++) c[j] = cases[i][j]; }
        else { for (int j = 0; j < 6; j++) { lcg = lcg * 1664525u + 1013904223u; c[j] = lcg; } }
        uint4 o = kat_engine::run(uint4{c[0], c[1], c[2], c[3]}, uint2{c[4], c[5]});
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c[0], c[1], c[2], c[3], c[4], c[5], o.x, o.y, o.z, o.w, i + 1 < n + 16 ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
