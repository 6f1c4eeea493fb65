// Writes tests/golden/philox_kat.json: Philox4x32-10 blocks computed by
// rocRAND's open-source engine (rocrand_philox4x32_10.h, host side) for a
// set of (counter, key) pairs, used to pin the oracle's Philox.
// Build and run (CPU only):
//   hipcc --offload-arch=gfx950 -O1 -o /tmp/mkphilox tests/golden/make_philox_vectors.cpp && /tmp/mkphilox > tests/golden/philox_kat.json
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

// ten_rounds is a protected member; expose it through a subclass
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main()
{
    const unsigned cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {7u, 1234u, 99u, (3u << 8) | (1u << 7) | (2u << 4), 0x5678u, 0x1234ABCDu},
        {255u, 40000u, 100u, (1u << 7) | (1u << 4), 0u, 0u},
    };
    Engine eng;
    std::printf("{\"source\": \"rocrand_device::philox4x32_10_engine::ten_rounds\", \"cases\": [\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        uint4 c = {cases[i][0], cases[i][1], cases[i][2], cases[i][3]};
        uint2 k = {cases[i][4], cases[i][5]};
        uint4 r = eng.ten_rounds(c, k);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c.x, c.y, c.z, c.w, k.x, k.y, r.x, r.y, r.z, r.w, i + 1 < n ? "," : "");
    }
    std::printf("]}\n");
    return 0;
}
