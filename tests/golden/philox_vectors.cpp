// Prints Philox4x32-10 (counter, key) -> output vectors computed by rocRAND's
// open-source implementation (rocrand_philox4x32_10.h, ten_rounds), host side.
// Built and run by make_philox_vectors.py; the output is a test fixture.
#include <cstdio>
#include <cstdint>
#include <rocrand/rocrand_philox4x32_10.h>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    Probe p;
    uint64_t s = 0x0123456789abcdefull;  // xorshift64* stream for the inputs
    auto next = [&]() -> uint32_t {
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        return (uint32_t)((s * 0x2545F4914F6CDD1Dull) >> 32);
    };
    std::printf("[\n");
    const uint32_t fixed[3][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u}};
    for (int i = 0; i < 67; ++i) {
        uint32_t v[6];
        for (int j = 0; j < 6; ++j) v[j] = (i < 3) ? fixed[i][j] : next();
        uint4 c = {v[0], v[1], v[2], v[3]};
        uint2 k = {v[4], v[5]};
        uint4 o = p.run(c, k);
        std::printf(" {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w, i < 66 ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
