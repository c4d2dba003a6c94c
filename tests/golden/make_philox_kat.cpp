// Writes philox_kat.json: Philox4x32-10 known-answer vectors produced by
// rocRAND's engine (rocrand/rocrand_philox4x32_10.h, ten_rounds), the
// published-algorithm implementation the oracle's and the product's Philox
// are pinned against. Build and run (host code only):
//   hipcc -O1 tests/golden/make_philox_kat.cpp -o /tmp/kat && /tmp/kat > tests/golden/philox_kat.json
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>
#include <cstdint>

// ten_rounds is a protected member; a derived class exposes it.
struct KatEngine : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    KatEngine eng;
    struct Case {
        uint32_t c[4];
        uint32_t k[2];
    };
    Case cases[64];
    int n = 0;
    cases[n++] = {{0u, 0u, 0u, 0u}, {0u, 0u}};
    cases[n++] = {{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}, {0xffffffffu, 0xffffffffu}};
    cases[n++] = {{0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}, {0xa4093822u, 0x299f31d0u}};
    // the counter/key shapes the PatchMatch stream uses: (pixel, phase, block, 0), (3407, view)
    uint32_t s = 12345u;
    for (int i = 0; i < 40; ++i) {
        s = s * 1664525u + 1013904223u;
        const uint32_t pixel = s % 7680000u;
        s = s * 1664525u + 1013904223u;
        const uint32_t phase = (i % 5 == 0) ? 0x40000000u : (s % 16u);
        cases[n++] = {{pixel, phase, uint32_t(i % 14), 0u}, {3407u, uint32_t(i % 49)}};
    }
    std::printf("{\"generator\": \"rocrand_device::philox4x32_10_engine::ten_rounds (ROCm 7.2 header)\",\n \"cases\": [\n");
    for (int i = 0; i < n; ++i) {
        uint4 ctr = {cases[i].c[0], cases[i].c[1], cases[i].c[2], cases[i].c[3]};
        uint2 key = {cases[i].k[0], cases[i].k[1]};
        uint4 o = eng.block(ctr, key);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    ctr.x, ctr.y, ctr.z, ctr.w, key.x, key.y, o.x, o.y, o.z, o.w,
                    i + 1 < n ? "," : "");
    }
    std::printf(" ]}\n");
    return 0;
}
