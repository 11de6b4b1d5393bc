// Host-only unit checks of the C++ codec schemas (no GPU call is reached:
// every queue here is at or under the schema's threshold, or the check is
// about configuration).  Built and run by tests/test_host.py.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "g5_compression.hh"

using namespace g5host;

static int failures = 0;
#define EXPECT(cond)                                                   \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::fprintf(stderr, "FAILED line %d: %s\n", __LINE__, #cond); \
            ++failures;                                                \
        }                                                              \
    } while (0)

template <typename F>
static bool throws(F f)
{
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

int main()
{
    // nocompression: exact round trip, never "compressed"
    {
        NoCompression<int64_t, uint32_t> nc;
        int64_t q[5] = {1, 3, 5, 1039, 1045};
        uint32_t *c = nullptr;
        size_t cs = 0;
        nc.compress(q, 5, &c, cs);
        EXPECT(cs == 5);
        EXPECT(((uintptr_t)c & 15) == 0);
        int64_t *back = nullptr;
        size_t n = 5;
        nc.decompress(c, (int)cs, &back, n);
        EXPECT(n == 5);
        for (int i = 0; i < 5; ++i) EXPECT(back[i] == q[i]);
        EXPECT(!nc.isCompressed(5, 5));
        EXPECT(nc.name() == "nocompression");
        free(c);
        free(back);
    }
    // hipsimd: the gate n > threshold, the receiver's rule, configuration
    {
        HipSimd<int64_t, uint32_t> hs;
        EXPECT(hs.threshold() == 64);
        EXPECT(!hs.isCompressed(64, 10));   // at the threshold: raw
        EXPECT(hs.isCompressed(65, 10));    // above it and shorter: coded
        EXPECT(!hs.isCompressed(65, 65));   // same length: read as raw
        hs.reconfigure(8, "s4-bp128-d4");
        EXPECT(hs.threshold() == 8);
        EXPECT(throws([&] { hs.reconfigure(8, "s4-bp128-dm"); }));
        EXPECT(throws([&] { hs.reconfigure(0, "s4-bp128-d4"); }));
        // a queue at the threshold travels raw without touching the device
        int64_t q[8] = {0, 2, 4, 6, 8, 10, 12, 14};
        uint32_t *c = nullptr;
        size_t cs = 0;
        hs.compress(q, 8, &c, cs);
        EXPECT(cs == 8);
        for (int i = 0; i < 8; ++i) EXPECT(c[i] == (uint32_t)q[i]);
        int64_t *back = nullptr;
        size_t n = 8;
        hs.decompress(c, (int)cs, &back, n);
        EXPECT(n == 8);
        for (int i = 0; i < 8; ++i) EXPECT(back[i] == q[i]);
        free(c);
        free(back);
    }
    // factory names
    {
        typedef CompressionFactory<int64_t, uint32_t> Factory;
        EXPECT(Factory::getFromName("nocompression")->name() == "nocompression");
        EXPECT(Factory::getFromName("cpusimd")->name() == "hipsimd");
        EXPECT(throws([] { Factory::getFromName("simdplus"); }));
    }
    // the C ABI's worst-case stream size (host-only function)
    EXPECT(g5_s4_max_words(0) >= 4);
    EXPECT(g5_s4_max_words(128) >= 4 + 4 + 128);
    if (failures) return 1;
    std::printf("host_unit: all checks passed\n");
    return 0;
}
