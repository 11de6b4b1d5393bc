// Thin extern "C" entry points over the reference codec library, compiled
// against the reference's own sources where they lie (see oracle/Makefile).
// Used only to write and check golden fixtures.
#include "codecfactory.h"
using namespace SIMDCompressionLib;

extern "C" int ref_encode(const char *name, const uint32_t *in, size_t n, uint32_t *out,
                          size_t *nwords)
{
    if (!CODECFactory::valid(name)) return -1;
    IntegerCODEC &c = *CODECFactory::getFromName(name);
    // the codec API takes a non-const input; work on a 16-byte aligned copy
    std::vector<uint32_t, std::allocator<uint32_t>> tmp(in, in + n);
    uint32_t *buf = nullptr;
    if (posix_memalign((void **)&buf, 16, (n + 4) * 4)) return -2;
    std::copy(tmp.begin(), tmp.end(), buf);
    size_t m = *nwords;
    try {
        c.encodeArray(buf, n, out, m);
    } catch (...) {
        free(buf);
        return -3;
    }
    free(buf);
    *nwords = m;
    return 0;
}

extern "C" int ref_decode(const char *name, const uint32_t *in, size_t nwords, uint32_t *out,
                          size_t *n)
{
    if (!CODECFactory::valid(name)) return -1;
    IntegerCODEC &c = *CODECFactory::getFromName(name);
    size_t m = *n;
    try {
        c.decodeArray(in, nwords, out, m);
    } catch (...) {
        return -3;
    }
    *n = m;
    return 0;
}
