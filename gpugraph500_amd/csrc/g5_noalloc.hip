// g5_noalloc.hip -- the codec round trip for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): g5_s4_roundtrip_noalloc.  It runs the same kernels as
// g5_s4_roundtrip, with the codec's scratch laid out in one file-scope
// device array instead of memory the library allocates.
//
// Every GPU the library runs on holds its own copy of the array, at its own
// address, until the process ends.  On the host the array's name only stands
// in for the symbol: passing it to a kernel would fault.  Each device's
// address therefore comes from hipGetSymbolAddress, taken once per device
// with that device current (g5_init, or the entry point's first call on a
// device), and only that address is used.  Every call on one GPU shares the
// array: calls that overlap on one GPU race.  Nothing in the array must
// start at zero: the header search clears the decoder's record (and its
// error and lost-values flag) before anything reads it, as in the allocating
// round trip, whose scratch is also reused from call to call.
#include <mutex>

#include "g5_codec.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_ptr[kMaxDevices] = {};

}  // namespace

namespace g5 {

// Once per device, with that device current.
int noalloc_setup(int device)
{
    G5_REQUIRE(device >= 0 && device < kMaxDevices, G5_E_ARG, "noalloc scratch: device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return G5_OK;
    const size_t need = S4Codec::scratch_bytes(G5_S4_NOALLOC_MAX_VALUES);
    G5_REQUIRE(need <= kScratchBytes, G5_E_STATE, "noalloc scratch: %zu bytes needed for %d values, %zu reserved",
               need, (int)G5_S4_NOALLOC_MAX_VALUES, kScratchBytes);
    void *p = nullptr;
    G5_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[device] = (uint8_t *)p;
    return G5_OK;
}

}  // namespace g5

extern "C" int g5_s4_roundtrip_noalloc(const uint32_t *d_in, size_t n, uint32_t *d_words, size_t cap_words,
                                       uint32_t *d_out, size_t *nwords)
{
    G5_REQUIRE(d_in && d_words && d_out && nwords, G5_E_ARG, "g5_s4_roundtrip_noalloc: null argument");
    G5_REQUIRE(n <= (size_t)G5_S4_NOALLOC_MAX_VALUES, G5_E_ARG,
               "g5_s4_roundtrip_noalloc: %zu values, at most %d (the size of its static scratch)", n,
               (int)G5_S4_NOALLOC_MAX_VALUES);
    int dev = 0;
    G5_HIP(hipGetDevice(&dev));
    int rc = g5::noalloc_setup(dev);
    if (rc) return rc;
    // a codec laid out in this device's scratch for the maximum size (so its
    // reserve() never allocates); its record comes home through a host array
    thread_local int64_t h_meta[16];
    g5::S4Codec c;
    c.bind(g_ptr[dev], G5_S4_NOALLOC_MAX_VALUES);
    c.h_meta = h_meta;
    rc = c.roundtrip(d_in, (int64_t)n, d_words, cap_words, d_out, nullptr, nullptr);
    if (rc) return rc;
    return c.roundtrip_result(nwords, "g5_s4_roundtrip_noalloc");
}
