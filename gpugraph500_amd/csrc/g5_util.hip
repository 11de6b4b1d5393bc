// g5_util.hip -- library plumbing and device-wide primitives: error state,
// device memory helpers, exclusive scan, bitmap <-> sorted queue.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "g5_internal.hpp"

namespace g5 {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail_hip(hipError_t e, const char *what, const char *file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? G5_E_NOMEM : G5_E_HIP;
}

int grid_for(int64_t n, int block, int max_grid)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_grid) g = max_grid;
    return (int)g;
}

// ---------------------------------------------------------------- fills

template <typename T>
__global__ void k_fill(T *d, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

hipError_t fill_u64(uint64_t *d, int64_t n, uint64_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<uint64_t>, grid_for(n, 256, 8192), 256, 0, s, d, n, v);
    return hipGetLastError();
}
hipError_t fill_u32(uint32_t *d, int64_t n, uint32_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<uint32_t>, grid_for(n, 256, 8192), 256, 0, s, d, n, v);
    return hipGetLastError();
}
hipError_t fill_i64(int64_t *d, int64_t n, int64_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<int64_t>, grid_for(n, 256, 8192), 256, 0, s, d, n, v);
    return hipGetLastError();
}
hipError_t fill_i32(int32_t *d, int64_t n, int32_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<int32_t>, grid_for(n, 256, 8192), 256, 0, s, d, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scan
// Three-phase exclusive scan over 1024-element tiles (256 threads x 4):
// tile sums -> recursive scan of the sums -> tile scan plus offset.

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *lds_waves, int64_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t inc = g5_wave_incl_scan(v);
    if (lane == 63) lds_waves[wave] = inc;
    __syncthreads();
    int64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        const int64_t t = lds_waves[w];
        off += (w < wave) ? t : 0;
        tot += t;
    }
    *total = tot;
    return off + inc - v;
}

__global__ void k_tile_sums(const int64_t *d, int64_t n, int64_t *sums)
{
    __shared__ int64_t lw[kScanBlock / 64];
    const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += d[base + i];
    int64_t tot;
    block_excl_scan(s, lw, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_tile_scan(int64_t *d, int64_t n, const int64_t *offs, int64_t *d_total)
{
    __shared__ int64_t lw[kScanBlock / 64];
    const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? d[base + i] : 0;
        s += v[i];
    }
    int64_t tot;
    int64_t run = block_excl_scan(s, lw, &tot) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) d[base + i] = run;
        run += v[i];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) *d_total = run;
}

size_t scan_scratch_words(int64_t n)
{
    size_t w = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        w += (size_t)n + 1;
    }
    return w + 1;
}

hipError_t exclusive_scan_i64(int64_t *d, int64_t n, int64_t *d_total, int64_t *scratch,
                              hipStream_t s)
{
    if (n <= 0) return d_total ? fill_i64(d_total, 1, 0, s) : hipSuccess;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_tile_scan, 1, kScanBlock, 0, s, d, n, (const int64_t *)nullptr,
                           d_total);
        return hipGetLastError();
    }
    int64_t *sums = scratch;
    hipLaunchKernelGGL(k_tile_sums, (unsigned)tiles, kScanBlock, 0, s, (const int64_t *)d, n, sums);
    hipError_t e = exclusive_scan_i64(sums, tiles, nullptr, scratch + tiles + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_scan, (unsigned)tiles, kScanBlock, 0, s, d, n,
                       (const int64_t *)sums, d_total);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bitmap <-> queue

// One thread per word.  Word popcounts are scanned per 256-word block; the
// block offsets come from a scan of the block totals.
__global__ void k_bm_counts(const uint64_t *bits, int64_t n_words, int64_t *blk)
{
    __shared__ int64_t lw[4];
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t c = (w < n_words) ? __popcll(bits[w]) : 0;
    int64_t tot;
    block_excl_scan(c, lw, &tot);
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void k_bm_write(const uint64_t *bits, int64_t n_words, const int64_t *blk_off,
                           uint32_t *queue, int64_t *d_count)
{
    __shared__ int64_t lw[4];
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint64_t word = (w < n_words) ? bits[w] : 0;
    int64_t tot;
    int64_t pos = block_excl_scan(__popcll(word), lw, &tot) + blk_off[blockIdx.x];
    while (word) {
        const int b = __ffsll((long long)word) - 1;
        queue[pos++] = (uint32_t)(w * 64 + b);
        word &= word - 1;
    }
    if (d_count && blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) *d_count = pos;
}

hipError_t bitmap_to_queue(const uint64_t *bits, int64_t n_words, uint32_t *queue,
                           int64_t *d_count, int64_t *scratch, hipStream_t s)
{
    if (n_words <= 0) return fill_i64(d_count, 1, 0, s);
    const int64_t blocks = (n_words + 255) / 256;
    int64_t *blk = scratch;
    hipLaunchKernelGGL(k_bm_counts, (unsigned)blocks, 256, 0, s, bits, n_words, blk);
    hipError_t e = exclusive_scan_i64(blk, blocks, nullptr, scratch + blocks + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bm_write, (unsigned)blocks, 256, 0, s, bits, n_words,
                       (const int64_t *)blk, queue, d_count);
    return hipGetLastError();
}

// Sorted queue: the first entry of each run of equal words ORs the run's bits
// with one atomic (runs that cross a wave boundary simply take two atomics).
__global__ void k_q2bm(const uint32_t *q, int64_t n, uint64_t *bits)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = q[i];
    const uint32_t w = v >> 6;
    const bool lead = (i == 0) || ((q[i - 1] >> 6) != w) || ((threadIdx.x & 63) == 0);
    if (!lead) return;
    uint64_t acc = 0;
    int64_t j = i;
    do {
        acc |= 1ull << (q[j] & 63);
        ++j;
    } while (j < n && (q[j] >> 6) == w && (((threadIdx.x + (j - i)) & 63) != 0));
    atomicOr((unsigned long long *)&bits[w], (unsigned long long)acc);
}

hipError_t queue_to_bitmap(const uint32_t *q, int64_t n, uint64_t *bits, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_q2bm, (unsigned)((n + 255) / 256), 256, 0, s, q, n, bits);
    return hipGetLastError();
}

}  // namespace g5

// ---------------------------------------------------------------- C ABI

extern "C" const char *g5_last_error(void) { return g5::g_err; }

// ---------------------------------------------------------------- exchange accounting

namespace {
std::mutex g_ex_mu;
bool g_ex_on = false;
double g_ex_ms[G5_UNIT_COUNT] = {}, g_ex_bytes[G5_UNIT_COUNT] = {};
int64_t g_ex_calls[G5_UNIT_COUNT] = {};
}  // namespace

namespace g5 {

ExchangeTimer::ExchangeTimer(hipStream_t s) : s_(s)
{
    {
        std::lock_guard<std::mutex> lk(g_ex_mu);
        if (!g_ex_on) return;
    }
    if (hipEventCreate(&e0_) != hipSuccess || hipEventCreate(&e1_) != hipSuccess) {
        if (e0_) (void)hipEventDestroy(e0_);
        if (e1_) (void)hipEventDestroy(e1_);
        e0_ = e1_ = nullptr;
    }
}

ExchangeTimer::~ExchangeTimer()
{
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
}

hipError_t ExchangeTimer::start() { return e0_ ? hipEventRecord(e0_, s_) : hipSuccess; }

hipError_t ExchangeTimer::stop() { return e1_ ? hipEventRecord(e1_, s_) : hipSuccess; }

hipError_t ExchangeTimer::commit(int unit, double bytes, int64_t calls)
{
    if (!e0_) return hipSuccess;
    hipError_t e = hipEventSynchronize(e1_);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0_, e1_);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_ex_mu);
    g_ex_ms[unit] += ms;
    g_ex_bytes[unit] += bytes;
    g_ex_calls[unit] += calls;
    return hipSuccess;
}

}  // namespace g5

extern "C" int g5_exchange_profile(int enable)
{
    std::lock_guard<std::mutex> lk(g_ex_mu);
    g_ex_on = enable != 0;
    for (int u = 0; u < G5_UNIT_COUNT; ++u) {
        g_ex_ms[u] = g_ex_bytes[u] = 0;
        g_ex_calls[u] = 0;
    }
    return G5_OK;
}

extern "C" int g5_exchange_stats(int unit, double *ms, double *bytes, int64_t *calls)
{
    G5_REQUIRE(unit >= 0 && unit < G5_UNIT_COUNT, G5_E_ARG, "g5_exchange_stats: unit %d out of range", unit);
    std::lock_guard<std::mutex> lk(g_ex_mu);
    if (ms) *ms = g_ex_ms[unit];
    if (bytes) *bytes = g_ex_bytes[unit];
    if (calls) *calls = g_ex_calls[unit];
    return G5_OK;
}

extern "C" int g5_init(int device)
{
    G5_HIP(hipSetDevice(device));
    G5_HIP(hipFree(nullptr));
    // the device's address of the no-allocation round trip's scratch
    return g5::noalloc_setup(device);
}

extern "C" int g5_mem_info(int64_t *free_bytes, int64_t *total_bytes)
{
    size_t f = 0, t = 0;
    G5_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return G5_OK;
}

extern "C" int g5_device_sync(void)
{
    G5_HIP(hipDeviceSynchronize());
    return G5_OK;
}

extern "C" int g5_dmalloc(void **p, size_t bytes)
{
    G5_REQUIRE(p != nullptr, G5_E_ARG, "g5_dmalloc: null output pointer");
    G5_HIP(hipMalloc(p, bytes ? bytes : 1));
    return G5_OK;
}

extern "C" int g5_dfree(void *p)
{
    if (p) G5_HIP(hipFree(p));
    return G5_OK;
}

extern "C" int g5_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) G5_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return G5_OK;
}

extern "C" int g5_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes) G5_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return G5_OK;
}

extern "C" int g5_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes)
{
    if (bytes) G5_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
    return G5_OK;
}

extern "C" int g5_fill_u64(uint64_t *d, int64_t n, uint64_t value)
{
    G5_REQUIRE(n >= 0 && (d || n == 0), G5_E_ARG, "g5_fill_u64: bad arguments");
    G5_HIP(g5::fill_u64(d, n, value, nullptr));
    G5_HIP(hipStreamSynchronize(nullptr));
    return G5_OK;
}

namespace {
// *ndiff += the positions where a and b differ
__global__ __launch_bounds__(256) void k_diff_u32(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                  int64_t n, unsigned long long *__restrict__ ndiff)
{
    int64_t c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += a[i] != b[i];
    c = g5_block_sum256(c);
    if (threadIdx.x == 0 && c) atomicAdd(ndiff, (unsigned long long)c);
}
}  // namespace

extern "C" int g5_diff_u32(const uint32_t *d_a, const uint32_t *d_b, int64_t n, int64_t *ndiff)
{
    G5_REQUIRE(n >= 0 && ndiff && (n == 0 || (d_a && d_b)), G5_E_ARG, "g5_diff_u32: bad arguments");
    *ndiff = 0;
    if (n == 0) return G5_OK;
    unsigned long long *d = nullptr, h = 0;
    G5_HIP(hipMalloc(&d, 8));
    hipError_t e = g5::fill_u64((uint64_t *)d, 1, 0, nullptr);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_diff_u32, g5::grid_for(n, 256, 8192), 256, 0, nullptr, d_a, d_b, n, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, 8, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    G5_HIP(e);
    *ndiff = (int64_t)h;
    return G5_OK;
}

extern "C" int g5_bitmap_to_queue(const uint64_t *d_bits, int64_t n_words, uint32_t *d_queue, int64_t *n)
{
    G5_REQUIRE(n_words >= 0 && n && (n_words == 0 || (d_bits && d_queue)), G5_E_ARG,
               "g5_bitmap_to_queue: bad arguments");
    int64_t *d_tmp = nullptr;
    const size_t words = g5::scan_scratch_words((n_words + 255) / 256 + 1) + (n_words + 255) / 256 + 8;
    G5_HIP(hipMalloc(&d_tmp, words * 8));
    hipError_t e = g5::bitmap_to_queue(d_bits, n_words, d_queue, d_tmp, d_tmp + 1, nullptr);
    if (e == hipSuccess) e = hipMemcpy(n, d_tmp, 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_tmp);
    G5_HIP(e);
    return G5_OK;
}

namespace {
__global__ void k_popcount(const uint64_t *__restrict__ bits, int64_t n_words, unsigned long long *__restrict__ out)
{
    int64_t c = 0;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n_words; w += (int64_t)gridDim.x * blockDim.x)
        c += __popcll(bits[w]);
    c = g5_block_sum256(c);
    if (threadIdx.x == 0 && c) atomicAdd(out, (unsigned long long)c);
}
}  // namespace

extern "C" int g5_bitmap_count(const uint64_t *d_bits, int64_t n_words, int64_t *count)
{
    G5_REQUIRE(n_words >= 0 && count && (n_words == 0 || d_bits), G5_E_ARG, "g5_bitmap_count: bad arguments");
    *count = 0;
    if (n_words == 0) return G5_OK;
    unsigned long long *d_c = nullptr;
    G5_HIP(hipMalloc(&d_c, 8));
    hipError_t e = g5::fill_u64((uint64_t *)d_c, 1, 0, nullptr);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_popcount, g5::grid_for(n_words, 256, 2048), 256, 0, nullptr, d_bits, n_words, d_c);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, d_c, 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_c);
    G5_HIP(e);
    *count = (int64_t)h;
    return G5_OK;
}

namespace {
// ids at or beyond n_words * 64 are counted in *bad and skipped
__global__ void k_q2bm_checked(const uint32_t *__restrict__ q, int64_t n, uint64_t *__restrict__ bits,
                               int64_t n_words, unsigned long long *__restrict__ bad)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = q[i];
        if ((int64_t)(v >> 6) >= n_words) {
            atomicAdd(bad, 1ull);
            continue;
        }
        atomicOr((unsigned long long *)&bits[v >> 6], 1ull << (v & 63));
    }
}
}  // namespace

extern "C" int g5_queue_to_bitmap(const uint32_t *d_queue, int64_t n, uint64_t *d_bits, int64_t n_words)
{
    G5_REQUIRE(n >= 0 && n_words >= 0 && (n == 0 || d_queue) && (n_words == 0 || d_bits), G5_E_ARG,
               "g5_queue_to_bitmap: bad arguments");
    G5_HIP(g5::fill_u64(d_bits, n_words, 0, nullptr));
    if (n == 0) {
        G5_HIP(hipStreamSynchronize(nullptr));
        return G5_OK;
    }
    unsigned long long *d_bad = nullptr, h_bad = 0;
    G5_HIP(hipMalloc(&d_bad, 8));
    hipError_t e = hipMemset(d_bad, 0, 8);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_q2bm_checked, g5::grid_for(n, 256, 16384), 256, 0, nullptr, d_queue, n, d_bits, n_words,
                           d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h_bad, d_bad, 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_bad);
    G5_HIP(e);
    G5_REQUIRE(h_bad == 0, G5_E_ARG, "g5_queue_to_bitmap: %llu ids beyond %lld words", h_bad, (long long)n_words);
    return G5_OK;
}
