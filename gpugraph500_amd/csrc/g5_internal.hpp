// g5_internal.hpp -- shared host/device definitions of the gfx950 library.
//
// Layout in HBM (one rank's DistMatrix2d block, cpu_2d/distmatrix2d.hh:30-130):
//   rowptr  int64[n_rows + 1]   CSR row offsets
//   cols    uint32[nnz]         local column ids, ascending and unique per row
//   gdeg    uint32[n_rows]      generated directed degree (duplicates and
//                               self-loops counted) -- the TEPS numerator
// Bitmaps are uint64 words, bit (v & 63) of word (v >> 6); n_rows and n_cols
// are padded to multiples of 64 as the reference's ALG = 64 does.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "g5hip.h"

namespace g5 {

void set_error(const char *fmt, ...);
int fail_hip(hipError_t e, const char *what, const char *file, int line);

}  // namespace g5

#define G5_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return g5::fail_hip(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define G5_REQUIRE(cond, code, ...)         \
    do {                                            \
        if (!(cond)) {                              \
            g5::set_error(__VA_ARGS__);             \
            return (code);                          \
        }                                           \
    } while (0)

struct g5_graph {
    int device = 0;
    int scale = 0;          // 0 when uploaded from a host CSR
    int64_t n_rows = 0;     // local rows, multiple of 64
    int64_t n_cols = 0;     // local columns, multiple of 64
    int64_t nnz = 0;
    int64_t nglobal = 0;    // global vertex count (max id + 1)
    int64_t n_gen_edges = 0;
    int64_t row_start = 0;  // global id of local row 0
    int64_t col_start = 0;  // global id of local column 0
    int64_t *d_rowptr = nullptr;
    uint32_t *d_cols = nullptr;
    uint32_t *d_gdeg = nullptr;
};

namespace g5 {

constexpr int kWave = 64;

// ---- device-wide primitives (g5_util.hip) ----
// Exclusive scan of n int64 values in place; returns the total through *d_total
// (device pointer, may be null).  scratch must hold scan_scratch_words(n) int64.
size_t scan_scratch_words(int64_t n);
hipError_t exclusive_scan_i64(int64_t *d_data, int64_t n, int64_t *d_total, int64_t *scratch,
                              hipStream_t s);
// Bitmap -> sorted queue of set-bit indices (+ base).  Writes the count to
// *d_count.  scratch as for exclusive_scan_i64 over n_words.
hipError_t bitmap_to_queue(const uint64_t *d_bits, int64_t n_words, uint32_t *d_queue,
                           int64_t *d_count, int64_t *scratch, hipStream_t s);
// Queue -> bitmap (bitmap must be cleared first).
hipError_t queue_to_bitmap(const uint32_t *d_queue, int64_t n, uint64_t *d_bits, hipStream_t s);
hipError_t fill_u64(uint64_t *d, int64_t n, uint64_t v, hipStream_t s);
hipError_t fill_u32(uint32_t *d, int64_t n, uint32_t v, hipStream_t s);
hipError_t fill_i32(int32_t *d, int64_t n, int32_t v, hipStream_t s);
hipError_t fill_i64(int64_t *d, int64_t n, int64_t v, hipStream_t s);
int grid_for(int64_t n, int block, int max_grid = 1 << 20);
// Once per device, with that device current: the device's address of the
// static scratch of g5_s4_roundtrip_noalloc (g5_noalloc.hip).
int noalloc_setup(int device);
// Device-time accounting of the exchange entry points (g5_exchange_profile).
// An ExchangeTimer on stream s is live only while accounting is on (else
// every call does nothing and begin()/end() are null): start() opens an
// interval (an event on s), stop() closes it (a second event, no wait), and
// commit(unit, bytes, calls) waits for the closing event and adds the
// interval, the algorithmic bytes and the calls to the unit's totals.  A
// single kernel is timed exactly by passing begin() and end() to
// hipExtLaunchKernelGGL instead of start() / stop() (events on an idle
// stream would also count the launch's host latency).  An interval that
// spans a host round trip is split in two: start, stop, commit(unit, 0, 0),
// start, stop, commit.
class ExchangeTimer
{
public:
    explicit ExchangeTimer(hipStream_t s);
    ~ExchangeTimer();
    hipError_t start();
    hipError_t stop();
    hipError_t commit(int unit, double bytes, int64_t calls);
    hipEvent_t begin() const { return e0_; }
    hipEvent_t end() const { return e1_; }

private:
    hipStream_t s_;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
};
// Ascending radix sort of n keys on bits [0, end_bit); tmp holds n keys.
// The result ends in keys.
hipError_t sort_u64(uint64_t *keys, uint64_t *tmp, int64_t n, int end_bit, hipStream_t s);

}  // namespace g5

// ---- device helpers ----
__device__ __forceinline__ uint32_t g5_lane() { return __lane_id(); }

__device__ __forceinline__ int64_t g5_wave_incl_scan(int64_t v)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int64_t t = __shfl_up(v, d, 64);
        if ((int)__lane_id() >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t g5_wave_or(uint64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ uint32_t g5_wave_or32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ int64_t g5_wave_sum(int64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// The same modulo 2^64.
__device__ __forceinline__ uint64_t g5_block_sum256_u64(uint64_t v)
{
    __shared__ uint64_t s_part_u[4];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (__lane_id() == 0) s_part_u[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint64_t t = s_part_u[0] + s_part_u[1] + s_part_u[2] + s_part_u[3];
    __syncthreads();
    return t;
}

// 64-bit finaliser of SplitMix64 (Steele, Lea, Flood 2014; constants as
// published by Vigna): a bijective mix, used for predecessor-map checksums
// sum over reached v of mix64(v << 32 | (pred(v) + 1)) mod 2^64.
__device__ __forceinline__ uint64_t g5_mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Sum of v over a 256-thread workgroup, returned to every thread (every
// thread must call it).  Kernels add it to a global counter from one thread:
// same-address atomics serialise in one L2 channel, so one per workgroup
// rather than one per wave.
__device__ __forceinline__ int64_t g5_block_sum256(int64_t v)
{
    __shared__ int64_t s_part[4];
    v = g5_wave_sum(v);
    if (__lane_id() == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    const int64_t t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    __syncthreads();  // s_part is free again for the next call
    return t;
}
