// g5_vreduce.hip -- the column reduction's combine step on the device.
//
// vreduce (cpu_2d/vreduce.hpp:41-544) folds the R column contributions of a
// level with the bitmap functor (cpu_2d/cpubfs_bin.cpp:55-65, a word OR);
// allReduceBitCompressed (cpu_2d/globalbfs.hh:235-524) merges predecessor
// lists, where this build takes the minimum present id (-1 = absent) instead
// of the reference's rank precedence, so the combined parent is the
// canonical minimum-id parent whatever the arrival order.  On one GPU the
// contributions a rank would receive are laid out contiguously, or named by
// a list of up to kMaxList pointers (the _list forms, passed by value as a
// kernel argument), and combined in one pass: k reads + 1 write per element,
// 16-byte loads, grid-stride.
#include <hip/hip_ext.h>

#include "g5_internal.hpp"

namespace {

__global__ __launch_bounds__(256) void k_or(const uint64_t *__restrict__ maps, int k, int64_t n,
                                            uint64_t *__restrict__ out)
{
    const int64_t n2 = n / 2;
    const ulonglong2 *m2 = reinterpret_cast<const ulonglong2 *>(maps);
    ulonglong2 *o2 = reinterpret_cast<ulonglong2 *>(out);
    const bool vec = (n % 2) == 0;  // every map starts 16-byte aligned
    if (vec) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
             i += (int64_t)gridDim.x * blockDim.x) {
            ulonglong2 a = m2[i];
            for (int j = 1; j < k; ++j) {
                const ulonglong2 b = m2[(int64_t)j * n2 + i];
                a.x |= b.x;
                a.y |= b.y;
            }
            o2[i] = a;
        }
    } else {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            uint64_t a = maps[i];
            for (int j = 1; j < k; ++j) a |= maps[(int64_t)j * n + i];
            out[i] = a;
        }
    }
}

__device__ __forceinline__ uint64_t pkey(int64_t p) { return p < 0 ? ~0ull : (uint64_t)p; }

__global__ __launch_bounds__(256) void k_min_pred(const int64_t *__restrict__ preds, int k, int64_t n,
                                                  int64_t *__restrict__ out)
{
    const int64_t n2 = n / 2;
    const bool vec = (n % 2) == 0;
    if (vec) {
        const longlong2 *p2 = reinterpret_cast<const longlong2 *>(preds);
        longlong2 *o2 = reinterpret_cast<longlong2 *>(out);
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
             i += (int64_t)gridDim.x * blockDim.x) {
            longlong2 a = p2[i];
            uint64_t x = pkey(a.x), y = pkey(a.y);
            for (int j = 1; j < k; ++j) {
                const longlong2 b = p2[(int64_t)j * n2 + i];
                x = min(x, pkey(b.x));
                y = min(y, pkey(b.y));
            }
            o2[i] = make_longlong2(x == ~0ull ? -1 : (int64_t)x, y == ~0ull ? -1 : (int64_t)y);
        }
    } else {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            uint64_t x = pkey(preds[i]);
            for (int j = 1; j < k; ++j) x = min(x, pkey(preds[(int64_t)j * n + i]));
            out[i] = x == ~0ull ? -1 : (int64_t)x;
        }
    }
}

constexpr int kMaxList = 64;
template <typename T>
struct PtrList {
    const T *p[kMaxList];
};

__global__ __launch_bounds__(256) void k_or_list(PtrList<uint64_t> m, int k, int64_t n, uint64_t *__restrict__ out)
{
    const bool vec = (n % 2) == 0;  // every map starts 16-byte aligned
    if (vec) {
        const int64_t n2 = n / 2;
        ulonglong2 *o2 = reinterpret_cast<ulonglong2 *>(out);
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
             i += (int64_t)gridDim.x * blockDim.x) {
            ulonglong2 a = reinterpret_cast<const ulonglong2 *>(m.p[0])[i];
            for (int j = 1; j < k; ++j) {
                const ulonglong2 b = reinterpret_cast<const ulonglong2 *>(m.p[j])[i];
                a.x |= b.x;
                a.y |= b.y;
            }
            o2[i] = a;
        }
    } else {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            uint64_t a = m.p[0][i];
            for (int j = 1; j < k; ++j) a |= m.p[j][i];
            out[i] = a;
        }
    }
}

__global__ __launch_bounds__(256) void k_min_pred_list(PtrList<int64_t> m, int k, int64_t n,
                                                       int64_t *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t x = pkey(m.p[0][i]);
        for (int j = 1; j < k; ++j) x = min(x, pkey(m.p[j][i]));
        out[i] = x == ~0ull ? -1 : (int64_t)x;
    }
}

}  // namespace

namespace g5 {

// e0 / e1 (both or neither): events the runtime records at the kernel's
// start and end (hipExtLaunchKernelGGL), for the exchange accounting
hipError_t vreduce_or(const uint64_t *d_maps, int k, int64_t n, uint64_t *d_out, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1)
{
    hipExtLaunchKernelGGL(k_or, grid_for(n / 2 + 1, 256, 8192), 256, 0, s, e0, e1, 0, d_maps, k, n, d_out);
    return hipGetLastError();
}

hipError_t vreduce_min_pred(const int64_t *d_preds, int k, int64_t n, int64_t *d_out, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1)
{
    hipExtLaunchKernelGGL(k_min_pred, grid_for(n / 2 + 1, 256, 8192), 256, 0, s, e0, e1, 0, d_preds, k, n, d_out);
    return hipGetLastError();
}

}  // namespace g5

extern "C" int g5_vreduce_or(const uint64_t *d_maps, int k, int64_t n_words, uint64_t *d_out)
{
    G5_REQUIRE(d_maps && d_out && k >= 1 && n_words >= 0, G5_E_ARG, "g5_vreduce_or: bad arguments");
    G5_REQUIRE(((uintptr_t)d_maps & 15) == 0 && ((uintptr_t)d_out & 15) == 0, G5_E_ARG,
               "g5_vreduce_or: buffers must be 16-byte aligned");
    if (n_words == 0) return G5_OK;
    g5::ExchangeTimer t(nullptr);
    G5_HIP(g5::vreduce_or(d_maps, k, n_words, d_out, nullptr, t.begin(), t.end()));
    G5_HIP(hipStreamSynchronize(nullptr));
    G5_HIP(t.commit(G5_UNIT_OR, 8.0 * (k + 1) * n_words, 1));
    return G5_OK;
}

extern "C" int g5_vreduce_min_pred(const int64_t *d_preds, int k, int64_t n, int64_t *d_out)
{
    G5_REQUIRE(d_preds && d_out && k >= 1 && n >= 0, G5_E_ARG, "g5_vreduce_min_pred: bad arguments");
    G5_REQUIRE(((uintptr_t)d_preds & 15) == 0 && ((uintptr_t)d_out & 15) == 0, G5_E_ARG,
               "g5_vreduce_min_pred: buffers must be 16-byte aligned");
    if (n == 0) return G5_OK;
    g5::ExchangeTimer t(nullptr);
    G5_HIP(g5::vreduce_min_pred(d_preds, k, n, d_out, nullptr, t.begin(), t.end()));
    G5_HIP(hipStreamSynchronize(nullptr));
    G5_HIP(t.commit(G5_UNIT_MIN_PRED, 8.0 * (k + 1) * n, 1));
    return G5_OK;
}

extern "C" int g5_vreduce_or_list(const uint64_t *const *d_maps, int k, int64_t n_words, uint64_t *d_out)
{
    G5_REQUIRE(d_maps && d_out && k >= 1 && k <= kMaxList && n_words >= 0, G5_E_ARG,
               "g5_vreduce_or_list: bad arguments (1 <= k <= %d)", kMaxList);
    PtrList<uint64_t> m{};
    for (int j = 0; j < k; ++j) {
        G5_REQUIRE(d_maps[j] && ((uintptr_t)d_maps[j] & 15) == 0, G5_E_ARG,
                   "g5_vreduce_or_list: map %d null or not 16-byte aligned", j);
        m.p[j] = d_maps[j];
    }
    G5_REQUIRE(((uintptr_t)d_out & 15) == 0, G5_E_ARG, "g5_vreduce_or_list: output must be 16-byte aligned");
    if (n_words == 0) return G5_OK;
    g5::ExchangeTimer t(nullptr);
    hipExtLaunchKernelGGL(k_or_list, g5::grid_for(n_words / 2 + 1, 256, 8192), 256, 0, nullptr, t.begin(), t.end(), 0,
                          m, k, n_words, d_out);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(nullptr));
    G5_HIP(t.commit(G5_UNIT_OR, 8.0 * (k + 1) * n_words, 1));
    return G5_OK;
}

extern "C" int g5_vreduce_min_pred_list(const int64_t *const *d_preds, int k, int64_t n, int64_t *d_out)
{
    G5_REQUIRE(d_preds && d_out && k >= 1 && k <= kMaxList && n >= 0, G5_E_ARG,
               "g5_vreduce_min_pred_list: bad arguments (1 <= k <= %d)", kMaxList);
    PtrList<int64_t> m{};
    for (int j = 0; j < k; ++j) {
        G5_REQUIRE(d_preds[j], G5_E_ARG, "g5_vreduce_min_pred_list: array %d null", j);
        m.p[j] = d_preds[j];
    }
    if (n == 0) return G5_OK;
    g5::ExchangeTimer t(nullptr);
    hipExtLaunchKernelGGL(k_min_pred_list, g5::grid_for(n, 256, 8192), 256, 0, nullptr, t.begin(), t.end(), 0, m, k,
                          n, d_out);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(nullptr));
    G5_HIP(t.commit(G5_UNIT_MIN_PRED, 8.0 * (k + 1) * n, 1));
    return G5_OK;
}
