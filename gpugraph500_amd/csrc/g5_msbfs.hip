// g5_msbfs.hip -- batched multi-source BFS: up to 64 roots searched together
// with one uint64 lane mask per vertex (bit r = root r).
//
// Per level every vertex v with an unvisited lane pulls from its neighbours
// in ascending id order: f = frontier[u] & ~visited[v].  The first neighbour
// that shows lane r is v's parent for root r -- the minimum-id vertex of the
// previous level, the parent CPUBFS_bin's ascending scan keeps
// (cpu_2d/cpubfs_bin.cpp:150-173) -- so the 64 results equal 64 separate
// single-root searches.  The scan stops as soon as every unvisited lane is
// found.  A vertex is only written by the work item that owns it, so the
// level needs no atomics on the masks: visited[v] |= found, next[v] = found.
//
// Work is binned statically by degree (per graph, once): one thread per
// vertex up to 32 neighbours, one wavefront (64 neighbours per step, first
// hits per lane from an exclusive prefix-OR across lanes) up to 2048, one
// 256-thread workgroup beyond.  Each level's launch sequence is captured once
// into a hipGraph per frontier/next parity and replayed; the level number
// lives in device memory.
//
// Output layout in HBM: parent and level int32[n][64] (vertex-major: a
// vertex's 64 lanes share one 256-byte line).  Per-root arrays are gathered
// by g5_msbfs_result.
#include <algorithm>

#include "g5_internal.hpp"

namespace {

constexpr int kSmallMax = 32;
constexpr int kMediumMax = 2048;
constexpr int kBlock = 256;

// counters: per degree class c (0 small, 1 medium, 2 large) slots 3c+0 new
// (vertex, root) pairs, 3c+1 edges read, 3c+2 vertices with an unvisited lane
enum MsCnt { kNewBits = 0, kEdgesRead = 1, kVisitedVerts = 2, kDepth = 9, kMsN = 16 };

struct Lists {
    uint32_t *small = nullptr, *medium = nullptr, *large = nullptr;
    int64_t n_small = 0, n_medium = 0, n_large = 0;
};

}  // namespace

struct g5_msbfs {
    g5_graph *g = nullptr;
    hipStream_t s = nullptr;
    int64_t n = 0;                       // vertices (rows == cols)
    uint64_t *d_visited = nullptr, *d_fa = nullptr, *d_fb = nullptr;
    int32_t *d_parent = nullptr, *d_level = nullptr;
    int64_t *d_cnt = nullptr, *h_cnt = nullptr;
    int64_t *d_scratch = nullptr;
    int64_t *d_roots = nullptr;
    Lists lists;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nroots = 0;
    int64_t roots[64];
    // accounting
    bool prof = false;
    hipEvent_t ek[4] = {nullptr, nullptr, nullptr, nullptr};
    double k_ms[3] = {0, 0, 0}, k_bytes[3] = {0, 0, 0};
    int64_t k_launches[3] = {0, 0, 0};
};

namespace {

__device__ __forceinline__ void write_parents(int32_t *__restrict__ parent, int64_t v, uint64_t bits, uint32_t u)
{
    while (bits) {
        const int r = __ffsll((long long)bits) - 1;
        parent[v * 64 + r] = (int32_t)u;
        bits &= bits - 1;
    }
}

// exclusive prefix OR over the lanes of a wave
__device__ __forceinline__ uint64_t lanes_excl_or(uint64_t f)
{
    uint64_t inc = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(inc, d, 64);
        if ((int)__lane_id() >= d) inc |= t;
    }
    const uint64_t prev = __shfl_up(inc, 1, 64);
    return __lane_id() == 0 ? 0ull : prev;
}

// One thread per vertex (degree <= kSmallMax).
__global__ __launch_bounds__(kBlock) void k_pull_thread(
    const uint32_t *__restrict__ list, int64_t nlist, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    uint64_t *__restrict__ next, int32_t *__restrict__ parent, int32_t *__restrict__ level,
    int64_t *__restrict__ cnt_all)
{
    const int32_t depth = (int32_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 0;
    int64_t nbits = 0, nedges = 0, nverts = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nlist;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = list[i];
        const uint64_t vis = visited[v];
        if (vis == ~0ull) {
            next[v] = 0;
            continue;
        }
        const uint64_t unv = ~vis;
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t j = b;
        for (; j < e; ++j) {
            const uint32_t u = cols[j];
            const uint64_t f = frontier[u] & unv & ~found;
            if (f) {
                found |= f;
                write_parents(parent, v, f, u);
                if (found == unv) {
                    ++j;
                    break;
                }
            }
        }
        next[v] = found;
        nedges += j - b;
        ++nverts;
        if (found) {
            visited[v] = vis | found;
            nbits += __popcll(found);
            uint64_t bits = found;
            while (bits) {
                const int r = __ffsll((long long)bits) - 1;
                level[(int64_t)v * 64 + r] = depth;
                bits &= bits - 1;
            }
        }
    }
    nbits = g5_wave_sum(nbits);
    nedges = g5_wave_sum(nedges);
    nverts = g5_wave_sum(nverts);
    if (__lane_id() == 0) {
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

// One wavefront per vertex (kSmallMax < degree <= kMediumMax).
__global__ __launch_bounds__(kBlock) void k_pull_wave(
    const uint32_t *__restrict__ list, int64_t nlist, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    uint64_t *__restrict__ next, int32_t *__restrict__ parent, int32_t *__restrict__ level,
    int64_t *__restrict__ cnt_all)
{
    const int32_t depth = (int32_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 3;
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
    int64_t nbits = 0, nedges = 0, nverts = 0;
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < nlist; i += waves) {
        const uint32_t v = list[i];
        const uint64_t vis = visited[v];
        if (vis == ~0ull) {
            if (lane == 0) next[v] = 0;
            continue;
        }
        const uint64_t unv = ~vis;
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t base = b;
        for (; base < e; base += 64) {
            const int64_t j = base + lane;
            uint32_t u = 0;
            uint64_t f = 0;
            if (j < e) {
                u = cols[j];
                f = frontier[u] & unv;
            }
            const uint64_t first = f & ~lanes_excl_or(f) & ~found;
            if (first) write_parents(parent, v, first, u);
            found |= g5_wave_or(f);
            if (found == unv) {
                base += 64;
                break;
            }
        }
        if (found >> lane & 1) level[(int64_t)v * 64 + lane] = depth;
        if (lane == 0) {
            next[v] = found;
            if (found) visited[v] = vis | found;
        }
        nbits += __popcll(found);
        nedges += min(base, e) - b;
        ++nverts;
    }
    if (lane == 0) {
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

// One workgroup per vertex (degree > kMediumMax): 256 neighbours per round.
__global__ __launch_bounds__(kBlock) void k_pull_block(
    const uint32_t *__restrict__ list, int64_t nlist, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    uint64_t *__restrict__ next, int32_t *__restrict__ parent, int32_t *__restrict__ level,
    int64_t *__restrict__ cnt_all)
{
    __shared__ uint64_t s_wave[kBlock / 64];
    __shared__ uint64_t s_vis;
    const int32_t depth = (int32_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 6;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    int64_t nbits = 0, nedges = 0, nverts = 0;
    for (int64_t i = blockIdx.x; i < nlist; i += gridDim.x) {
        const uint32_t v = list[i];
        if (threadIdx.x == 0) s_vis = visited[v];
        __syncthreads();
        const uint64_t vis = s_vis;
        if (vis == ~0ull) {
            if (threadIdx.x == 0) next[v] = 0;
            __syncthreads();
            continue;
        }
        const uint64_t unv = ~vis;
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t base = b;
        for (; base < e; base += kBlock) {
            const int64_t j = base + threadIdx.x;
            uint32_t u = 0;
            uint64_t f = 0;
            if (j < e) {
                u = cols[j];
                f = frontier[u] & unv;
            }
            const uint64_t in_wave = lanes_excl_or(f);
            const uint64_t wave_all = __shfl(in_wave | f, 63, 64);
            if (lane == 0) s_wave[wv] = wave_all;
            __syncthreads();
            uint64_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) {
                const uint64_t t = s_wave[w];
                before |= (w < wv) ? t : 0ull;
                all |= t;
            }
            const uint64_t first = f & ~(in_wave | before) & ~found;
            if (first) write_parents(parent, v, first, u);
            found |= all;
            __syncthreads();
            if (found == unv) {
                base += kBlock;
                break;
            }
        }
        if (threadIdx.x < 64 && (found >> threadIdx.x & 1)) level[(int64_t)v * 64 + threadIdx.x] = depth;
        if (threadIdx.x == 0) {
            next[v] = found;
            if (found) visited[v] = vis | found;
            nbits += __popcll(found);
            nedges += min(base, e) - b;
            ++nverts;
        }
    }
    if (threadIdx.x == 0) {
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

__global__ void k_level_begin(int64_t *cnt)
{
    for (int i = 0; i < 9; ++i) cnt[i] = 0;
}

__global__ void k_level_end(int64_t *cnt) { cnt[kDepth] += 1; }

__global__ void k_ms_init(uint64_t *__restrict__ visited, uint64_t *__restrict__ fa, uint64_t *__restrict__ fb,
                          int64_t n, uint64_t unused_lanes)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        visited[i] = unused_lanes;
        fa[i] = 0;
        fb[i] = 0;
    }
}

__global__ void k_ms_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    const int64_t n4 = n / 4;
    int4 *p4 = reinterpret_cast<int4 *>(p);
    const int4 q = make_int4(v, v, v, v);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        p4[i] = q;
}

// Roots: one thread; lanes of duplicate roots share the vertex.
__global__ void k_ms_roots(const int64_t *__restrict__ roots, int nroots, uint64_t *__restrict__ visited,
                           uint64_t *__restrict__ frontier, int32_t *__restrict__ parent,
                           int32_t *__restrict__ level, int64_t *__restrict__ cnt)
{
    for (int r = 0; r < nroots; ++r) {
        const int64_t v = roots[r];
        visited[v] |= 1ull << r;
        frontier[v] |= 1ull << r;
        parent[v * 64 + r] = (int32_t)v;
        level[v * 64 + r] = 0;
    }
    cnt[kDepth] = 1;
}

// Classification of vertices with nonzero degree into three bitmaps.
__global__ void k_classify(const int64_t *__restrict__ rowptr, int64_t nwords, uint64_t *__restrict__ bs,
                           uint64_t *__restrict__ bm, uint64_t *__restrict__ bl)
{
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nwords;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint64_t s = 0, m = 0, l = 0;
        for (int i = 0; i < 64; ++i) {
            const int64_t v = w * 64 + i;
            const int64_t d = rowptr[v + 1] - rowptr[v];
            if (d <= 0) continue;
            if (d <= kSmallMax) s |= 1ull << i;
            else if (d <= kMediumMax) m |= 1ull << i;
            else l |= 1ull << i;
        }
        bs[w] = s;
        bm[w] = m;
        bl[w] = l;
    }
}

// per-lane extraction: pred/level of root `lane` as int64/int32[n]
__global__ void k_ms_extract(const int32_t *__restrict__ parent, const int32_t *__restrict__ level, int64_t n,
                             int lane, int64_t *__restrict__ pred, int32_t *__restrict__ lev)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t l = level[i * 64 + lane];
        lev[i] = l;
        pred[i] = l < 0 ? -1 : (int64_t)parent[i * 64 + lane];
    }
}

// nedge per lane: sum of generated degrees over reached vertices
__global__ __launch_bounds__(256) void k_ms_nedge(const int32_t *__restrict__ level, const uint32_t *__restrict__ gdeg,
                                                  int64_t n, unsigned long long *__restrict__ out)
{
    // thread t of a wave handles lane t for a strip of vertices
    const int lane = __lane_id();
    unsigned long long acc = 0;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t v = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6); v < n; v += waves) {
        const uint32_t d = gdeg[v];
        if (d && level[v * 64 + lane] >= 0) acc += d;
    }
    atomicAdd(&out[lane], acc);
}

// prof: events between the three pull kernels (the uncaptured path)
int level_launches(g5_msbfs *m, hipStream_t s, const uint64_t *front, uint64_t *next, bool prof = false)
{
    g5_graph *g = m->g;
    hipLaunchKernelGGL(k_level_begin, 1, 1, 0, s, m->d_cnt);
    if (prof) G5_HIP(hipEventRecord(m->ek[0], s));
    if (m->lists.n_large)
        hipLaunchKernelGGL(k_pull_block, (unsigned)std::min<int64_t>(m->lists.n_large, 4096), kBlock, 0, s,
                           (const uint32_t *)m->lists.large, m->lists.n_large, (const int64_t *)g->d_rowptr,
                           (const uint32_t *)g->d_cols, m->d_visited, front, next, m->d_parent, m->d_level,
                           m->d_cnt);
    if (prof) G5_HIP(hipEventRecord(m->ek[1], s));
    if (m->lists.n_medium)
        hipLaunchKernelGGL(k_pull_wave, g5::grid_for((m->lists.n_medium + 3) / 4, 1, 16384), kBlock, 0, s,
                           (const uint32_t *)m->lists.medium, m->lists.n_medium, (const int64_t *)g->d_rowptr,
                           (const uint32_t *)g->d_cols, m->d_visited, front, next, m->d_parent, m->d_level,
                           m->d_cnt);
    if (prof) G5_HIP(hipEventRecord(m->ek[2], s));
    if (m->lists.n_small)
        hipLaunchKernelGGL(k_pull_thread, g5::grid_for(m->lists.n_small, kBlock, 16384), kBlock, 0, s,
                           (const uint32_t *)m->lists.small, m->lists.n_small, (const int64_t *)g->d_rowptr,
                           (const uint32_t *)g->d_cols, m->d_visited, front, next, m->d_parent, m->d_level,
                           m->d_cnt);
    if (prof) G5_HIP(hipEventRecord(m->ek[3], s));
    hipLaunchKernelGGL(k_level_end, 1, 1, 0, s, m->d_cnt);
    return hipGetLastError() == hipSuccess ? G5_OK : G5_E_HIP;
}

int capture_graphs(g5_msbfs *m)
{
    for (int p = 0; p < 2; ++p) {
        const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
        uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
        hipGraph_t gr = nullptr;
        G5_HIP(hipStreamBeginCapture(m->s, hipStreamCaptureModeThreadLocal));
        int rc = level_launches(m, m->s, front, next);
        hipError_t e = hipStreamEndCapture(m->s, &gr);
        if (rc) return rc;
        G5_HIP(e);
        e = hipGraphInstantiate(&m->graph[p], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        G5_HIP(e);
    }
    return G5_OK;
}

int build_lists(g5_msbfs *m)
{
    g5_graph *g = m->g;
    hipStream_t s = m->s;
    const int64_t nw = m->n / 64;
    uint64_t *bits = nullptr;
    G5_HIP(hipMalloc(&bits, (size_t)std::max<int64_t>(nw, 1) * 8 * 3));
    hipLaunchKernelGGL(k_classify, g5::grid_for(nw, 256, 8192), 256, 0, s, (const int64_t *)g->d_rowptr, nw,
                       bits, bits + nw, bits + 2 * nw);
    uint32_t **dst[3] = {&m->lists.small, &m->lists.medium, &m->lists.large};
    int64_t *cntp[3] = {&m->lists.n_small, &m->lists.n_medium, &m->lists.n_large};
    for (int c = 0; c < 3; ++c) {
        uint32_t *tmp = nullptr;
        hipError_t e = hipMalloc(&tmp, (size_t)std::max<int64_t>(m->n, 1) * 4);
        if (e == hipSuccess) e = g5::bitmap_to_queue(bits + c * nw, nw, tmp, m->d_cnt, m->d_scratch, s);
        if (e == hipSuccess) e = hipMemcpyAsync(m->h_cnt, m->d_cnt, 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            (void)hipFree(bits);
            return g5::fail_hip(e, "msbfs lists", __FILE__, __LINE__);
        }
        *cntp[c] = m->h_cnt[0];
        e = hipMalloc(dst[c], (size_t)std::max<int64_t>(*cntp[c], 1) * 4);
        if (e == hipSuccess && *cntp[c])
            e = hipMemcpyAsync(*dst[c], tmp, (size_t)*cntp[c] * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        if (e != hipSuccess) {
            (void)hipFree(bits);
            return g5::fail_hip(e, "msbfs lists", __FILE__, __LINE__);
        }
    }
    (void)hipFree(bits);
    return G5_OK;
}

}  // namespace

extern "C" int g5_msbfs_free(g5_msbfs *m);

extern "C" int g5_msbfs_create(g5_graph *g, g5_msbfs **out)
{
    G5_REQUIRE(g && out, G5_E_ARG, "g5_msbfs_create: null argument");
    G5_REQUIRE(g->n_rows == g->n_cols && g->row_start == g->col_start, G5_E_ARG,
               "g5_msbfs_create: needs a diagonal (1x1) block");
    g5_msbfs *m = new g5_msbfs();
    m->g = g;
    m->n = g->n_rows;
    const int64_t n = std::max<int64_t>(m->n, 64);
    auto fail = [&](int rc) {
        g5_msbfs_free(m);
        return rc;
    };
    hipError_t e = hipStreamCreate(&m->s);
    if (e == hipSuccess) e = hipMalloc(&m->d_visited, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_fa, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_fb, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_parent, (size_t)n * 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_level, (size_t)n * 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_cnt, kMsN * 8);
    if (e == hipSuccess) e = hipHostMalloc(&m->h_cnt, kMsN * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&m->d_scratch, (g5::scan_scratch_words(n / 64 + 1) + 16) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_roots, 64 * 8);
    if (e == hipSuccess) e = hipEventCreate(&m->ev0);
    if (e == hipSuccess) e = hipEventCreate(&m->ev1);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&m->ek[k]);
    if (e != hipSuccess) return fail(g5::fail_hip(e, "g5_msbfs_create allocation", __FILE__, __LINE__));
    int rc = build_lists(m);
    if (rc) return fail(rc);
    rc = capture_graphs(m);
    if (rc) return fail(rc);
    *out = m;
    return G5_OK;
}

extern "C" int g5_msbfs_free(g5_msbfs *m)
{
    if (!m) return G5_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    for (int p = 0; p < 2; ++p)
        if (m->graph[p]) (void)hipGraphExecDestroy(m->graph[p]);
    void *dev[] = {m->d_visited, m->d_fa,        m->d_fb,           m->d_parent,       m->d_level,
                   m->d_cnt,     m->d_scratch,   m->d_roots,        m->lists.small,    m->lists.medium,   m->lists.large};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (m->h_cnt) (void)hipHostFree(m->h_cnt);
    hipEvent_t evs[] = {m->ev0, m->ev1, m->ek[0], m->ek[1], m->ek[2], m->ek[3]};
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return G5_OK;
}

extern "C" int g5_msbfs_run(g5_msbfs *m, const int64_t *roots, int nroots, int32_t *nlevels, float *ms)
{
    G5_REQUIRE(m && roots && nroots >= 1 && nroots <= 64, G5_E_ARG, "g5_msbfs_run: 1..64 roots required");
    for (int r = 0; r < nroots; ++r)
        G5_REQUIRE(roots[r] >= 0 && roots[r] < m->n, G5_E_ARG, "g5_msbfs_run: root %lld out of range",
                   (long long)roots[r]);
    hipStream_t s = m->s;
    m->nroots = nroots;
    std::copy(roots, roots + nroots, m->roots);
    int64_t *d_roots = m->d_roots;
    G5_HIP(hipMemcpyAsync(d_roots, roots, (size_t)nroots * 8, hipMemcpyHostToDevice, s));
    G5_HIP(hipStreamSynchronize(s));
    G5_HIP(hipEventRecord(m->ev0, s));
    const uint64_t unused = nroots == 64 ? 0ull : ~((1ull << nroots) - 1);
    hipLaunchKernelGGL(k_ms_init, g5::grid_for(m->n, 256, 16384), 256, 0, s, m->d_visited, m->d_fa, m->d_fb,
                       m->n, unused);
    hipLaunchKernelGGL(k_ms_fill_i32, g5::grid_for(m->n * 16, 256, 16384), 256, 0, s, m->d_level, m->n * 64,
                       (int32_t)-1);
    hipLaunchKernelGGL(k_ms_roots, 1, 1, 0, s, (const int64_t *)d_roots, nroots, m->d_visited, m->d_fa,
                       m->d_parent, m->d_level, m->d_cnt);
    G5_HIP(hipGetLastError());
    int p = 0, levels = 1;
    for (;;) {
        if (m->prof) {
            int rc = level_launches(m, s, p == 0 ? m->d_fa : m->d_fb, p == 0 ? m->d_fb : m->d_fa, true);
            if (rc) return rc;
        } else {
            G5_HIP(hipGraphLaunch(m->graph[p], s));
        }
        G5_HIP(hipMemcpyAsync(m->h_cnt, m->d_cnt, kMsN * 8, hipMemcpyDeviceToHost, s));
        G5_HIP(hipStreamSynchronize(s));
        int64_t new_bits = 0;
        for (int c = 0; c < 3; ++c) new_bits += m->h_cnt[3 * c + kNewBits];
        if (m->prof) {
            // per class c (events ek[2-c] .. ek[3-c]): 16 B row offsets + 8 B
            // visited read + 8 B next write per vertex with an unvisited lane,
            // 4 B column + 8 B frontier word per edge read, 4 B parent + 4 B
            // level per new (vertex, root) pair
            const int order[3] = {2, 1, 0};  // launch order: large, medium, small
            for (int k = 0; k < 3; ++k) {
                const int c = order[k];
                const int64_t nl = c == 0 ? m->lists.n_small : c == 1 ? m->lists.n_medium : m->lists.n_large;
                if (!nl) continue;
                float kms = 0;
                G5_HIP(hipEventElapsedTime(&kms, m->ek[k], m->ek[k + 1]));
                m->k_ms[c] += kms;
                m->k_launches[c] += 1;
                m->k_bytes[c] += 32.0 * m->h_cnt[3 * c + kVisitedVerts] + 12.0 * m->h_cnt[3 * c + kEdgesRead] +
                                 8.0 * m->h_cnt[3 * c + kNewBits];
            }
        }
        if (new_bits == 0) break;
        ++levels;
        p ^= 1;
    }
    G5_HIP(hipEventRecord(m->ev1, s));
    G5_HIP(hipEventSynchronize(m->ev1));
    if (ms) G5_HIP(hipEventElapsedTime(ms, m->ev0, m->ev1));
    if (nlevels) *nlevels = levels;
    return G5_OK;
}

extern "C" int g5_msbfs_result(g5_msbfs *m, int lane, int64_t *pred, int32_t *level)
{
    G5_REQUIRE(m && lane >= 0 && lane < m->nroots, G5_E_ARG, "g5_msbfs_result: lane out of range");
    int64_t *dp = nullptr;
    int32_t *dl = nullptr;
    G5_HIP(hipMalloc(&dp, (size_t)m->n * 8));
    hipError_t e = hipMalloc(&dl, (size_t)m->n * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ms_extract, g5::grid_for(m->n, 256, 16384), 256, 0, m->s,
                           (const int32_t *)m->d_parent, (const int32_t *)m->d_level, m->n, lane, dp, dl);
        e = hipGetLastError();
    }
    if (e == hipSuccess && pred) e = hipMemcpyAsync(pred, dp, (size_t)m->n * 8, hipMemcpyDeviceToHost, m->s);
    if (e == hipSuccess && level) e = hipMemcpyAsync(level, dl, (size_t)m->n * 4, hipMemcpyDeviceToHost, m->s);
    if (e == hipSuccess) e = hipStreamSynchronize(m->s);
    (void)hipFree(dp);
    if (dl) (void)hipFree(dl);
    G5_HIP(e);
    return G5_OK;
}

extern "C" int g5_msbfs_nedge(g5_msbfs *m, int64_t *nedge)
{
    G5_REQUIRE(m && nedge, G5_E_ARG, "g5_msbfs_nedge: null argument");
    unsigned long long *d = nullptr;
    G5_HIP(hipMalloc(&d, 64 * 8));
    hipError_t e = hipMemsetAsync(d, 0, 64 * 8, m->s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ms_nedge, g5::grid_for(m->n / 4 + 1, 1, 8192), 256, 0, m->s,
                           (const int32_t *)m->d_level, (const uint32_t *)m->g->d_gdeg, m->n, d);
        e = hipGetLastError();
    }
    unsigned long long h[64];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, 64 * 8, hipMemcpyDeviceToHost, m->s);
    if (e == hipSuccess) e = hipStreamSynchronize(m->s);
    (void)hipFree(d);
    G5_HIP(e);
    for (int r = 0; r < m->nroots; ++r) nedge[r] = (int64_t)(h[r] / 2);
    return G5_OK;
}

extern "C" int g5_msbfs_profile(g5_msbfs *m, int enable)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_profile: null handle");
    m->prof = enable != 0;
    for (int c = 0; c < 3; ++c) {
        m->k_ms[c] = m->k_bytes[c] = 0;
        m->k_launches[c] = 0;
    }
    return G5_OK;
}

extern "C" int g5_msbfs_stats(g5_msbfs *m, int kernel, double *ms, int64_t *launches, double *bytes)
{
    G5_REQUIRE(m && kernel >= 0 && kernel < 3, G5_E_ARG, "g5_msbfs_stats: bad arguments");
    if (ms) *ms = m->k_ms[kernel];
    if (launches) *launches = m->k_launches[kernel];
    if (bytes) *bytes = m->k_bytes[kernel];
    return G5_OK;
}

extern "C" int g5_msbfs_info(g5_msbfs *m, int64_t *n_small, int64_t *n_medium, int64_t *n_large)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_info: null handle");
    if (n_small) *n_small = m->lists.n_small;
    if (n_medium) *n_medium = m->lists.n_medium;
    if (n_large) *n_large = m->lists.n_large;
    return G5_OK;
}
