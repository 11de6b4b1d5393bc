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
// Sparse frontiers (fewer than nnz/16 frontier edges) are pushed instead:
// next[v] |= frontier[u] & ~visited[v] along the frontier's edges
// (edge-balanced tiles, atomicOr), then the new (vertex, root) parent slots
// are lowered with atomicMin over the same edges -- the same minimum-id
// parent.  A 1-bit-per-vertex active bitmap lets gathers skip vertices with
// an empty frontier word.
//
// Relabelling: the search runs on a copy of the block whose vertices are
// renumbered by descending degree (ties by id).  Frontier words of the hubs,
// which most neighbour lists point at, then sit in a small id range that the
// L2 and the Infinity Cache hold, instead of being scattered over the whole
// n * 8 bytes.  Every row keeps its neighbours in ascending ORIGINAL id
// order and parents are recorded as original ids, so the first hit of a
// scan is still the minimum-id parent.  The degree bins are contiguous id
// ranges: one thread per vertex up to 32 neighbours, one wavefront (64
// neighbours per step, first hits from an exclusive prefix-OR across lanes)
// up to 2048, one 256-thread workgroup beyond.  Each pull level's launch
// sequence is captured once into a hipGraph per frontier parity and
// replayed; the level number lives in device memory.
//
// Output layout in HBM: parent int32[n][64] (original ids) and level
// uint16[n][64] (the reference validator's depth width, validate.h:48-58),
// vertex-major in the relabelled order, valid where the vertex's visited bit
// for that lane is set.  Per-root arrays in original ids are gathered by
// g5_msbfs_result.
#include <algorithm>
#include <vector>

#include "g5_internal.hpp"

namespace {

constexpr int kMediumMax = 2048;
// degree bins in launch order: one workgroup per vertex (> 2048), one
// wavefront per vertex (33..2048), one thread per vertex (1..32)
constexpr int kBins = 3;
constexpr int kBinMaxDeg[kBins] = {1 << 30, 2048, 32};
constexpr int kPushUnit = kBins;  // accounting unit of push levels
constexpr int kUnits = kBins + 1;
constexpr int kBlock = 256;
constexpr int kLevelRec = 12;  // per level: ms[4], edges[4], new pairs, pushed, frontier vertices, edges
// push when the frontier's edges are fewer than nnz / kPushRatio
constexpr int64_t kPushRatio = 16;

// counters: per accounting unit b (the kBins pull bins, then kPushUnit)
// slots 4b+0 new (vertex, root) pairs, 4b+1 edges read, 4b+2 vertices
// worked on, 4b+3 frontier words gathered; then the level number, the next
// frontier's active vertices and their edges, the push queue length, and the
// push frontier queue length / edge total.
enum MsCnt {
    kNewBits = 0, kEdgesRead = 1, kVisitedVerts = 2, kGathers = 3,
    kDepth = 16, kActVerts = 17, kActEdges = 18, kNq = 19, kFqLen = 20, kFqEdges = 21, kMsN = 24
};

// The degree-ordered copy of the block.
struct Relabeled {
    uint32_t *new2old = nullptr, *old2new = nullptr;
    int64_t *rowptr = nullptr;
    uint32_t *cols = nullptr;  // new ids, rows in ascending original-id order
    uint32_t *gdeg = nullptr;
    int64_t bin_first[kBins + 1] = {0};  // bin b holds new ids [bin_first[b], bin_first[b+1])
};

}  // namespace

struct g5_msbfs {
    g5_graph *g = nullptr;
    hipStream_t s = nullptr;
    int64_t n = 0;                       // vertices (rows == cols)
    uint64_t *d_visited = nullptr, *d_fa = nullptr, *d_fb = nullptr;
    uint64_t *d_act = nullptr;           // two n/64-word active bitmaps (parity a, b)
    uint32_t *d_fq = nullptr, *d_nq = nullptr;  // push: frontier queue, newly reached queue
    int64_t *d_qoff = nullptr;                  // push: frontier degree prefix
    int32_t *d_parent = nullptr;
    uint16_t *d_level = nullptr;
    int64_t *d_cnt = nullptr, *h_cnt = nullptr;
    int64_t *d_scratch = nullptr;
    int64_t *d_roots = nullptr;
    Relabeled rl;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nroots = 0;
    int64_t roots[64];
    // accounting
    bool prof = false;
    hipEvent_t ek[kBins + 1] = {};
    double k_ms[kUnits] = {}, k_bytes[kUnits] = {};
    int64_t k_launches[kUnits] = {};
    // per level of the last profiled run (kLevelRec doubles each)
    std::vector<double> lvl;
    int mode = 0;  // 0 direction chosen per level, 1 pull only, 2 push only
};

namespace {

// parent slots of vertex v for the roots in `bits` := u (an original id)
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

// One thread per vertex (degree <= 32): a sequential ascending scan.
__global__ __launch_bounds__(kBlock) void k_pull_thread(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint32_t *__restrict__ new2old, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint16_t *__restrict__ level, int64_t *__restrict__ cnt_all, int bin)
{
    const uint16_t depth = (uint16_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 4 * bin;
    int64_t ngath = 0;
    int64_t nbits = 0, nedges = 0, nverts = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nlist;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(first_v + i);
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
            if (!(active[u >> 6] >> (u & 63) & 1)) continue;
            ++ngath;
            const uint64_t f = frontier[u] & unv & ~found;
            if (f) {
                found |= f;
                write_parents(parent, v, f, new2old[u]);
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
    ngath = g5_wave_sum(ngath);
    if (__lane_id() == 0) {
        if (ngath) atomicAdd((unsigned long long *)&cnt[kGathers], (unsigned long long)ngath);
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

// G lanes per vertex (used with G = 64 for degrees 33..2048): a step loads
// G neighbours at once; the group's exclusive prefix-OR gives each root's
// first hit, the group's inclusive OR the new bits.  Control flow is uniform
// within a group (v, found and the step count are the group's), so the
// width-G shuffles only read lanes of the same group.
template <int G>
__global__ __launch_bounds__(kBlock) void k_pull_group(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint32_t *__restrict__ new2old, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint16_t *__restrict__ level, int64_t *__restrict__ cnt_all, int bin)
{
    constexpr int kGroups = kBlock / G;
    const uint16_t depth = (uint16_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 4 * bin;
    const int gl = threadIdx.x & (G - 1);
    const int gbase = (int)(__lane_id() & ~(G - 1));
    const int64_t stride = (int64_t)gridDim.x * kGroups;
    int64_t nbits = 0, nedges = 0, nverts = 0, ngath = 0;
    for (int64_t i = blockIdx.x * (int64_t)kGroups + threadIdx.x / G; i < nlist; i += stride) {
        const uint32_t v = (uint32_t)(first_v + i);
        const uint64_t vis = visited[v];
        if (vis == ~0ull) {
            if (gl == 0) next[v] = 0;
            continue;
        }
        const uint64_t unv = ~vis;
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t base = b;
        for (; base < e; base += G) {
            const int64_t j = base + gl;
            uint32_t u = 0;
            uint64_t fresh = 0;
            if (j < e) {
                u = cols[j];
                if (active[u >> 6] >> (u & 63) & 1) {
                    fresh = frontier[u] & unv & ~found;
                    ++ngath;
                }
            }
            const uint64_t bal = __ballot(fresh != 0);
            const uint64_t gmask = G == 64 ? bal : (bal >> gbase) & ((1ull << (G & 63)) - 1);
            if (gmask == 0) continue;
            uint64_t inc = fresh;
#pragma unroll
            for (int d = 1; d < G; d <<= 1) {
                const uint64_t t = __shfl_up(inc, d, G);
                if (gl >= d) inc |= t;
            }
            const uint64_t prev = __shfl_up(inc, 1, G);
            const uint64_t first = fresh & ~(gl ? prev : 0ull);
            if (first) write_parents(parent, v, first, new2old[u]);
            found |= __shfl(inc, G - 1, G);
            if (found == unv) {
                base += G;
                break;
            }
        }
        if (found) {
#pragma unroll
            for (int r = gl; r < 64; r += G)
                if (found >> r & 1) level[(int64_t)v * 64 + r] = depth;
        }
        if (gl == 0) {
            next[v] = found;
            if (found) visited[v] = vis | found;
            nbits += __popcll(found);
            nedges += min(base, e) - b;
            ++nverts;
        }
    }
    nbits = g5_wave_sum(nbits);
    nedges = g5_wave_sum(nedges);
    nverts = g5_wave_sum(nverts);
    ngath = g5_wave_sum(ngath);
    if (__lane_id() == 0) {
        if (ngath) atomicAdd((unsigned long long *)&cnt[kGathers], (unsigned long long)ngath);
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

// One workgroup per vertex (degree > kMediumMax): 256 neighbours per round.
__global__ __launch_bounds__(kBlock) void k_pull_block(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint32_t *__restrict__ new2old, uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint16_t *__restrict__ level, int64_t *__restrict__ cnt_all, int bin)
{
    __shared__ uint64_t s_wave[kBlock / 64];
    __shared__ uint64_t s_vis;
    const uint16_t depth = (uint16_t)cnt_all[kDepth];
    int64_t *__restrict__ cnt = cnt_all + 4 * bin;
    int64_t ngath = 0;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    int64_t nbits = 0, nedges = 0, nverts = 0;
    for (int64_t i = blockIdx.x; i < nlist; i += gridDim.x) {
        const uint32_t v = (uint32_t)(first_v + i);
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
                if (active[u >> 6] >> (u & 63) & 1) {
                    f = frontier[u] & unv;
                    ++ngath;
                }
            }
            f &= ~found;
            const bool any = __ballot(f != 0) != 0;
            const uint64_t in_wave = any ? lanes_excl_or(f) : 0ull;
            const uint64_t wave_all = any ? __shfl(in_wave | f, 63, 64) : 0ull;
            if (lane == 0) s_wave[wv] = wave_all;
            __syncthreads();
            uint64_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) {
                const uint64_t t = s_wave[w];
                before |= (w < wv) ? t : 0ull;
                all |= t;
            }
            const uint64_t first = f & ~(in_wave | before);
            if (first) write_parents(parent, v, first, new2old[u]);
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
    ngath = g5_wave_sum(ngath);
    if (lane == 0 && ngath) atomicAdd((unsigned long long *)&cnt[kGathers], (unsigned long long)ngath);
    if (threadIdx.x == 0) {
        if (nbits) atomicAdd((unsigned long long *)&cnt[kNewBits], (unsigned long long)nbits);
        if (nedges) atomicAdd((unsigned long long *)&cnt[kEdgesRead], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kVisitedVerts], (unsigned long long)nverts);
    }
}

// ---- push (top-down) levels for sparse frontiers ----
// The frontier's vertices u (active bitmap -> sorted queue) push their lane
// bits along their edges: next[v] |= frontier[u] & ~visited[v].  Parents are
// then the minimum frontier neighbour per (v, root): every slot of a newly
// reached pair is set to INT32_MAX, and a second pass over the same edges
// lowers it with atomicMin.  Work is balanced by edges: each block tile takes
// 2048 consecutive edges of the frontier and finds each edge's source by
// binary search in the frontier's degree prefix.
constexpr int kPushTile = 2048;
constexpr int kPushGrid = 4096;

__device__ __forceinline__ int64_t fq_src(const int64_t *__restrict__ qoff, int64_t lo, int64_t hi, int64_t e)
{
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (qoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_fq_degrees(const uint32_t *__restrict__ fq, const int64_t *__restrict__ cnt,
                             const int64_t *__restrict__ rowptr, int64_t *__restrict__ qoff)
{
    const int64_t n = cnt[kFqLen];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = fq[i];
        qoff[i] = rowptr[u + 1] - rowptr[u];
    }
}

template <bool kMin>
__global__ __launch_bounds__(kBlock) void k_push_edges(
    const uint32_t *__restrict__ fq, const int64_t *__restrict__ qoff, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, const uint32_t *__restrict__ new2old, const uint64_t *__restrict__ visited,
    const uint64_t *__restrict__ frontier, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint32_t *__restrict__ nq, int64_t *__restrict__ cnt)
{
    __shared__ int64_t s_lo, s_hi;
    const int64_t qn = cnt[kFqLen], E = cnt[kFqEdges];
    for (int64_t t0 = blockIdx.x * (int64_t)kPushTile; t0 < E; t0 += (int64_t)gridDim.x * kPushTile) {
        const int64_t t1 = min(E, t0 + (int64_t)kPushTile);
        if (threadIdx.x == 0) {
            s_lo = fq_src(qoff, 0, qn - 1, t0);
            s_hi = fq_src(qoff, s_lo, qn - 1, t1 - 1);
        }
        __syncthreads();
        const int64_t lo = s_lo, hi = s_hi;
        for (int64_t e = t0 + threadIdx.x; e < t1; e += kBlock) {
            const int64_t si = fq_src(qoff, lo, hi, e);
            const uint32_t u = fq[si];
            const uint32_t v = cols[rowptr[u] + (e - qoff[si])];
            const uint64_t f = frontier[u] & ~visited[v];
            if (!f) continue;
            if (!kMin) {
                if ((next[v] & f) != f) {
                    const uint64_t old = atomicOr((unsigned long long *)&next[v], (unsigned long long)f);
                    if (old == 0) nq[atomicAdd((unsigned long long *)&cnt[kNq], 1ull)] = v;
                }
            } else {
                uint64_t bits = f;  // every bit of f is new at this level (visited excludes older ones)
                const int32_t uo = (int32_t)new2old[u];
                while (bits) {
                    const int r = __ffsll((long long)bits) - 1;
                    int32_t *slot = &parent[(int64_t)v * 64 + r];
                    if (*slot > uo) atomicMin(slot, uo);
                    bits &= bits - 1;
                }
            }
        }
        __syncthreads();
    }
    if (!kMin && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)&cnt[4 * kPushUnit + kEdgesRead], (unsigned long long)E);
}

// wave per newly reached vertex: stage = 0 parent slots := INT32_MAX,
// stage = 1 visited |= next, level := depth, counts
template <int kStage>
__global__ __launch_bounds__(kBlock) void k_push_vertices(const uint32_t *__restrict__ nq, uint64_t *__restrict__ next,
                                                          uint64_t *__restrict__ visited, int32_t *__restrict__ parent,
                                                          uint16_t *__restrict__ level, int64_t *__restrict__ cnt)
{
    const int64_t n = cnt[kNq];
    const uint16_t depth = (uint16_t)cnt[kDepth];
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
    int64_t nbits = 0, nv = 0;
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t v = nq[i];
        const uint64_t w = next[v];
        if (kStage == 0) {
            if (w >> lane & 1) parent[(int64_t)v * 64 + lane] = 0x7FFFFFFF;
        } else {
            if (w >> lane & 1) level[(int64_t)v * 64 + lane] = depth;
            if (lane == 0) {
                visited[v] |= w;
                nbits += __popcll(w);
                ++nv;
            }
        }
    }
    if (kStage == 1 && lane == 0 && nbits) {
        atomicAdd((unsigned long long *)&cnt[4 * kPushUnit + kNewBits], (unsigned long long)nbits);
        atomicAdd((unsigned long long *)&cnt[4 * kPushUnit + kVisitedVerts], (unsigned long long)nv);
    }
}

__global__ void k_level_begin(int64_t *cnt)
{
    for (int i = 0; i < kDepth; ++i) cnt[i] = 0;
    cnt[kActVerts] = 0;
    cnt[kActEdges] = 0;
    cnt[kNq] = 0;
}

// active[w] bit i = (next[64w + i] != 0): one wavefront per 64-vertex word
__global__ __launch_bounds__(256) void k_build_active(const uint64_t *__restrict__ next, int64_t nwords,
                                                      const int64_t *__restrict__ rowptr,
                                                      uint64_t *__restrict__ active, int64_t *__restrict__ cnt)
{
    const int64_t waves = (int64_t)gridDim.x * 4;
    int64_t nv = 0, ne = 0;
    for (int64_t w = blockIdx.x * 4 + (threadIdx.x >> 6); w < nwords; w += waves) {
        const int64_t v = w * 64 + __lane_id();
        const bool on = next[v] != 0;
        const uint64_t bits = __ballot(on);
        if (__lane_id() == 0) active[w] = bits;
        if (on) {
            ++nv;
            ne += rowptr[v + 1] - rowptr[v];
        }
    }
    nv = g5_wave_sum(nv);
    ne = g5_wave_sum(ne);
    if (__lane_id() == 0 && nv) {
        atomicAdd((unsigned long long *)&cnt[kActVerts], (unsigned long long)nv);
        atomicAdd((unsigned long long *)&cnt[kActEdges], (unsigned long long)ne);
    }
}

__global__ void k_level_end(int64_t *cnt) { cnt[kDepth] += 1; }

__global__ void k_ms_init(uint64_t *__restrict__ visited, uint64_t *__restrict__ fa, uint64_t *__restrict__ fb,
                          uint64_t *__restrict__ act, int64_t n, uint64_t unused_lanes)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        visited[i] = unused_lanes;
        fa[i] = 0;
        fb[i] = 0;
        if (i < 2 * (n / 64)) act[i] = 0;
    }
}

// Roots: one thread; lanes of duplicate roots share the vertex.
__global__ void k_ms_roots(const int64_t *__restrict__ roots, int nroots, const uint32_t *__restrict__ old2new,
                           uint64_t *__restrict__ visited, uint64_t *__restrict__ frontier,
                           uint64_t *__restrict__ active, int32_t *__restrict__ parent,
                           uint16_t *__restrict__ level, int64_t *__restrict__ cnt)
{
    for (int r = 0; r < nroots; ++r) {
        const int64_t v = old2new[roots[r]];
        visited[v] |= 1ull << r;
        frontier[v] |= 1ull << r;
        active[v >> 6] |= 1ull << (v & 63);
        parent[v * 64 + r] = (int32_t)roots[r];
        level[v * 64 + r] = 0;
    }
    cnt[kDepth] = 1;
}

// Degree-order keys: (2^32-1 - degree) << 32 | id, so an ascending sort
// gives descending degree, ties by id; also counts the degree bins.
__global__ void k_degree_keys(const int64_t *__restrict__ rowptr, int64_t n, uint64_t *__restrict__ keys,
                              unsigned long long *__restrict__ bins)
{
    unsigned long long c[kBins] = {0, 0, 0};
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = rowptr[v + 1] - rowptr[v];
        keys[v] = ((uint64_t)(0xFFFFFFFFu - (uint32_t)d) << 32) | (uint64_t)v;
        for (int k = 0; k < kBins; ++k)
            if (d > 0 && d <= kBinMaxDeg[k] && (k + 1 == kBins || d > kBinMaxDeg[k + 1])) ++c[k];
    }
    for (int k = 0; k < kBins; ++k) {
        const unsigned long long t = (unsigned long long)g5_wave_sum((int64_t)c[k]);
        if (__lane_id() == 0 && t) atomicAdd(&bins[k], t);
    }
}

__global__ void k_perm(const uint64_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ new2old,
                       uint32_t *__restrict__ old2new, const int64_t *__restrict__ rowptr,
                       const uint32_t *__restrict__ gdeg, int64_t *__restrict__ ndeg, uint32_t *__restrict__ ngdeg)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = (uint32_t)keys[i];
        new2old[i] = o;
        old2new[o] = (uint32_t)i;
        ndeg[i] = rowptr[o + 1] - rowptr[o];
        ngdeg[i] = gdeg[o];
    }
}

// one wavefront per relabelled row: copy the original row (ascending
// original ids) mapped to new ids
__global__ __launch_bounds__(256) void k_perm_rows(const int64_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ cols,
                                                   const uint32_t *__restrict__ new2old,
                                                   const uint32_t *__restrict__ old2new, int64_t n,
                                                   const int64_t *__restrict__ nrowptr, uint32_t *__restrict__ ncols)
{
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t o = new2old[i];
        const int64_t b = rowptr[o], d = rowptr[o + 1] - b, dst = nrowptr[i];
        for (int64_t j = __lane_id(); j < d; j += 64) ncols[dst + j] = old2new[cols[b + j]];
    }
}

// per-lane extraction: pred/level of root `lane` as int64/int32[n]; a lane's
// entries are meaningful only where its visited bit is set
__global__ void k_ms_extract(const uint32_t *__restrict__ old2new, const int32_t *__restrict__ parent,
                             const uint16_t *__restrict__ level, const uint64_t *__restrict__ visited, int64_t n,
                             int lane, int64_t *__restrict__ pred, int32_t *__restrict__ lev)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = old2new[i];
        const bool reached = visited[v] >> lane & 1;
        lev[i] = reached ? (int32_t)level[v * 64 + lane] : -1;
        pred[i] = reached ? (int64_t)parent[v * 64 + lane] : -1;
    }
}

// nedge per lane: sum of generated degrees over reached vertices
__global__ __launch_bounds__(256) void k_ms_nedge(const uint64_t *__restrict__ visited, const uint32_t *__restrict__ gdeg,
                                                  int64_t n, unsigned long long *__restrict__ out)
{
    // thread t of a wave handles lane t for a strip of vertices
    const int lane = __lane_id();
    unsigned long long acc = 0;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t v = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6); v < n; v += waves) {
        const uint32_t d = gdeg[v];
        if (d && (visited[v] >> lane & 1)) acc += d;
    }
    atomicAdd(&out[lane], acc);
}

// prof: events between the three pull kernels (the uncaptured path)
int level_launches(g5_msbfs *m, hipStream_t s, int p, bool prof = false)
{
    const Relabeled &R = m->rl;
    const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    const uint64_t *act = m->d_act + (p == 0 ? 0 : nw);
    uint64_t *act_next = m->d_act + (p == 0 ? nw : 0);
    hipLaunchKernelGGL(k_level_begin, 1, 1, 0, s, m->d_cnt);
    for (int k = 0; k < kBins; ++k) {
        if (prof) G5_HIP(hipEventRecord(m->ek[k], s));
        const int64_t first = R.bin_first[k], cnt = R.bin_first[k + 1] - first;
        if (!cnt) continue;
#define G5_PULL_ARGS first, cnt, (const int64_t *)R.rowptr, (const uint32_t *)R.cols, (const uint32_t *)R.new2old, \
                     m->d_visited, front, act, next, m->d_parent, m->d_level, m->d_cnt, k
        switch (k) {
        case 0:
            hipLaunchKernelGGL(k_pull_block, (unsigned)std::min<int64_t>(cnt, 4096), kBlock, 0, s, G5_PULL_ARGS);
            break;
        case 1:
            hipLaunchKernelGGL(k_pull_group<64>, g5::grid_for(cnt, kBlock / 64, 16384), kBlock, 0, s, G5_PULL_ARGS);
            break;
        default:
            hipLaunchKernelGGL(k_pull_thread, g5::grid_for(cnt, kBlock, 16384), kBlock, 0, s, G5_PULL_ARGS);
            break;
        }
#undef G5_PULL_ARGS
    }
    if (prof) G5_HIP(hipEventRecord(m->ek[kBins], s));
    hipLaunchKernelGGL(k_build_active, g5::grid_for((nw + 3) / 4, 1, 16384), 256, 0, s, (const uint64_t *)next, nw,
                       (const int64_t *)R.rowptr, act_next, m->d_cnt);
    hipLaunchKernelGGL(k_level_end, 1, 1, 0, s, m->d_cnt);
    return hipGetLastError() == hipSuccess ? G5_OK : G5_E_HIP;
}

// One push level on parity p: direct launches (the frontier queue's length
// sizes the degree scan on the host).  fqn = active vertices of the frontier.
int push_level(g5_msbfs *m, hipStream_t s, int p, int64_t fqn, bool prof)
{
    const Relabeled &R = m->rl;
    const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    const uint64_t *act = m->d_act + (p == 0 ? 0 : nw);
    uint64_t *act_next = m->d_act + (p == 0 ? nw : 0);
    hipLaunchKernelGGL(k_level_begin, 1, 1, 0, s, m->d_cnt);
    G5_HIP(g5::fill_u64(next, m->n, 0, s));
    if (prof) G5_HIP(hipEventRecord(m->ek[0], s));
    G5_HIP(g5::bitmap_to_queue(act, nw, m->d_fq, m->d_cnt + kFqLen, m->d_scratch, s));
    hipLaunchKernelGGL(k_fq_degrees, g5::grid_for(fqn, 256, 8192), 256, 0, s, (const uint32_t *)m->d_fq,
                       (const int64_t *)m->d_cnt, (const int64_t *)R.rowptr, m->d_qoff);
    G5_HIP(g5::exclusive_scan_i64(m->d_qoff, fqn, m->d_cnt + kFqEdges, m->d_scratch, s));
    hipLaunchKernelGGL(k_push_edges<false>, kPushGrid, kBlock, 0, s, (const uint32_t *)m->d_fq,
                       (const int64_t *)m->d_qoff, (const int64_t *)R.rowptr, (const uint32_t *)R.cols,
                       (const uint32_t *)R.new2old, (const uint64_t *)m->d_visited, front, next, m->d_parent,
                       m->d_nq, m->d_cnt);
    hipLaunchKernelGGL(k_push_vertices<0>, 2048, kBlock, 0, s, (const uint32_t *)m->d_nq, next, m->d_visited,
                       m->d_parent, m->d_level, m->d_cnt);
    hipLaunchKernelGGL(k_push_edges<true>, kPushGrid, kBlock, 0, s, (const uint32_t *)m->d_fq,
                       (const int64_t *)m->d_qoff, (const int64_t *)R.rowptr, (const uint32_t *)R.cols,
                       (const uint32_t *)R.new2old, (const uint64_t *)m->d_visited, front, next, m->d_parent,
                       m->d_nq, m->d_cnt);
    hipLaunchKernelGGL(k_push_vertices<1>, 2048, kBlock, 0, s, (const uint32_t *)m->d_nq, next, m->d_visited,
                       m->d_parent, m->d_level, m->d_cnt);
    if (prof) G5_HIP(hipEventRecord(m->ek[1], s));
    hipLaunchKernelGGL(k_build_active, g5::grid_for((nw + 3) / 4, 1, 16384), 256, 0, s, (const uint64_t *)next, nw,
                       (const int64_t *)R.rowptr, act_next, m->d_cnt);
    hipLaunchKernelGGL(k_level_end, 1, 1, 0, s, m->d_cnt);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

int capture_graphs(g5_msbfs *m)
{
    for (int p = 0; p < 2; ++p) {
        hipGraph_t gr = nullptr;
        G5_HIP(hipStreamBeginCapture(m->s, hipStreamCaptureModeThreadLocal));
        int rc = level_launches(m, m->s, p);
        hipError_t e = hipStreamEndCapture(m->s, &gr);
        if (rc) return rc;
        G5_HIP(e);
        e = hipGraphInstantiate(&m->graph[p], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        G5_HIP(e);
    }
    return G5_OK;
}

// The degree-ordered copy of the block (see the file comment).
int build_relabel(g5_msbfs *m)
{
    g5_graph *g = m->g;
    hipStream_t s = m->s;
    Relabeled &R = m->rl;
    const int64_t n = m->n;
    uint64_t *keys = nullptr, *tmp = nullptr;
    unsigned long long *bins = nullptr;
    int64_t *ndeg = nullptr;
    auto cleanup = [&]() {
        if (keys) (void)hipFree(keys);
        if (tmp) (void)hipFree(tmp);
        if (bins) (void)hipFree(bins);
        if (ndeg) (void)hipFree(ndeg);
    };
    hipError_t e = hipMalloc(&keys, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&tmp, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&bins, kBins * 8);
    if (e == hipSuccess) e = hipMalloc(&ndeg, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&R.new2old, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.old2new, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.gdeg, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.rowptr, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&R.cols, (size_t)std::max<int64_t>(g->nnz, 1) * 4);
    if (e == hipSuccess) e = hipMemsetAsync(bins, 0, kBins * 8, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_degree_keys, g5::grid_for(n, 256, 16384), 256, 0, s, (const int64_t *)g->d_rowptr, n,
                           keys, bins);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = g5::sort_u64(keys, tmp, n, 64, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_perm, g5::grid_for(n, 256, 16384), 256, 0, s, (const uint64_t *)keys, n, R.new2old,
                           R.old2new, (const int64_t *)g->d_rowptr, (const uint32_t *)g->d_gdeg, ndeg, R.gdeg);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = g5::exclusive_scan_i64(ndeg, n, ndeg + n, m->d_scratch, s);
    if (e == hipSuccess) e = hipMemcpyAsync(R.rowptr, ndeg, (size_t)(n + 1) * 8, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_perm_rows, g5::grid_for((n + 3) / 4, 1, 65536), 256, 0, s,
                           (const int64_t *)g->d_rowptr, (const uint32_t *)g->d_cols, (const uint32_t *)R.new2old,
                           (const uint32_t *)R.old2new, n, (const int64_t *)R.rowptr, R.cols);
        e = hipGetLastError();
    }
    unsigned long long hb[kBins] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(hb, bins, kBins * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return g5::fail_hip(e, "msbfs relabel", __FILE__, __LINE__);
    R.bin_first[0] = 0;
    for (int k = 0; k < kBins; ++k) R.bin_first[k + 1] = R.bin_first[k] + (int64_t)hb[k];
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
    if (e == hipSuccess) e = hipMalloc(&m->d_act, (size_t)(n / 64) * 2 * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_fq, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_nq, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_qoff, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_parent, (size_t)n * 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_level, (size_t)n * 64 * 2);
    if (e == hipSuccess) e = hipMalloc(&m->d_cnt, kMsN * 8);
    if (e == hipSuccess) e = hipHostMalloc(&m->h_cnt, kMsN * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&m->d_scratch, (g5::scan_scratch_words(n + 1) + 16) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_roots, 64 * 8);
    if (e == hipSuccess) e = hipEventCreate(&m->ev0);
    if (e == hipSuccess) e = hipEventCreate(&m->ev1);
    for (int k = 0; k <= kBins && e == hipSuccess; ++k) e = hipEventCreate(&m->ek[k]);
    if (e != hipSuccess) return fail(g5::fail_hip(e, "g5_msbfs_create allocation", __FILE__, __LINE__));
    int rc = build_relabel(m);
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
    void *dev[] = {m->d_fq, m->d_nq, m->d_qoff, m->d_act, m->d_visited, m->d_fa,        m->d_fb,           m->d_parent,       m->d_level,
                   m->d_cnt,     m->d_scratch,   m->d_roots,        m->rl.new2old,     m->rl.old2new,     m->rl.rowptr,
                   m->rl.cols,   m->rl.gdeg};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (m->h_cnt) (void)hipHostFree(m->h_cnt);
    hipEvent_t evs[] = {m->ev0, m->ev1};
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : m->ek)
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
                       m->d_act, m->n, unused);
    hipLaunchKernelGGL(k_ms_roots, 1, 1, 0, s, (const int64_t *)d_roots, nroots, (const uint32_t *)m->rl.old2new,
                       m->d_visited, m->d_fa, m->d_act, m->d_parent, m->d_level, m->d_cnt);
    G5_HIP(hipGetLastError());
    int p = 0, levels = 1;
    m->lvl.clear();
    // the first frontier is the set of distinct roots; its edge count is not
    // known on the host, and a handful of roots always takes the push path
    int64_t act_verts = 0, act_edges = -1;
    {
        int64_t tmp[64];
        std::copy(roots, roots + nroots, tmp);
        std::sort(tmp, tmp + nroots);
        act_verts = std::unique(tmp, tmp + nroots) - tmp;
    }
    const int64_t nnz = m->g->nnz;
    for (;;) {
        const bool push = m->mode == 2 || (m->mode == 0 && (act_edges < 0 || act_edges * kPushRatio < nnz));
        if (push) {
            int rc = push_level(m, s, p, act_verts, m->prof);
            if (rc) return rc;
        } else if (m->prof) {
            int rc = level_launches(m, s, p, true);
            if (rc) return rc;
        } else {
            G5_HIP(hipGraphLaunch(m->graph[p], s));
        }
        G5_HIP(hipMemcpyAsync(m->h_cnt, m->d_cnt, kMsN * 8, hipMemcpyDeviceToHost, s));
        G5_HIP(hipStreamSynchronize(s));
        int64_t new_bits = 0;
        for (int c = 0; c < kUnits; ++c) new_bits += m->h_cnt[4 * c + kNewBits];
        if (m->prof) {
            double rec[kLevelRec] = {0};
            rec[8] = (double)new_bits;
            rec[9] = push ? 1.0 : 0.0;
            rec[10] = (double)act_verts;
            rec[11] = (double)act_edges;
            if (push) {
                // 16 B row offsets per frontier vertex; per frontier edge and
                // pass (two passes) 4 B column + 8 B frontier word + 8 B
                // visited word + 8 B next word; 16 B visited read-modify-write
                // per newly reached vertex; 4 B parent init + 4 B parent + 2 B
                // level per new (vertex, root) pair
                float kms = 0;
                G5_HIP(hipEventElapsedTime(&kms, m->ek[0], m->ek[1]));
                const double E = (double)m->h_cnt[kFqEdges];
                m->k_ms[kPushUnit] += kms;
                m->k_launches[kPushUnit] += 1;
                m->k_bytes[kPushUnit] += 16.0 * m->h_cnt[kFqLen] + 56.0 * E +
                                         16.0 * m->h_cnt[4 * kPushUnit + kVisitedVerts] +
                                         10.0 * m->h_cnt[4 * kPushUnit + kNewBits];
                rec[kPushUnit] = kms;
                rec[4 + kPushUnit] = E;
            } else {
                // per bin: 16 B row offsets + 8 B visited read + 8 B next
                // write per vertex with an unvisited lane; 4 B column + 1 bit
                // of the active bitmap per edge read; 8 B per frontier word
                // gathered; 4 B parent + 2 B level per new (vertex, root) pair
                for (int k = 0; k < kBins; ++k) {
                    if (m->rl.bin_first[k + 1] == m->rl.bin_first[k]) continue;
                    float kms = 0;
                    G5_HIP(hipEventElapsedTime(&kms, m->ek[k], m->ek[k + 1]));
                    m->k_ms[k] += kms;
                    m->k_launches[k] += 1;
                    m->k_bytes[k] += 32.0 * m->h_cnt[4 * k + kVisitedVerts] + 4.125 * m->h_cnt[4 * k + kEdgesRead] +
                                     8.0 * m->h_cnt[4 * k + kGathers] + 6.0 * m->h_cnt[4 * k + kNewBits];
                    rec[k] = kms;
                    rec[4 + k] = (double)m->h_cnt[4 * k + kEdgesRead];
                }
            }
            m->lvl.insert(m->lvl.end(), rec, rec + kLevelRec);
        }
        if (new_bits == 0) break;
        G5_REQUIRE(levels < 65535, G5_E_STATE, "g5_msbfs_run: more than 65534 levels (uint16 depth limit)");
        act_verts = m->h_cnt[kActVerts];
        act_edges = m->h_cnt[kActEdges];
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
                           (const uint32_t *)m->rl.old2new, (const int32_t *)m->d_parent,
                           (const uint16_t *)m->d_level, (const uint64_t *)m->d_visited, m->n, lane, dp, dl);
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
                           (const uint64_t *)m->d_visited, (const uint32_t *)m->rl.gdeg, m->n, d);
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
    for (int c = 0; c < kUnits; ++c) {
        m->k_ms[c] = m->k_bytes[c] = 0;
        m->k_launches[c] = 0;
    }
    return G5_OK;
}

extern "C" int g5_msbfs_set_mode(g5_msbfs *m, int mode)
{
    G5_REQUIRE(m && mode >= 0 && mode <= 2, G5_E_ARG, "g5_msbfs_set_mode: mode must be 0, 1 or 2");
    m->mode = mode;
    return G5_OK;
}

extern "C" int g5_msbfs_stats(g5_msbfs *m, int kernel, double *ms, int64_t *launches, double *bytes)
{
    G5_REQUIRE(m && kernel >= 0 && kernel < kUnits, G5_E_ARG, "g5_msbfs_stats: bad arguments");
    if (ms) *ms = m->k_ms[kernel];
    if (launches) *launches = m->k_launches[kernel];
    if (bytes) *bytes = m->k_bytes[kernel];
    return G5_OK;
}

extern "C" int g5_msbfs_info(g5_msbfs *m, int64_t *bins, int cap)
{
    G5_REQUIRE(m && bins, G5_E_ARG, "g5_msbfs_info: null argument");
    for (int k = 0; k < kBins && k < cap; ++k) bins[k] = m->rl.bin_first[k + 1] - m->rl.bin_first[k];
    return G5_OK;
}

extern "C" int g5_msbfs_levels(g5_msbfs *m, double *out, int cap, int *nlev)
{
    G5_REQUIRE(m && nlev, G5_E_ARG, "g5_msbfs_levels: null argument");
    const int n = (int)(m->lvl.size() / kLevelRec);
    *nlev = n;
    if (out)
        for (int i = 0; i < std::min(n, cap) * kLevelRec; ++i) out[i] = m->lvl[i];
    return G5_OK;
}
