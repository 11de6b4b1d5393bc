// g5_bfs.hip -- single-root local expansion: the GPU "Derived" plug-in of
// GlobalBFS (cpu_2d/cuda/cuda_bfs.h:39-84) re-designed for gfx950.
//
// State per search lives in HBM as bitmaps (visited, out) plus a sorted
// uint32 in-queue.  A level is:
//   1. degree prefix over the in-queue (exclusive scan of row lengths),
//   2. k_expand: edge-balanced top-down expansion -- every block takes a
//      2048-edge tile of the level's edge range, finds each edge's source by
//      binary search in the degree prefix, and for an unvisited target sets
//      its out bit (atomicOr, skipped when already set) and lowers its parent
//      with atomicMin.  The surviving parent is the minimum-id frontier
//      neighbour: the same vertex CPUBFS_bin's ascending scan keeps
//      (cpu_2d/cpubfs_bin.cpp:150-173), independent of thread timing.
//   3. k_commit: visited |= out, level := depth + 1, count of new vertices.
//   4. out bitmap -> sorted queue (word popcount scan), the next in-queue.
// Queues are born sorted and unique, so the reference's thrust::sort and host
// unique_copy (cuda_bfs.cu:380-475) have no counterpart.
//
// g5_bfs_run (1x1 block, resident) also runs dense levels bottom-up
// (k_bottom_up): each unvisited vertex scans its sorted neighbour list and
// stops at the first frontier vertex, which is again the minimum-id frontier
// neighbour, so both directions give identical parents and levels.  It
// reads only the counters from the host per level, and builds the sorted
// queue of step 4 only for a top-down level or the codec.
#include <algorithm>

#include "g5_codec.hpp"
#include "g5_internal.hpp"

namespace {

constexpr uint32_t kNoParent = 0xFFFFFFFFu;
constexpr uint32_t kRootMark = 0xFFFFFFFEu;  // parent slot of the root column
constexpr int kExpandBlock = 256;
constexpr int kExpandTile = 2048;            // edges per block tile
constexpr int kExpandGrid = 4096;            // persistent grid, grid-stride over tiles

// kNewEdges: degree sum of the newly reached vertices (1x1 only); kBuEdges,
// kBuVerts: neighbours checked and vertices scanned by a bottom-up level;
// kCodecErr: set when a level queue's codec round trip failed (g5_bfs_run)
enum Cnt {
    kQLen = 0, kEdges = 1, kNew = 2, kOutLen = 3, kNewEdges = 4, kBuEdges = 5, kBuVerts = 6, kCodecErr = 7, kNCnt = 8
};
// direction switch of the 1x1 search (Beamer, Asanovic, Patterson, SC'12):
// top-down -> bottom-up when the frontier's edges exceed the unexplored
// edges / alpha, back when the frontier holds fewer than n / beta vertices
// (g5_bfs_tune).  As in the GAP benchmark's implementation of the rule, the
// unexplored edges are reduced only by the frontiers expanded top-down, so
// that the few vertices left at the tail of the search do not flip it back
// to bottom-up, where every unreached vertex would be scanned again.
// alpha 20: 3.32 ms per root at scale 26 against 3.51 ms with 14 and 3.94
// with 8 (scale 22: within 1 %); beta barely matters (12..96: within 0.5 %)
constexpr int64_t kAlpha = 20, kBeta = 24;
// frontier bits gathered ahead per lane in the bottom-up scan: 4 on graphs
// below 2^25 vertices, 2 above, where each gather past the first hit more
// likely misses the caches (measured at scales 22, 24 and 26; 1 and 8 were
// slower everywhere)
constexpr int64_t kBuLarge = (int64_t)1 << 25;

}  // namespace

struct g5_bfs {
    g5_graph *g = nullptr;
    hipStream_t s = nullptr;
    int64_t wr = 0, wc = 0;  // row / column bitmap words
    uint64_t *d_visited = nullptr, *d_out = nullptr;
    uint32_t *d_parent = nullptr;
    int32_t *d_level = nullptr;
    int64_t alpha = kAlpha, beta = kBeta;  // direction switch of g5_bfs_run
    uint32_t *d_queue = nullptr;  // in-queue (local rows), sorted
    uint32_t *d_outq = nullptr;   // out-queue (local columns), sorted
    int64_t *d_qoff = nullptr;    // degree prefix over the in-queue, qlen + 1
    int64_t *d_scratch = nullptr;
    int64_t *d_cnt = nullptr;
    int64_t *h_cnt = nullptr;     // pinned mirror of d_cnt
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    g5::S4Codec codec;
    uint32_t *d_words = nullptr;  // codec stream for G5_RUN_CODEC
    size_t words_cap = 0;
    int64_t root = -1;
    int depth = -1;
    int64_t qlen = 0;
    // per-kernel accounting of k_expand (g5_bfs_profile / g5_bfs_stats)
    bool prof = false;
    hipEvent_t ek0 = nullptr, ek1 = nullptr;
    double exp_ms = 0, exp_bytes = 0;
    int64_t exp_launches = 0, exp_edges = 0;
    // bottom-up levels of g5_bfs_run: the next-frontier bitmap and accounting
    uint64_t *d_next = nullptr;
    double bu_ms = 0, bu_bytes = 0;
    int64_t bu_launches = 0, bu_edges = 0;
    // g5_bfs_run: vertices without edges (a diagonal block), visited from
    // the start so that no bottom-up level reads their offsets
    uint64_t *d_isolated = nullptr;
    // hook path: the row frontier as a bitmap (g5_bfs_set_in_bitmap), for
    // g5_bfs_expand_bottom_up
    uint64_t *d_inbits = nullptr;
    bool inbits_valid = false;
};

namespace {

__device__ __forceinline__ int64_t src_of(const int64_t *__restrict__ qoff, int64_t lo, int64_t hi,
                                          int64_t e)
{
    // last s in [lo, hi] with qoff[s] <= e
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (qoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_queue_degrees(const uint32_t *__restrict__ q, const int64_t *__restrict__ d_qlen,
                                const int64_t *__restrict__ rowptr, int64_t *__restrict__ qoff)
{
    const int64_t n = *d_qlen;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = q[i];
        qoff[i] = rowptr[u + 1] - rowptr[u];
    }
}

__global__ __launch_bounds__(kExpandBlock) void k_expand(
    const uint32_t *__restrict__ q, const int64_t *__restrict__ qoff, const int64_t *__restrict__ cnt,
    const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint64_t *__restrict__ visited, uint64_t *__restrict__ out, uint32_t *__restrict__ parent)
{
    __shared__ int64_t s_lo, s_hi;
    const int64_t qn = cnt[kQLen];
    const int64_t E = cnt[kEdges];
    for (int64_t t0 = blockIdx.x * (int64_t)kExpandTile; t0 < E; t0 += (int64_t)gridDim.x * kExpandTile) {
        const int64_t t1 = min(E, t0 + (int64_t)kExpandTile);
        if (threadIdx.x == 0) {
            s_lo = src_of(qoff, 0, qn - 1, t0);
            s_hi = src_of(qoff, s_lo, qn - 1, t1 - 1);
        }
        __syncthreads();
        const int64_t lo = s_lo, hi = s_hi;
        for (int64_t e = t0 + threadIdx.x; e < t1; e += kExpandBlock) {
            const int64_t si = src_of(qoff, lo, hi, e);
            const uint32_t u = q[si];
            const uint32_t v = cols[rowptr[u] + (e - qoff[si])];
            const uint64_t bit = 1ull << (v & 63);
            if (visited[v >> 6] & bit) continue;
            if (!(out[v >> 6] & bit)) atomicOr((unsigned long long *)&out[v >> 6], (unsigned long long)bit);
            if (parent[v] > u) atomicMin(&parent[v], u);
        }
        __syncthreads();
    }
}

// visited |= out (minus already visited), level := depth for new bits, count;
// with rowptr (1x1 block) also the degree sum of the new vertices.  One
// wavefront per 64 consecutive words: lane j commits word w0 + j; then the
// wave walks the words with new bits 8 at a time, lane i taking vertex
// 64 w + i, so a word's level stores and offset reads are coalesced, and the
// 8 words' offset loads are all issued before any is used (a vertex's degree
// is the next lane's offset minus its own; lane 63 loads the next word's
// first offset).
constexpr int kCommitGroup = 8;

__device__ __forceinline__ uint64_t bcast64(uint64_t x, int l)
{
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
}

__global__ __launch_bounds__(256) void k_commit(uint64_t *__restrict__ out, uint64_t *__restrict__ visited,
                                                int32_t *__restrict__ level, int64_t wc, int32_t depth,
                                                int64_t *__restrict__ cnt, const int64_t *__restrict__ rowptr)
{
    const int lane = __lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int64_t c = 0, ce = 0;
    for (int64_t w0 = (blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; w0 < wc;
         w0 += nwaves * 64) {
        const int64_t w = w0 + lane;
        uint64_t nw = 0;
        if (w < wc) {
            const uint64_t o = out[w];
            if (o) {
                const uint64_t vis = visited[w];
                nw = o & ~vis;
                if (nw != o) out[w] = nw;
                if (nw) visited[w] = vis | nw;
                c += __popcll(nw);
            }
        }
        const uint64_t todo = __ballot(nw != 0);
        for (int j0 = 0; j0 < 64; j0 += kCommitGroup) {
            if (((todo >> j0) & ((1ull << kCommitGroup) - 1)) == 0) continue;
            uint64_t bits[kCommitGroup];
            int64_t lo[kCommitGroup], hi[kCommitGroup];
#pragma unroll
            for (int k = 0; k < kCommitGroup; ++k) {
                bits[k] = bcast64(nw, j0 + k);  // zero for words past wc
                const int64_t v = (w0 + j0 + k) * 64 + lane;
                const bool any = bits[k] != 0;  // wave-uniform
                lo[k] = (rowptr && any) ? rowptr[v] : 0;
                hi[k] = (rowptr && any && lane == 63) ? rowptr[v + 1] : 0;
                if (bits[k] >> lane & 1) level[v] = depth;
            }
#pragma unroll
            for (int k = 0; k < kCommitGroup; ++k) {
                const int64_t nxt = __shfl_down(lo[k], 1, 64);
                const int64_t up = lane == 63 ? hi[k] : nxt;
                if (bits[k] >> lane & 1) ce += up - lo[k];
            }
        }
    }
    c = g5_block_sum256(c);
    ce = g5_block_sum256(ce);
    if (threadIdx.x == 0 && c) {
        atomicAdd((unsigned long long *)&cnt[kNew], (unsigned long long)c);
        if (ce) atomicAdd((unsigned long long *)&cnt[kNewEdges], (unsigned long long)ce);
    }
}

// Bottom-up level of the 1x1 search: every unvisited vertex scans its
// neighbours in ascending id order and stops at the first one whose frontier
// bit is set -- its minimum-id frontier neighbour, the parent CPUBFS_bin's
// ascending scan keeps (cpu_2d/cpubfs_bin.cpp:150-173).  One wavefront per
// 64-vertex bitmap word (lane = vertex), so the next-frontier word is the
// wave's ballot and needs no atomics.  kBuUnroll frontier bits are gathered
// before any is tested, to keep several loads in flight per lane.
template <int kBuUnroll>
__global__ __launch_bounds__(256) void k_bottom_up(const int64_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ cols,
                                                   const uint64_t *__restrict__ visited,
                                                   const uint64_t *__restrict__ front, uint64_t *__restrict__ next,
                                                   uint32_t *__restrict__ parent, int64_t wc, int64_t *__restrict__ cnt)
{
    const int lane = __lane_id();
    int64_t nedges = 0, nverts = 0, newdeg = 0;
    for (int64_t w = blockIdx.x * 4 + (threadIdx.x >> 6); w < wc; w += (int64_t)gridDim.x * 4) {
        const uint64_t vis = visited[w];
        bool hit = false;
        if (!(vis >> lane & 1)) {
            const int64_t v = w * 64 + lane;
            const int64_t b = rowptr[v], e = rowptr[v + 1];
            int64_t j = b;
            while (j < e && !hit) {
                uint32_t u[kBuUnroll];
                uint64_t fw[kBuUnroll];
#pragma unroll
                for (int k = 0; k < kBuUnroll; ++k) u[k] = j + k < e ? cols[j + k] : 0u;
#pragma unroll
                for (int k = 0; k < kBuUnroll; ++k) fw[k] = j + k < e ? front[u[k] >> 6] : 0ull;
#pragma unroll
                for (int k = 0; k < kBuUnroll; ++k)
                    if (!hit && (fw[k] >> (u[k] & 63) & 1)) {
                        hit = true;
                        parent[v] = u[k];
                    }
                j += kBuUnroll;
            }
            nedges += min(j, e) - b;
            ++nverts;
            if (hit) newdeg += e - b;
        }
        const uint64_t bits = __ballot(hit);
        if (lane == 0) next[w] = bits;
    }
    nedges = g5_block_sum256(nedges);
    nverts = g5_block_sum256(nverts);
    newdeg = g5_block_sum256(newdeg);
    if (threadIdx.x == 0) {
        if (nedges) atomicAdd((unsigned long long *)&cnt[kBuEdges], (unsigned long long)nedges);
        if (nverts) atomicAdd((unsigned long long *)&cnt[kBuVerts], (unsigned long long)nverts);
        if (newdeg) atomicAdd((unsigned long long *)&cnt[kNewEdges], (unsigned long long)newdeg);
    }
}

__global__ void k_count_bits(const uint64_t *__restrict__ bits, int64_t n, int64_t *__restrict__ dst)
{
    int64_t c = 0;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n;
         w += (int64_t)gridDim.x * blockDim.x)
        c += __popcll(bits[w]);
    c = g5_block_sum256(c);
    if (threadIdx.x == 0 && c) atomicAdd((unsigned long long *)dst, (unsigned long long)c);
}

__global__ void k_start(int64_t root_row, int64_t root_col, uint32_t *queue, uint64_t *visited, uint64_t *out,
                        int32_t *level, uint32_t *parent, int64_t *cnt)
{
    cnt[kQLen] = root_row >= 0 ? 1 : 0;
    if (root_row >= 0) queue[0] = (uint32_t)root_row;
    if (root_col >= 0) {
        visited[root_col >> 6] |= 1ull << (root_col & 63);
        out[root_col >> 6] = 1ull << (root_col & 63);  // the frontier of a bottom-up first level
        level[root_col] = 0;
        parent[root_col] = kRootMark;
    }
}

__global__ void k_result(const uint32_t *__restrict__ parent, int64_t n, int64_t row_start,
                         int64_t col_start, int64_t *__restrict__ pred)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = parent[i];
        pred[i] = p == kNoParent ? -1 : (p == kRootMark ? i + col_start : (int64_t)p + row_start);
    }
}

__global__ void k_nedge(const uint32_t *__restrict__ gdeg, const int32_t *__restrict__ level, int64_t n,
                        int64_t *__restrict__ dst)
{
    int64_t c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (level[i] >= 0) c += gdeg[i];
    c = g5_block_sum256(c);
    if (threadIdx.x == 0 && c) atomicAdd((unsigned long long *)dst, (unsigned long long)c);
}

// nedge's generated-degree sum and the predecessor-map checksum of a 1x1
// search in one pass: sum over reached v of mix64(v << 32 | (pred(v) + 1))
__global__ void k_summary(const uint32_t *__restrict__ gdeg, const int32_t *__restrict__ level,
                          const uint32_t *__restrict__ parent, int64_t n, int64_t *__restrict__ dst)
{
    int64_t c = 0;
    uint64_t h = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (level[i] >= 0) {
            c += gdeg[i];
            const uint32_t p = parent[i];
            const uint64_t pred = p == kRootMark ? (uint64_t)i : (uint64_t)p;
            h += g5_mix64(((uint64_t)i << 32) | (uint32_t)(pred + 1));
        }
    c = g5_block_sum256(c);
    h = g5_block_sum256_u64(h);
    if (threadIdx.x == 0) {
        if (c) atomicAdd((unsigned long long *)dst, (unsigned long long)c);
        if (h) atomicAdd((unsigned long long *)(dst + 1), (unsigned long long)h);
    }
}

__global__ void k_set_cnt(int64_t *cnt, int idx, int64_t v) { cnt[idx] = v; }

// One level's expansion from the current in-queue (d_queue, cnt[kQLen]).
int expand_level(g5_bfs *b)
{
    g5_graph *g = b->g;
    hipStream_t s = b->s;
    G5_HIP(g5::fill_u64(b->d_out, b->wc, 0, s));
    hipLaunchKernelGGL(k_queue_degrees, g5::grid_for(g->n_rows, 256, 4096), 256, 0, s,
                       (const uint32_t *)b->d_queue, (const int64_t *)b->d_cnt + kQLen,
                       (const int64_t *)g->d_rowptr, b->d_qoff);
    G5_HIP(hipGetLastError());
    // the scan length must be known on the host: it is the in-queue length
    G5_HIP(g5::exclusive_scan_i64(b->d_qoff, b->qlen, b->d_cnt + kEdges, b->d_scratch, s));
    if (b->qlen > 0) {
        // qoff[qlen] = total, so src_of never reads past the prefix
        if (b->prof) G5_HIP(hipEventRecord(b->ek0, s));
        hipLaunchKernelGGL(k_expand, kExpandGrid, kExpandBlock, 0, s, (const uint32_t *)b->d_queue,
                           (const int64_t *)b->d_qoff, (const int64_t *)b->d_cnt,
                           (const int64_t *)g->d_rowptr, (const uint32_t *)g->d_cols,
                           (const uint64_t *)b->d_visited, b->d_out, b->d_parent);
        G5_HIP(hipGetLastError());
        if (b->prof) G5_HIP(hipEventRecord(b->ek1, s));
    }
    return G5_OK;
}

// After the level's counters are on the host: k_expand time and its
// algorithmic bytes, 16 B row-offset pair per frontier vertex, 4 B column id
// + 8 B visited word per frontier edge, 8 B out word + 4 B parent per newly
// reached vertex.
int account_level(g5_bfs *b, int64_t qlen, int64_t n_new)
{
    if (!b->prof || qlen == 0) return G5_OK;
    float ms = 0;
    G5_HIP(hipEventElapsedTime(&ms, b->ek0, b->ek1));
    const int64_t E = b->h_cnt[kEdges];
    b->exp_ms += ms;
    b->exp_launches += 1;
    b->exp_edges += E;
    b->exp_bytes += 16.0 * qlen + 12.0 * E + 12.0 * n_new;
    return G5_OK;
}

// After a bottom-up level: k_bottom_up time and its algorithmic bytes, 16 B
// row-offset pair per unvisited vertex scanned, 4 B column + 8 B frontier
// word per neighbour checked, 4 B parent per newly reached vertex, 16 B
// (visited read + next write) per 64-vertex word.
int account_bottom_up(g5_bfs *b, int64_t n_new)
{
    if (!b->prof) return G5_OK;
    float ms = 0;
    G5_HIP(hipEventElapsedTime(&ms, b->ek0, b->ek1));
    const int64_t E = b->h_cnt[kBuEdges], V = b->h_cnt[kBuVerts];
    b->bu_ms += ms;
    b->bu_launches += 1;
    b->bu_edges += E;
    b->bu_bytes += 16.0 * V + 12.0 * E + 4.0 * n_new + 16.0 * b->wc;
    return G5_OK;
}

int sync_counters(g5_bfs *b)
{
    G5_HIP(hipMemcpyAsync(b->h_cnt, b->d_cnt, kNCnt * 8, hipMemcpyDeviceToHost, b->s));
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

// bit v of word v / 64: row v of the (diagonal) block has no neighbours
__global__ void k_isolated(const int64_t *__restrict__ rowptr, int64_t wc, uint64_t *__restrict__ bits)
{
    const int lane = __lane_id();
    for (int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; w < wc;
         w += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const int64_t v = w * 64 + lane;
        const uint64_t word = __ballot(rowptr[v + 1] == rowptr[v]);
        if (lane == 0) bits[w] = word;
    }
}

// isolated: start with the block's edgeless vertices visited (g5_bfs_run on a
// diagonal block: they are never reached, so results do not change)
int reset_search(g5_bfs *b, int64_t root, bool isolated = false)
{
    g5_graph *g = b->g;
    hipStream_t s = b->s;
    const int64_t rr = root - g->row_start, rc = root - g->col_start;
    const int64_t root_row = (root >= 0 && rr >= 0 && rr < g->n_rows) ? rr : -1;
    const int64_t root_col = (root >= 0 && rc >= 0 && rc < g->n_cols) ? rc : -1;
    if (isolated) {
        if (!b->d_isolated) {
            G5_HIP(hipMalloc(&b->d_isolated, (size_t)(b->wc ? b->wc : 1) * 8));
            hipLaunchKernelGGL(k_isolated, g5::grid_for(b->wc * 64, 256, 4096), 256, 0, s,
                               (const int64_t *)g->d_rowptr, b->wc, b->d_isolated);
            G5_HIP(hipGetLastError());
        }
        G5_HIP(hipMemcpyAsync(b->d_visited, b->d_isolated, (size_t)b->wc * 8, hipMemcpyDeviceToDevice, s));
    } else {
        G5_HIP(g5::fill_u64(b->d_visited, b->wc, 0, s));
    }
    G5_HIP(g5::fill_u64(b->d_out, b->wc, 0, s));
    G5_HIP(g5::fill_u32(b->d_parent, g->n_cols, kNoParent, s));
    G5_HIP(g5::fill_i32(b->d_level, g->n_cols, -1, s));
    G5_HIP(g5::fill_i64(b->d_cnt, kNCnt, 0, s));
    hipLaunchKernelGGL(k_start, 1, 1, 0, s, root_row, root_col, b->d_queue, b->d_visited, b->d_out, b->d_level,
                       b->d_parent, b->d_cnt);
    G5_HIP(hipGetLastError());
    b->root = root;
    b->depth = 0;
    b->qlen = root_row >= 0 ? 1 : 0;
    b->inbits_valid = false;
    return G5_OK;
}

}  // namespace

extern "C" int g5_bfs_create(g5_graph *g, g5_bfs **out)
{
    G5_REQUIRE(g && out, G5_E_ARG, "g5_bfs_create: null argument");
    g5_bfs *b = new g5_bfs();
    b->g = g;
    b->wr = g->n_rows / 64;
    b->wc = g->n_cols / 64;
    auto fail = [&](int rc) {
        g5_bfs_free(b);
        return rc;
    };
    hipError_t e = hipStreamCreate(&b->s);
    const int64_t nmax = std::max(g->n_rows, g->n_cols);
    const size_t scr = g5::scan_scratch_words(nmax + 1) + g5::scan_scratch_words(b->wc + 1) + 64;
    if (e == hipSuccess) e = hipMalloc(&b->d_visited, (size_t)std::max<int64_t>(b->wc, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_out, (size_t)std::max<int64_t>(b->wc, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_next, (size_t)std::max<int64_t>(b->wc, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_parent, (size_t)std::max<int64_t>(g->n_cols, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&b->d_level, (size_t)std::max<int64_t>(g->n_cols, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&b->d_queue, (size_t)std::max<int64_t>(nmax, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&b->d_outq, (size_t)std::max<int64_t>(nmax, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&b->d_qoff, (size_t)(g->n_rows + 2) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_scratch, scr * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_cnt, kNCnt * 8);
    if (e == hipSuccess) e = hipHostMalloc(&b->h_cnt, kNCnt * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreate(&b->ek0);
    if (e == hipSuccess) e = hipEventCreate(&b->ek1);
    if (e != hipSuccess) return fail(g5::fail_hip(e, "g5_bfs_create allocation", __FILE__, __LINE__));
    *out = b;
    return G5_OK;
}

extern "C" int g5_bfs_free(g5_bfs *b)
{
    if (!b) return G5_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    void *dev[] = {b->d_visited, b->d_out,     b->d_next, b->d_parent, b->d_level, b->d_queue,
                   b->d_outq,    b->d_qoff,    b->d_scratch, b->d_cnt, b->d_words, b->d_isolated, b->d_inbits};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (b->h_cnt) (void)hipHostFree(b->h_cnt);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ek0) (void)hipEventDestroy(b->ek0);
    if (b->ek1) (void)hipEventDestroy(b->ek1);
    b->codec.release();
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return G5_OK;
}

extern "C" int g5_bfs_start(g5_bfs *b, int64_t root)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_start: null handle");
    int rc = reset_search(b, root);
    if (rc) return rc;
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_expand(g5_bfs *b, int *any_new)
{
    G5_REQUIRE(b && any_new, G5_E_ARG, "g5_bfs_expand: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_expand: g5_bfs_start not called");
    int rc = expand_level(b);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kOutLen, (int64_t)0);
    hipLaunchKernelGGL(k_count_bits, g5::grid_for(b->wc, 256, 2048), 256, 0, b->s,
                       (const uint64_t *)b->d_out, b->wc, b->d_cnt + kOutLen);
    G5_HIP(hipGetLastError());
    rc = sync_counters(b);
    if (rc) return rc;
    // the out bits are this block's newly reached columns
    rc = account_level(b, b->qlen, b->h_cnt[kOutLen]);
    if (rc) return rc;
    *any_new = b->h_cnt[kOutLen] > 0;
    return G5_OK;
}

extern "C" int g5_bfs_outqueue(g5_bfs *b, uint32_t *q, int64_t *n)
{
    G5_REQUIRE(b && n, G5_E_ARG, "g5_bfs_outqueue: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_outqueue: g5_bfs_start not called");
    G5_HIP(g5::bitmap_to_queue(b->d_out, b->wc, b->d_outq, b->d_cnt + kOutLen, b->d_scratch, b->s));
    int rc = sync_counters(b);
    if (rc) return rc;
    const int64_t len = b->h_cnt[kOutLen];
    G5_REQUIRE(len <= *n, G5_E_OVERFLOW, "g5_bfs_outqueue: queue of %lld entries exceeds capacity %lld",
               (long long)len, (long long)*n);
    *n = len;
    if (len) {
        G5_REQUIRE(q, G5_E_ARG, "g5_bfs_outqueue: null queue buffer");
        G5_HIP(hipMemcpy(q, b->d_outq, (size_t)len * 4, hipMemcpyDeviceToHost));
    }
    return G5_OK;
}

extern "C" int g5_bfs_set_outqueue(g5_bfs *b, const uint32_t *q, int64_t n)
{
    G5_REQUIRE(b && n >= 0, G5_E_ARG, "g5_bfs_set_outqueue: bad arguments");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_set_outqueue: g5_bfs_start not called");
    g5_graph *g = b->g;
    if (q) {
        G5_REQUIRE(n <= g->n_cols, G5_E_OVERFLOW, "g5_bfs_set_outqueue: queue longer than the column slice");
        for (int64_t i = 0; i < n; ++i)
            G5_REQUIRE(q[i] < (uint64_t)g->n_cols && (i == 0 || q[i] > q[i - 1]), G5_E_ARG,
                       "g5_bfs_set_outqueue: queue not sorted/unique/in range at %lld", (long long)i);
        G5_HIP(g5::fill_u64(b->d_out, b->wc, 0, b->s));
        if (n) {
            G5_HIP(hipMemcpyAsync(b->d_outq, q, (size_t)n * 4, hipMemcpyHostToDevice, b->s));
            G5_HIP(g5::queue_to_bitmap(b->d_outq, n, b->d_out, b->s));
        }
    }
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kNew, (int64_t)0);
    hipLaunchKernelGGL(k_commit, g5::grid_for(b->wc, 256, 2048), 256, 0, b->s, b->d_out, b->d_visited,
                       b->d_level, b->wc, (int32_t)(b->depth + 1), b->d_cnt,
                       (const int64_t *)nullptr);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_set_inqueue(g5_bfs *b, const uint32_t *q, int64_t n)
{
    G5_REQUIRE(b && n >= 0, G5_E_ARG, "g5_bfs_set_inqueue: bad arguments");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_set_inqueue: g5_bfs_start not called");
    g5_graph *g = b->g;
    if (q) {
        G5_REQUIRE(n <= g->n_rows, G5_E_OVERFLOW, "g5_bfs_set_inqueue: queue longer than the row slice");
        for (int64_t i = 0; i < n; ++i)
            G5_REQUIRE(q[i] < (uint64_t)g->n_rows && (i == 0 || q[i] > q[i - 1]), G5_E_ARG,
                       "g5_bfs_set_inqueue: queue not sorted/unique/in range at %lld", (long long)i);
        if (n) G5_HIP(hipMemcpyAsync(b->d_queue, q, (size_t)n * 4, hipMemcpyHostToDevice, b->s));
        hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kQLen, n);
        b->qlen = n;
    } else {
        G5_REQUIRE(g->n_rows == g->n_cols && g->row_start == g->col_start, G5_E_STATE,
                   "g5_bfs_set_inqueue: own out-queue as in-queue needs a diagonal block");
        G5_HIP(g5::bitmap_to_queue(b->d_out, b->wc, b->d_queue, b->d_cnt + kQLen, b->d_scratch, b->s));
        int rc = sync_counters(b);
        if (rc) return rc;
        b->qlen = b->h_cnt[kQLen];
    }
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(b->s));
    b->inbits_valid = false;
    b->depth++;
    return G5_OK;
}

extern "C" int g5_bfs_result(g5_bfs *b, int64_t *pred, int32_t *level)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_result: null handle");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_result: no search");
    g5_graph *g = b->g;
    if (pred) {
        int64_t *d = nullptr;
        G5_HIP(hipMallocAsync((void **)&d, (size_t)std::max<int64_t>(g->n_cols, 1) * 8, b->s));
        hipLaunchKernelGGL(k_result, g5::grid_for(g->n_cols, 256, 8192), 256, 0, b->s,
                           (const uint32_t *)b->d_parent, g->n_cols, g->row_start, g->col_start, d);
        G5_HIP(hipGetLastError());
        G5_HIP(hipMemcpyAsync(pred, d, (size_t)g->n_cols * 8, hipMemcpyDeviceToHost, b->s));
        G5_HIP(hipFreeAsync(d, b->s));
    }
    if (level) G5_HIP(hipMemcpyAsync(level, b->d_level, (size_t)g->n_cols * 4, hipMemcpyDeviceToHost, b->s));
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_out_bitmap(g5_bfs *b, const uint64_t **d_bits, int64_t *n_words)
{
    G5_REQUIRE(b && d_bits && n_words, G5_E_ARG, "g5_bfs_out_bitmap: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_out_bitmap: g5_bfs_start not called");
    *d_bits = b->d_out;
    *n_words = b->wc;
    return G5_OK;
}

extern "C" int g5_bfs_set_out_bitmap(g5_bfs *b, const uint64_t *d_bits)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_set_out_bitmap: null handle");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_set_out_bitmap: g5_bfs_start not called");
    if (d_bits && d_bits != b->d_out && b->wc)
        G5_HIP(hipMemcpyAsync(b->d_out, d_bits, (size_t)b->wc * 8, hipMemcpyDeviceToDevice, b->s));
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kNew, (int64_t)0);
    hipLaunchKernelGGL(k_commit, g5::grid_for(b->wc, 256, 2048), 256, 0, b->s, b->d_out, b->d_visited,
                       b->d_level, b->wc, (int32_t)(b->depth + 1), b->d_cnt,
                       (const int64_t *)nullptr);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_set_in_bitmap(g5_bfs *b, const uint64_t *d_bits)
{
    G5_REQUIRE(b && d_bits, G5_E_ARG, "g5_bfs_set_in_bitmap: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_set_in_bitmap: g5_bfs_start not called");
    G5_HIP(g5::bitmap_to_queue(d_bits, b->wr, b->d_queue, b->d_cnt + kQLen, b->d_scratch, b->s));
    if (!b->d_inbits) G5_HIP(hipMalloc(&b->d_inbits, (size_t)(b->wr ? b->wr : 1) * 8));
    if (b->wr) G5_HIP(hipMemcpyAsync(b->d_inbits, d_bits, (size_t)b->wr * 8, hipMemcpyDeviceToDevice, b->s));
    int rc = sync_counters(b);
    if (rc) return rc;
    b->qlen = b->h_cnt[kQLen];
    b->inbits_valid = true;
    b->depth++;
    return G5_OK;
}

extern "C" int g5_bfs_expand_bottom_up(g5_bfs *b, const g5_graph *gt, int *any_new)
{
    G5_REQUIRE(b && gt && any_new, G5_E_ARG, "g5_bfs_expand_bottom_up: null argument");
    G5_REQUIRE(b->depth >= 0 && b->inbits_valid, G5_E_STATE,
               "g5_bfs_expand_bottom_up: needs a row frontier from g5_bfs_set_in_bitmap");
    const g5_graph *g = b->g;
    G5_REQUIRE(gt->n_rows == g->n_cols && gt->n_cols == g->n_rows && gt->row_start == g->col_start &&
                   gt->col_start == g->row_start,
               G5_E_ARG, "g5_bfs_expand_bottom_up: gt is not the transposed block");
    hipStream_t s = b->s;
    G5_HIP(g5::fill_i64(b->d_cnt + kNew, kBuVerts - kNew + 1, 0, s));
    if (b->prof) G5_HIP(hipEventRecord(b->ek0, s));
    // rows of gt = this block's columns, its columns = this block's rows: the
    // kernel's frontier bits are the row frontier, its parents local row ids
    hipLaunchKernelGGL(gt->n_rows < kBuLarge ? k_bottom_up<4> : k_bottom_up<2>, g5::grid_for(b->wc, 4, 4096), 256, 0,
                       s, (const int64_t *)gt->d_rowptr, (const uint32_t *)gt->d_cols,
                       (const uint64_t *)b->d_visited, (const uint64_t *)b->d_inbits, b->d_out, b->d_parent, b->wc,
                       b->d_cnt);
    G5_HIP(hipGetLastError());
    if (b->prof) G5_HIP(hipEventRecord(b->ek1, s));
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, s, b->d_cnt, (int)kOutLen, (int64_t)0);
    hipLaunchKernelGGL(k_count_bits, g5::grid_for(b->wc, 256, 2048), 256, 0, s, (const uint64_t *)b->d_out, b->wc,
                       b->d_cnt + kOutLen);
    G5_HIP(hipGetLastError());
    int rc = sync_counters(b);
    if (rc) return rc;
    rc = account_bottom_up(b, b->h_cnt[kOutLen]);
    if (rc) return rc;
    *any_new = b->h_cnt[kOutLen] > 0;
    return G5_OK;
}

extern "C" int g5_bfs_pred_device(g5_bfs *b, int64_t *d_pred)
{
    G5_REQUIRE(b && d_pred, G5_E_ARG, "g5_bfs_pred_device: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_pred_device: no search");
    g5_graph *g = b->g;
    hipLaunchKernelGGL(k_result, g5::grid_for(g->n_cols, 256, 8192), 256, 0, b->s, (const uint32_t *)b->d_parent,
                       g->n_cols, g->row_start, g->col_start, d_pred);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_level_device(g5_bfs *b, int32_t *d_level)
{
    G5_REQUIRE(b && d_level, G5_E_ARG, "g5_bfs_level_device: null argument");
    G5_REQUIRE(b->depth >= 0, G5_E_STATE, "g5_bfs_level_device: no search");
    G5_HIP(hipMemcpyAsync(d_level, b->d_level, (size_t)b->g->n_cols * 4, hipMemcpyDeviceToDevice, b->s));
    G5_HIP(hipStreamSynchronize(b->s));
    return G5_OK;
}

extern "C" int g5_bfs_nedge(g5_bfs *b, int64_t *nedge)
{
    G5_REQUIRE(b && nedge, G5_E_ARG, "g5_bfs_nedge: null argument");
    G5_REQUIRE(b->g->n_rows == b->g->n_cols, G5_E_STATE, "g5_bfs_nedge: needs a 1x1 block");
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kOutLen, (int64_t)0);
    hipLaunchKernelGGL(k_nedge, g5::grid_for(b->g->n_rows, 256, 4096), 256, 0, b->s,
                       (const uint32_t *)b->g->d_gdeg, (const int32_t *)b->d_level, b->g->n_rows,
                       b->d_cnt + kOutLen);
    G5_HIP(hipGetLastError());
    int rc = sync_counters(b);
    if (rc) return rc;
    *nedge = b->h_cnt[kOutLen] / 2;
    return G5_OK;
}

extern "C" int g5_bfs_summary(g5_bfs *b, int64_t *nedge, uint64_t *checksum)
{
    G5_REQUIRE(b && nedge && checksum, G5_E_ARG, "g5_bfs_summary: null argument");
    G5_REQUIRE(b->g->n_rows == b->g->n_cols && b->g->row_start == b->g->col_start, G5_E_STATE,
               "g5_bfs_summary: needs a diagonal (1x1) block");
    // kOutLen and kNewEdges as the two sums (free between levels)
    static_assert(kNewEdges == kOutLen + 1, "the summary's two slots must be adjacent");
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kOutLen, (int64_t)0);
    hipLaunchKernelGGL(k_set_cnt, 1, 1, 0, b->s, b->d_cnt, (int)kNewEdges, (int64_t)0);
    hipLaunchKernelGGL(k_summary, g5::grid_for(b->g->n_rows, 256, 4096), 256, 0, b->s,
                       (const uint32_t *)b->g->d_gdeg, (const int32_t *)b->d_level, (const uint32_t *)b->d_parent,
                       b->g->n_rows, b->d_cnt + kOutLen);
    G5_HIP(hipGetLastError());
    int rc = sync_counters(b);
    if (rc) return rc;
    *nedge = b->h_cnt[kOutLen] / 2;
    *checksum = (uint64_t)b->h_cnt[kNewEdges];
    return G5_OK;
}

extern "C" int g5_bfs_run(g5_bfs *b, int64_t root, int flags, int32_t *nlevels, float *ms)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_run: null handle");
    g5_graph *g = b->g;
    G5_REQUIRE(g->n_rows == g->n_cols && g->row_start == g->col_start, G5_E_STATE,
               "g5_bfs_run: the resident loop needs a diagonal (1x1) block");
    G5_REQUIRE(root >= 0 && root < g->n_rows, G5_E_ARG, "g5_bfs_run: root %lld out of range", (long long)root);
    const bool codec = (flags & G5_RUN_CODEC) != 0;
    const bool topdown_only = (flags & G5_RUN_TOPDOWN) != 0;
    if (codec && !b->d_words) {
        b->words_cap = g5_s4_max_words((size_t)g->n_cols);
        G5_HIP(hipMalloc(&b->d_words, b->words_cap * 4));
        int rc = b->codec.reserve((int64_t)g->n_cols);
        if (rc) return rc;
    }
    hipStream_t s = b->s;
    G5_HIP(hipStreamSynchronize(s));
    G5_HIP(hipEventRecord(b->ev0, s));
    int rc = reset_search(b, root, true);
    if (rc) return rc;
    const int64_t n = g->n_rows;
    int64_t m_u = g->nnz;  // edges not yet checked top-down
    bool bu = false;       // the first level is always top-down
    for (;;) {
        G5_HIP(g5::fill_i64(b->d_cnt + kNew, kBuVerts - kNew + 1, 0, s));
        uint64_t *fresh = b->d_out;
        if (bu) {
            if (b->prof) G5_HIP(hipEventRecord(b->ek0, s));
            hipLaunchKernelGGL(g->n_rows < kBuLarge ? k_bottom_up<4> : k_bottom_up<2>, g5::grid_for(b->wc, 4, 4096),
                               256, 0, s, (const int64_t *)g->d_rowptr,
                               (const uint32_t *)g->d_cols, (const uint64_t *)b->d_visited,
                               (const uint64_t *)b->d_out, b->d_next, b->d_parent, b->wc, b->d_cnt);
            G5_HIP(hipGetLastError());
            if (b->prof) G5_HIP(hipEventRecord(b->ek1, s));
            fresh = b->d_next;
        } else {
            rc = expand_level(b);
            if (rc) return rc;
        }
        // a bottom-up level already summed its new vertices' degrees
        hipLaunchKernelGGL(k_commit, g5::grid_for(b->wc, 256, 2048), 256, 0, s, fresh, b->d_visited, b->d_level,
                           b->wc, (int32_t)(b->depth + 1), b->d_cnt,
                           bu ? (const int64_t *)nullptr : (const int64_t *)g->d_rowptr);
        G5_HIP(hipGetLastError());
        if (bu) std::swap(b->d_out, b->d_next);
        rc = sync_counters(b);
        if (rc) return rc;
        // the previous level's codec round trip has completed by now
        G5_REQUIRE(b->h_cnt[kCodecErr] == 0, G5_E_STREAM, "g5_bfs_run: codec round trip lost values");
        const int64_t n_new = b->h_cnt[kNew];
        rc = bu ? account_bottom_up(b, n_new) : account_level(b, b->qlen, n_new);
        if (rc) return rc;
        if (n_new == 0) break;
        // the next level's direction
        const int64_t m_f = b->h_cnt[kNewEdges];
        bool bu_next = bu;
        if (!topdown_only) {
            if (!bu && m_f > m_u / b->alpha) bu_next = true;
            else if (bu && n_new < n / b->beta) bu_next = false;
        }
        if (!bu_next) m_u -= m_f;
        // the out-bitmap holds exactly the new vertices; a queue of them is
        // needed by a top-down level and by the codec
        b->qlen = n_new;
        if (!bu_next || codec)
            G5_HIP(g5::bitmap_to_queue(b->d_out, b->wc, b->d_queue, b->d_cnt + kQLen, b->d_scratch, s));
        if (codec) {
            // encode -> decode in place with no host synchronisation; checked
            // at the next level's counter read
            rc = b->codec.roundtrip(b->d_queue, b->qlen, b->d_words, b->words_cap, b->d_queue, b->d_cnt + kCodecErr, s);
            if (rc) return rc;
        }
        bu = bu_next;
        b->depth++;
    }
    G5_HIP(hipEventRecord(b->ev1, s));
    G5_HIP(hipEventSynchronize(b->ev1));
    if (ms) G5_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
    if (nlevels) *nlevels = b->depth + 1;
    return G5_OK;
}

extern "C" int g5_bfs_tune(g5_bfs *b, int64_t alpha, int64_t beta)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_tune: null handle");
    G5_REQUIRE(alpha >= 1 && beta >= 1, G5_E_ARG, "g5_bfs_tune: alpha and beta must be >= 1");
    b->alpha = alpha;
    b->beta = beta;
    return G5_OK;
}

extern "C" int g5_bfs_profile(g5_bfs *b, int enable)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_profile: null handle");
    b->prof = enable != 0;
    b->exp_ms = b->exp_bytes = 0;
    b->exp_launches = b->exp_edges = 0;
    b->bu_ms = b->bu_bytes = 0;
    b->bu_launches = b->bu_edges = 0;
    return G5_OK;
}

extern "C" int g5_bfs_kernel_stats(g5_bfs *b, int kernel, double *ms, int64_t *launches, double *bytes,
                                   int64_t *edges)
{
    G5_REQUIRE(b && (kernel == 0 || kernel == 1), G5_E_ARG, "g5_bfs_kernel_stats: kernel must be 0 or 1");
    if (kernel == 0) return g5_bfs_stats(b, ms, launches, bytes, edges);
    if (ms) *ms = b->bu_ms;
    if (launches) *launches = b->bu_launches;
    if (bytes) *bytes = b->bu_bytes;
    if (edges) *edges = b->bu_edges;
    return G5_OK;
}

extern "C" int g5_bfs_stats(g5_bfs *b, double *ms, int64_t *launches, double *bytes, int64_t *edges)
{
    G5_REQUIRE(b, G5_E_ARG, "g5_bfs_stats: null handle");
    if (ms) *ms = b->exp_ms;
    if (launches) *launches = b->exp_launches;
    if (bytes) *bytes = b->exp_bytes;
    if (edges) *edges = b->exp_edges;
    return G5_OK;
}
