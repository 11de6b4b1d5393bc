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
// (edge-balanced tiles, atomicOr; the edge that sets a pair's bit first
// leaves its source in the pair's parent slot), then the new (vertex, root)
// parent slots are lowered with atomicMin over the same edges -- the same
// minimum-id parent -- or, when the new vertices have fewer edges than the
// frontier, by a scan of the new vertices' rows (k_min_rows).  A 1-bit-per-vertex active bitmap lets gathers skip vertices with
// an empty frontier word.
//
// Split levels: a pull scan stops only when all of the vertex's unvisited
// lanes are found, so one lane whose frontier is still small (a root of low
// degree, a root that joined late) makes most vertices scan their whole
// neighbour list.  Pull levels therefore work only on the lanes whose
// frontier is non-empty, and right after a push level (which counts each
// lane's frontier edges) or after lanes joined (whose frontier is their
// root), the lanes with fewer than nnz / lane_ratio frontier edges are
// pushed instead in the same level: the block and wave kernels run on the
// other lanes, then the push kernels, restricted to the pushed lanes and to
// targets in the block and wave bins, OR their bits into next and set their
// parents with atomicMin.  The thread kernel's rows are short and scanned to
// the end at such a level anyway, so it pulls every lane, the pushed ones
// included, and the push skips its vertices.  Each (vertex, root) pair is
// found by one kernel, so it gets the same level and minimum-id parent as
// before.  As every row is scanned to its end there, split levels bound the
// thread bin by a degree of their own (split_max, at most thread_max; the
// wave kernel's coalesced steps serve the mid-degree rows better then, while
// on pull levels, where most scans end early, the thread bin pays up to
// higher degrees), start the wave kernel with
// a whole step, and, when the block and wave bins' vertices times the pushed
// lanes fit in n words, push through compact per-pair candidates (one
// atomicMin pass and a commit) instead of the two passes of a push level.
//
// Relabelling: the search runs on a copy of the block whose vertices are
// renumbered by descending degree class (powers of two; ties by id).
// Frontier words of the hubs,
// which most neighbour lists point at, then sit in a small id range that the
// L2 and the Infinity Cache hold, instead of being scattered over the whole
// n * 8 bytes.  Every row keeps its neighbours in ascending ORIGINAL id
// order and parents are recorded as original ids, so the first hit of a
// scan is still the minimum-id parent.  The degree bins are contiguous id
// ranges (exactly so for bounds that are powers of two, as the autotuner's
// are; other bounds move some vertices into a neighbouring bin, which gives
// the same results): one thread per vertex up to thread_max neighbours, one wavefront
// (64 neighbours per step, first hits from an exclusive prefix-OR across
// lanes) up to wave_max, one 256-thread workgroup beyond (Tune; 32 and 2048
// by default).  Each pull level's launch
// sequence is captured once into a hipGraph per frontier parity and
// replayed; the level number lives in device memory.
//
// Output in HBM, final when g5_msbfs_run returns: the 64 predecessor maps
// as one vertex-major array parent int32[n][64] and the 64 level maps as
// level uint8[n][64], both indexed by the ORIGINAL vertex id, slot = lane.
// A parent is an original id (root -> root, -1 unreached); a level byte is
// the distance from the root (255 unreached).  The kernels that find a
// (vertex, root) pair write both at once: the pull kernels take a parent's
// original id from cols_orig at the hit's row position (the rows they scan),
// once per distinct parent (a hit), not once per slot, and store each
// vertex's new slots with one masked,
// non-temporal store of its 256 B record and one of its 64 B level record
// (the records are not read again during the search, so they should not
// displace frontier words from the caches).  Slots of pairs the search did
// not reach still hold an earlier search's values until k_ms_finalize, the
// last pass of a batch run, sets them; edgeless vertices' records stay at
// the unreached values except where a root sat, which the next run clears.
// Levels need one byte: a search deeper than 254 levels reports its level
// bytes as not final, and g5_msbfs_result then derives levels from the
// parent tree by pointer jumping, as the reference validator derives
// depths from the predecessor map (validate/validate.h:107-232).
#include <algorithm>
#include <vector>

#include "g5_internal.hpp"

namespace {

// degree bins in launch order: one workgroup per vertex (degree above
// wave_max), one wavefront per vertex (thread_max < degree <= wave_max), one
// thread per vertex (1..thread_max); the bounds are tunable (Tune)
constexpr int kBins = 3;
constexpr int kPushUnit = kBins;  // accounting unit of push levels
constexpr int kUnits = kBins + 1;
constexpr int kBlock = 256;
constexpr int kLevelRec = 12;  // per level: ms[4], edges[4], new pairs, pushed, frontier vertices, edges
// Shape parameters (g5_msbfs_tune / g5_msbfs_autotune).  Any values give the
// same results; they only move work between kernels.
struct Tune {
    int64_t thread_max = 32;   // degree bound of the thread-per-vertex bin
    int64_t wave_max = 2048;   // degree bound of the wavefront-per-vertex bin
    int64_t push_ratio = 16;   // push when the frontier's edges < nnz / push_ratio
    int64_t active_ratio = 2;  // pull levels consult the active bitmap when the
                               // frontier's edges < nnz / active_ratio
    int64_t lane_ratio = 64;   // a pull level pushes the lanes whose frontier
                               // edges < nnz / lane_ratio (0: never)
    int64_t split_max = 16;    // split levels bound the thread bin at
                               // min(split_max, thread_max)
};

// Level counters.  Per accounting unit u (the kBins pull bins, then
// kPushUnit) kUnitSlots accumulated slots at u * kUnitSlots: new (vertex,
// root) pairs, edges read, vertices worked on, frontier words gathered, and
// the next frontier's vertices and edges; then the OR of the next frontier's
// words (the lanes still searching).  Then set-once slots: the level number,
// the newly reached queue length (push), the frontier queue length and its
// edge total (push), the lanes the block and wave kernels work on, the lanes
// in use, the first word of level flags the level before wrote, the lanes
// the thread kernel works on and the wave kernel's first step (the last five
// written before each pull level, k_set_pull_lanes).
enum MsCnt {
    kNewBits = 0, kEdgesRead = 1, kVisitedVerts = 2, kGathers = 3, kNextVerts = 4, kNextEdges = 5,
    kUnitSlots = 6,
    kLive = kUnits * kUnitSlots,
    kAccum = kLive + 1,
    kDepth = kAccum, kNq = kAccum + 1, kFqLen = kAccum + 2, kFqEdges = kAccum + 3, kPullLanes = kAccum + 4,
    kUsed = kAccum + 5, kFlagWord = kAccum + 6, kThreadLanes = kAccum + 7, kWaveStep = kAccum + 8,
    kCntStride = 40
};
// Many workgroups adding into one address serialise in one L2 channel, so
// the accumulated slots are striped: workgroup b adds into stripe b %
// kStripes (stripes 320 B apart) and k_level_end folds the stripes into
// stripe 0, which also holds the set-once slots.
constexpr int kStripes = 32;
static_assert(kWaveStep < kCntStride, "counter stripe too small");
// Per-lane frontier edges (counted by push levels, from their new-vertex
// queue): kLaneStripes stripes of 64 slots, summed on the host.
constexpr int kLaneStripes = 8;
constexpr int kHostSlots = kCntStride + 64;  // stripe 0, then the folded per-lane edges
// edgeless roots remembered between batch runs (beyond: every edgeless
// vertex's record is cleared)
constexpr int64_t kDirtyCap = 1 << 16;
static_assert(kStripes * kCntStride >= kHostSlots, "per-lane edges do not fit the folded stripes");

// The degree-ordered copy of the block.
struct Relabeled {
    uint32_t *new2old = nullptr, *old2new = nullptr;
    int64_t *rowptr = nullptr;
    uint32_t *cols = nullptr;  // new ids, rows in ascending original-id order
    // the same entries' original ids: a hit at row position j names its
    // parent as cols_orig[j], read from the rows the wave is scanning anyway
    // instead of by a random new2old gather
    uint32_t *cols_orig = nullptr;
    // per vertex, its row's first two neighbours (low word first; 0 past the
    // row's end): the thread kernel's first step reads them from here, 64
    // vertices' in 512 contiguous bytes, instead of from 64 row starts
    uint64_t *head = nullptr;
    // the same two neighbours' original ids: the parent a hit among them
    // names, read with the vertex's other coalesced words instead of from
    // cols_orig at 64 row starts
    uint64_t *head_orig = nullptr;
    uint32_t *gdeg = nullptr;
    int64_t bin_first[kBins + 1] = {0};  // bin b holds new ids [bin_first[b], bin_first[b+1])
    int64_t split_first[kBins + 1] = {0};  // the same on split levels (min(split_max, thread_max))
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
    // records [n][64] by original vertex id: slot r of v holds lane r's
    // parent of v (original id) and its level byte
    int32_t *d_parent = nullptr;
    uint8_t *d_level = nullptr;
    int64_t *d_cnt = nullptr, *h_cnt = nullptr;
    int64_t *d_lane = nullptr;  // [kLaneStripes][64] per-lane frontier edges
    int64_t *d_scratch = nullptr;
    int64_t *d_roots = nullptr;
    Relabeled rl;
    hipGraphExec_t graph[2][2] = {};  // [frontier parity][active-bitmap check]
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nroots = 0;
    int64_t roots[64];
    int32_t lane_of[64];  // batch run: root r's lane (continuous mode: lane r itself)
    // accounting
    bool prof = false;
    hipEvent_t ek[kBins + 1] = {};
    double k_ms[kUnits] = {}, k_bytes[kUnits] = {};
    int64_t k_launches[kUnits] = {};
    // per level of the last profiled run (kLevelRec doubles each)
    std::vector<double> lvl;
    int mode = 0;  // 0 direction chosen per level, 1 pull only, 2 push only
    Tune tune;
    // level state: the parity of the frontier the next level reads and the
    // frontier's active vertices / edges (edges < 0: unknown, take push)
    int p = 0;
    int64_t act_verts = 0, act_edges = -1;
    // the lanes with a non-empty frontier, and the lanes whose frontier edges
    // for the frontier the next level reads lane_edges holds (every lane
    // after a push level, which counts them from its new-vertex queue; the
    // lanes that joined since)
    uint64_t front_lanes = 0;
    uint64_t lane_known = 0;
    // a split may pay: the previous level pushed (a lane may still lag
    // behind the others) or lanes joined since
    bool split_ok = false;
    int64_t lane_edges[64] = {};
    hipEvent_t ek_push[2] = {};  // split levels (profiled): around the pushed lanes' kernels
    int64_t split_levels = 0;
    // the last push level's frontier queue, which lists the only non-zero
    // words of the buffer the next push level writes (listed_ok: the level
    // just run was that push level, or both buffers are clear)
    uint32_t *d_fq2 = nullptr;
    uint32_t *listed = nullptr;
    int64_t listed_len = 0;
    bool listed_ok = false;
    int64_t flag_word = INT64_MAX;     // the first word of lvflag the last level wrote (INT64_MAX: none)
    uint64_t *d_lvflag = nullptr;      // bit v: v's level record holds the next level's prediction
    bool last_empty = false;  // the level just run found no new (vertex, root) pair
    // continuous mode (g5_msbfs_open): lanes in use
    bool cont = false;
    uint64_t busy = 0;
    // join: roots [0, 64), lanes [64, 128), added frontier [128, 130), root
    // degrees [130, 194); batch run: the roots' reach [0, 64)
    int64_t *d_join = nullptr;
    unsigned long long *d_sums = nullptr;  // leave: checksums [0, 64), generated-degree sums [64, 128)
    // the last batch run: device time of the search and of its finalising
    // pass (ev_fin between them), and whether its level bytes are final
    // (depth below 255)
    hipEvent_t ev_fin = nullptr;
    float search_ms = 0, finalize_ms = 0;
    bool levels_final = false;
    int32_t run_levels = 0;
    // roots since the last batch run's start (batch roots and joined roots,
    // original ids): an edgeless one left its slot in a record that must read
    // "unreached" again; d_dirty holds them on the device for the clearing pass
    std::vector<int64_t> dirty;
    int64_t *d_dirty = nullptr;
    // continuous mode: each lane's root and the levels run since it joined
    int64_t lane_root[64] = {};
    int64_t lane_steps[64] = {};
};

namespace {

// Every slot holds an original id and is written by one level only, the
// level that reaches its (vertex, root) pair; within a level the pulled and
// pushed lanes are disjoint, and the push's atomicMin pass compares
// original ids.

// slots `bits` of the record of vertex o (original id): parent uo, level lv
__device__ __forceinline__ void write_slots(int32_t *__restrict__ parent, uint8_t *__restrict__ level, int64_t o,
                                            uint64_t bits, uint32_t uo, uint8_t lv)
{
    while (bits) {
        const int r = __ffsll((long long)bits) - 1;
        parent[o * 64 + r] = (int32_t)uo;
        level[o * 64 + r] = lv;
        bits &= bits - 1;
    }
}

// the level byte of the pairs the running level reaches (the level number
// lives in device memory, so captured level graphs replay unchanged)
__device__ __forceinline__ uint8_t level_byte(const int64_t *__restrict__ cnt) { return (uint8_t)cnt[kDepth]; }

// lane-to-wave broadcast; l must be wave-uniform
__device__ __forceinline__ uint32_t lane_read(uint32_t x, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}
__device__ __forceinline__ uint64_t lane_read64(uint64_t x, int l)
{
    return ((uint64_t)lane_read((uint32_t)(x >> 32), l) << 32) | lane_read((uint32_t)x, l);
}

// Adds the workgroup's totals of N per-thread counts to slots c[0..N-1] of
// the workgroup's counter stripe, one atomic per nonzero slot (every thread
// of the kBlock-thread workgroup must call it).
template <int N>
__device__ __forceinline__ void block_add(int64_t *__restrict__ c, const int64_t (&x)[N])
{
    __shared__ int64_t s_sum[kBlock / 64][N];
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int64_t t = g5_wave_sum(x[k]);
        if (__lane_id() == 0) s_sum[wv][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        int64_t t = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += s_sum[w][threadIdx.x];
        if (t)
            atomicAdd((unsigned long long *)&c[(blockIdx.x % kStripes) * kCntStride + threadIdx.x],
                      (unsigned long long)t);
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

// One thread per vertex (degree <= thread_max): a sequential ascending scan.  Each
// thread keeps its first kHits hits (the parent's original id, read from
// cols_orig at the hit's row position or from head_orig, and the new lanes)
// in registers; once the wave's 64 vertices are scanned, the wave writes them
// out vertex by vertex, lane r storing root r's parent and level, so every
// store instruction covers one vertex's new slots in its contiguous 256 B
// parent record (or 64 B level record) instead of 64 scattered slots.
// Hits beyond kHits (rare: they need more than kHits distinct first-hit
// neighbours) are stored directly.  The scan loads kThreadUnroll neighbours
// and their frontier words before using any of them, so that several
// gathers are in flight per thread; its first step loads only kThreadFirst,
// because on the dense levels most vertices are done after one or two
// neighbours and every gather past the early exit is a wasted line fill.
// That first step takes the two neighbours from the head array (one 8 B word
// per vertex, contiguous across the wave) rather than from the row starts,
// each in a line of its own once rows exceed 32 neighbours.
// kActive (all pull kernels): skip neighbours whose bit in the active bitmap
// is clear before gathering their frontier word (pays only when few sources
// are active).  The block and wave kernels work on the lanes in
// cnt[kPullLanes] only, the thread kernel on those in cnt[kThreadLanes] (the
// same, except on a split level: there it takes every lane, the pushed ones
// too, since its rows are scanned to the end anyway): a vertex is done when
// every unvisited lane among them is found.
constexpr int kHits = 6;  // 4: +0.5 ms at level 3 (spilled hits' scattered stores), 8: 3 waves per SIMD, slower
constexpr int kThreadUnroll = 8;  // 4: +0.6 ms at level 3, 16: +0.7 ms at level 4; 12 after the short first step: the same
constexpr int kThreadFirst = 2;  // 1: same time, more neighbours read

// the record outputs of a pull level: parent and level records by original
// id, the level byte of the pairs it reaches, new2old and the columns'
// original ids
struct PullOut {
    int32_t *__restrict__ parent;
    uint8_t *__restrict__ level;
    const uint32_t *__restrict__ new2old;
    const uint32_t *__restrict__ cols_orig;
    uint8_t lv;
};

// One step of a thread's scan: neighbours [j, j + K) of the row ending at e
// (kHead: the row's first two, j at the row's start, from the head array).
// A kept hit keeps its parent's original id in hu, loaded from cols_orig at
// the hit's row position, except a hit of the head step: that one keeps its
// head slot (0, 1) and sets its bit in from_head, and its original id comes
// from head_orig, which the wave reads coalesced.
template <int K, bool kActive, bool kHead>
__device__ __forceinline__ int64_t thread_scan_step(int64_t j, int64_t e, uint32_t v, uint64_t unv, uint64_t &found,
                                                    int &nh, uint32_t (&hu)[kHits], uint64_t (&hf)[kHits],
                                                    uint32_t &from_head, int64_t &ngath,
                                                    const uint32_t *__restrict__ cols,
                                                    const uint64_t *__restrict__ head,
                                                    const uint64_t *__restrict__ frontier,
                                                    const uint64_t *__restrict__ active, const PullOut &out)
{
    static_assert(!kHead || K == 2, "the head array holds two neighbours per vertex");
    uint32_t u[K];
    uint64_t fr[K];
    if (kHead) {
        const uint64_t h = head[v];
        u[0] = (uint32_t)h;
        u[K - 1] = (uint32_t)(h >> 32);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) u[k] = j + k < e ? cols[j + k] : 0u;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool take = j + k < e && (!kActive || (active[u[k] >> 6] >> (u[k] & 63) & 1));
        fr[k] = take ? frontier[u[k]] : 0ull;
        ngath += take;
    }
    // in neighbour order: the first hit of a lane is its parent
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t f = fr[k] & unv & ~found;
        if (!f) continue;
        found |= f;
        if (nh < kHits) {
#pragma unroll
            for (int h = 0; h < kHits; ++h)
                if (h == nh) {
                    hu[h] = kHead ? (uint32_t)k : out.cols_orig[j + k];
                    hf[h] = f;
                }
            if (kHead) from_head |= 1u << nh;
            ++nh;
        } else {
            write_slots(out.parent, out.level, out.new2old[v], f, out.cols_orig[j + k], out.lv);
        }
    }
    return j + K;
}

// The hits a wave step keeps for its write-out: per thread (vertex) the
// number kept, their parents (after translate(): all original ids), their new
// lanes, the vertex's original id and all lanes found at this level (spilled
// hits included).
constexpr uint32_t kLvFlagged = 1;  // the level bytes of this level's lanes were predicted
constexpr uint32_t kLvWhole = 2;    // store the whole level record with the next level predicted
struct KeptHits {
    int nh = 0;
    uint32_t hu[kHits] = {};
    uint64_t hf[kHits] = {};
    uint32_t o = 0;
    uint32_t lvmode = 0;
    uint64_t found = 0;
    uint64_t own = 0;  // kLvWhole: the lanes that reached the vertex at the level before
};

// Issues the loads of the head-step hits' original ids (from head_orig,
// coalesced; the other hits hold theirs already) and of the vertex's own.
__device__ __forceinline__ void translate(KeptHits &k, uint32_t v, uint32_t from_head,
                                          const uint32_t *__restrict__ new2old,
                                          const uint64_t *__restrict__ head_orig)
{
    if (!k.nh) return;
    k.o = new2old[v];
    if (!from_head) return;
    const uint64_t ho = head_orig[v];
#pragma unroll
    for (int h = 0; h < kHits; ++h)
        if (h < k.nh && (from_head >> h & 1)) k.hu[h] = (uint32_t)(ho >> (32 * k.hu[h]));
}

// The wave writes the kept hits of its 64 vertices out, one vertex per step,
// lane = root: one coalesced store of the vertex's new slots in its 256 B
// parent record and one in its 64 B level record.  Lanes found at this level
// through spilled hits already hold their parents and are left alone; slots
// no level reaches are set by k_ms_finalize.  (Storing a vertex's whole
// records at its first event, the other lanes as unreached, wrote more bytes
// than it saved in partial sectors once lanes are ranked by reach: the lanes
// a vertex gains at one level mostly share their sectors.)
// The level record is one 64 B sector that a vertex's lanes would write in
// pieces at two or three levels.  Where all of a vertex's earlier lanes came
// at the level before (or none did), the wave stores the whole record instead
// -- this level for the lanes found now, the level before for those, and the
// next level, predicted, for the lanes still open -- and flags the vertex in
// lvflag; at the next level (when it runs the pull kernels too) the lanes a
// flagged vertex gains need no level store, and k_ms_finalize overwrites the
// prediction of lanes that never come.  A lane found later than predicted is
// stored as usual.
__device__ __forceinline__ void write_out(const KeptHits &k, int lane, int32_t *__restrict__ parent,
                                          uint8_t *__restrict__ level, uint8_t lv)
{
    uint64_t todo = __ballot(k.nh != 0);
    while (todo) {
        const int src = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const int nhj = (int)lane_read((uint32_t)k.nh, src);
        uint32_t p = 0;
        bool mine = false;
#pragma unroll
        for (int h = 0; h < kHits; ++h) {
            if (h >= nhj) break;
            const uint64_t fk = lane_read64(k.hf[h], src);
            const uint32_t uk = lane_read(k.hu[h], src);
            if (fk >> lane & 1) {
                p = uk;
                mine = true;
            }
        }
        const uint32_t mode = lane_read(k.lvmode, src);
        const int64_t slot = (int64_t)lane_read(k.o, src) * 64 + lane;
        if (mine) __builtin_nontemporal_store((int32_t)p, &parent[slot]);
        if (mode & kLvWhole) {
            const uint64_t fs = lane_read64(k.found, src), os = lane_read64(k.own, src);
            const uint8_t b = (fs >> lane & 1) ? lv : (os >> lane & 1) ? (uint8_t)(lv - 1) : (uint8_t)(lv + 1);
            __builtin_nontemporal_store(b, &level[slot]);
        } else if (mine && !(mode & kLvFlagged)) {
            __builtin_nontemporal_store(lv, &level[slot]);
        }
    }
}

// The wave's steps are software-pipelined: a step's write-out waits until the
// end of the next step's scan, so the loads it needs (the vertices' original
// ids, the head-step parents' from head_orig and the cols_orig loads issued
// at the hits) complete under that scan instead of stalling the wave.
template <bool kActive>
__global__ __launch_bounds__(kBlock) void k_pull_thread(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint64_t *__restrict__ head, const uint64_t *__restrict__ head_orig,
    const uint32_t *__restrict__ cols_orig, uint64_t *__restrict__ visited,
    const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint8_t *__restrict__ level, const uint32_t *__restrict__ new2old, int64_t *__restrict__ cnt_all, int bin,
    uint64_t *__restrict__ act_next, uint64_t *__restrict__ lvflag)
{
    __shared__ uint64_t s_live[kBlock / 64];
    const int lane = __lane_id();
    const uint64_t lanes = (uint64_t)cnt_all[kThreadLanes];
    const uint64_t used = (uint64_t)cnt_all[kUsed];
    const int64_t flags_from = cnt_all[kFlagWord];
    const PullOut out{parent, level, new2old, cols_orig, level_byte(cnt_all)};
    int64_t ngath = 0, nbits = 0, nedges = 0, nverts = 0, nnext = 0, enext = 0;
    uint64_t live = 0;
    KeptHits pend;  // the previous step's hits, translated, not yet written
    // wave-uniform loop: the wave takes 64 consecutive vertices per step, one
    // word of the next frontier's active bitmap (the bin starts at a multiple
    // of 64)
    for (int64_t base = blockIdx.x * (int64_t)kBlock + (threadIdx.x & ~63); base < nlist;
         base += (int64_t)gridDim.x * kBlock) {
        const int64_t i = base + lane;
        const uint32_t v = (uint32_t)(first_v + i);
        const int64_t fword = (first_v + base) >> 6;
        const uint64_t fw = fword >= flags_from ? lvflag[fword] : 0ull;
        KeptHits cur;
        cur.lvmode = (fw >> lane & 1) ? kLvFlagged : 0u;
        uint32_t from_head = 0;
        if (i < nlist) {
            const uint64_t vis = visited[v];
            const uint64_t unv = ~vis & lanes;
            if (unv) {
                uint64_t found = 0;
                const int64_t b = rowptr[v], e = rowptr[v + 1];
                int64_t j = b;
                if (j < e)
                    j = thread_scan_step<kThreadFirst, kActive, true>(j, e, v, unv, found, cur.nh, cur.hu, cur.hf,
                                                                      from_head, ngath, cols, head, frontier, active,
                                                                      out);
                while (j < e && found != unv)
                    j = thread_scan_step<kThreadUnroll, kActive, false>(j, e, v, unv, found, cur.nh, cur.hu, cur.hf,
                                                                        from_head, ngath, cols, head, frontier,
                                                                        active, out);
                j = j < e ? j : e;
                if (found) {
                    visited[v] = vis | found;
                    ++nnext;
                    enext += e - b;
                }
                nedges += j - b;
                ++nverts;
                nbits += __popcll(found);
                cur.found = found;
                // the vertex's lanes so far all came at the level before (or
                // this is its first): its whole level record can be stored,
                // with the next level predicted for the lanes still open
                if (found && !cur.lvmode) {
                    const uint64_t own = frontier[v];
                    if ((vis & used & ~own) == 0) {
                        cur.lvmode = kLvWhole;
                        cur.own = own;
                    }
                }
            }
            next[v] = cur.found;
        }
        live |= cur.found;
        {
            const uint64_t nf = __ballot(cur.lvmode == kLvWhole);
            if (lane == 0) lvflag[fword] = nf;
        }
        const uint64_t word = __ballot(cur.found != 0);
        if (lane == 0) act_next[(first_v + base) >> 6] = word;
        translate(cur, v, from_head, new2old, head_orig);
        write_out(pend, lane, parent, level, out.lv);
        pend = cur;
    }
    write_out(pend, lane, parent, level, out.lv);
    live = g5_wave_or(live);
    if (lane == 0) s_live[threadIdx.x >> 6] = live;
    block_add(cnt_all + kUnitSlots * bin, {nbits, nedges, nverts, ngath, nnext, enext});  // its barrier orders s_live
    if (threadIdx.x == 0) {
        uint64_t t = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t |= s_live[w];
        if (t)
            atomicOr((unsigned long long *)&cnt_all[(blockIdx.x % kStripes) * kCntStride + kLive],
                     (unsigned long long)t);
    }
}

// One wavefront per vertex (thread_max < degree <= wave_max): a step loads
// 64 neighbours at once; the wave's exclusive prefix-OR gives each root's first hit, the
// inclusive OR the new bits.  First-hit lanes drop the parent's original id
// (cols_orig at their row position) into the
// wave's 64-slot LDS row (slot r = root r); when the vertex is done lane r
// stores slot r, so the vertex's parents leave in one contiguous store.
// Each iteration loads kWaveUnroll steps (64 neighbours each) and their
// frontier words before processing the steps in order.  The first step takes
// only kWaveFirst neighbours (the other lanes idle): on dense levels most
// vertices are done after a few, and every gather past the early exit is a
// wasted line fill.  On split levels every row is scanned to its end, so
// there the first step is a whole one (cnt[kWaveStep]).
constexpr int kWaveUnroll = 1;
constexpr int kWaveFirst = 16;  // 8: same time
static_assert(kWaveFirst <= 64, "a step spans at most one neighbour per lane");

template <bool kActive>
__global__ __launch_bounds__(kBlock) void k_pull_wave(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint64_t *__restrict__ head, const uint64_t *__restrict__ head_orig,
    const uint32_t *__restrict__ cols_orig, uint64_t *__restrict__ visited,
    const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint8_t *__restrict__ level, const uint32_t *__restrict__ new2old, int64_t *__restrict__ cnt_all, int bin,
    uint64_t *__restrict__ act_next, uint64_t *__restrict__ lvflag)
{
    __shared__ uint32_t s_par[kBlock / 64][64];
    uint32_t *__restrict__ sp = s_par[threadIdx.x >> 6];
    const uint8_t lv = level_byte(cnt_all);
    const int lane = __lane_id();
    const uint64_t lanes = (uint64_t)cnt_all[kPullLanes];
    const int64_t first_step = cnt_all[kWaveStep];
    const int64_t stride = (int64_t)gridDim.x * (kBlock / 64);
    int64_t nbits = 0, nedges = 0, nverts = 0, ngath = 0, nnext = 0, enext = 0;
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < nlist; i += stride) {
        const uint32_t v = (uint32_t)(first_v + i);
        const uint64_t vis = visited[v];
        const uint64_t unv = ~vis & lanes;
        if (!unv) {
            if (lane == 0) next[v] = 0;
            continue;
        }
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t base = b;
        int64_t step = first_step;
        while (base < e && found != unv) {
            const int64_t lim = min(e, base + step * kWaveUnroll);
            uint32_t u[kWaveUnroll];
            uint64_t fr[kWaveUnroll];
            const int64_t j0 = base + lane, jstep = step;
#pragma unroll
            for (int k = 0; k < kWaveUnroll; ++k) {
                const int64_t j = base + step * k + lane;
                u[k] = j < lim ? cols[j] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kWaveUnroll; ++k) {
                const int64_t j = base + step * k + lane;
                const bool take = j < lim && (!kActive || (active[u[k] >> 6] >> (u[k] & 63) & 1));
                fr[k] = take ? frontier[u[k]] : 0ull;
                ngath += take;
            }
            base += step * kWaveUnroll;
            step = 64;
#pragma unroll
            for (int k = 0; k < kWaveUnroll; ++k) {
                const uint64_t fresh = fr[k] & unv & ~found;
                if (__ballot(fresh != 0) == 0) continue;
                const uint64_t prev = lanes_excl_or(fresh);
                uint64_t first = fresh & ~prev;
                if (first) {
                    const uint32_t uo = cols_orig[j0 + jstep * k];
                    while (first) {
                        sp[__ffsll((long long)first) - 1] = uo;
                        first &= first - 1;
                    }
                }
                found |= __shfl(prev | fresh, 63, 64);
            }
        }
        if (found) {
            // the row's LDS writes come from other lanes of this wave; LDS
            // operations of one wave complete in issue order
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the slots hold original ids; lane r stores root r's new slot
            if (found >> lane & 1) {
                const int64_t slot = (int64_t)new2old[v] * 64 + lane;
                __builtin_nontemporal_store((int32_t)sp[lane], &parent[slot]);
                __builtin_nontemporal_store(lv, &level[slot]);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            next[v] = found;
            if (found) {
                visited[v] = vis | found;
                ++nnext;
                enext += e - b;
            }
            nbits += __popcll(found);
            nedges += min(base, e) - b;
            ++nverts;
        }
    }
    block_add(cnt_all + kUnitSlots * bin, {nbits, nedges, nverts, ngath, nnext, enext});
}

// One workgroup per vertex (degree > wave_max): kBlockUnroll sub-rounds of
// 256 neighbours per pair of barriers.
constexpr int kBlockUnroll = 1;

template <bool kActive>
__global__ __launch_bounds__(kBlock) void k_pull_block(
    int64_t first_v, int64_t nlist, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const uint64_t *__restrict__ head, const uint64_t *__restrict__ head_orig,
    const uint32_t *__restrict__ cols_orig, uint64_t *__restrict__ visited,
    const uint64_t *__restrict__ frontier,
    const uint64_t *__restrict__ active, uint64_t *__restrict__ next, int32_t *__restrict__ parent,
    uint8_t *__restrict__ level, const uint32_t *__restrict__ new2old, int64_t *__restrict__ cnt_all, int bin,
    uint64_t *__restrict__ act_next, uint64_t *__restrict__ lvflag)
{
    __shared__ uint64_t s_wave[kBlockUnroll][kBlock / 64];
    __shared__ uint64_t s_vis;
    __shared__ uint32_t s_par[64];
    const uint8_t lv = level_byte(cnt_all);
    int64_t ngath = 0;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    const uint64_t lanes = (uint64_t)cnt_all[kPullLanes];
    int64_t nbits = 0, nedges = 0, nverts = 0, nnext = 0, enext = 0;
    for (int64_t i = blockIdx.x; i < nlist; i += gridDim.x) {
        const uint32_t v = (uint32_t)(first_v + i);
        if (threadIdx.x == 0) s_vis = visited[v];
        __syncthreads();
        const uint64_t vis = s_vis;
        const uint64_t unv = ~vis & lanes;
        if (!unv) {
            if (threadIdx.x == 0) next[v] = 0;
            __syncthreads();
            continue;
        }
        uint64_t found = 0;
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int64_t base = b;
        // found is uniform across the workgroup: it changes only after the
        // barrier, from the shared per-wave ORs
        while (base < e && found != unv) {
            uint32_t u[kBlockUnroll];
            uint64_t f[kBlockUnroll], in_wave[kBlockUnroll];
#pragma unroll
            for (int k = 0; k < kBlockUnroll; ++k) {
                const int64_t j = base + kBlock * k + threadIdx.x;
                u[k] = j < e ? cols[j] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kBlockUnroll; ++k) {
                const int64_t j = base + kBlock * k + threadIdx.x;
                const bool take = j < e && (!kActive || (active[u[k] >> 6] >> (u[k] & 63) & 1));
                f[k] = take ? frontier[u[k]] & unv & ~found : 0ull;
                ngath += take;
            }
            base += kBlock * kBlockUnroll;
#pragma unroll
            for (int k = 0; k < kBlockUnroll; ++k) {
                const bool any = __ballot(f[k] != 0) != 0;
                in_wave[k] = any ? lanes_excl_or(f[k]) : 0ull;
                const uint64_t wave_all = any ? __shfl(in_wave[k] | f[k], 63, 64) : 0ull;
                if (lane == 0) s_wave[k][wv] = wave_all;
            }
            __syncthreads();
            // sub-round k's first hits exclude everything hit in earlier
            // sub-rounds and, within sub-round k, in earlier waves and lanes
            uint64_t earlier = 0;
#pragma unroll
            for (int k = 0; k < kBlockUnroll; ++k) {
                uint64_t before = earlier, all = 0;
#pragma unroll
                for (int w = 0; w < kBlock / 64; ++w) {
                    const uint64_t t = s_wave[k][w];
                    before |= (w < wv) ? t : 0ull;
                    all |= t;
                }
                uint64_t first = f[k] & ~(in_wave[k] | before);
                if (first) {
                    const uint32_t uo = cols_orig[base - kBlock * kBlockUnroll + kBlock * k + threadIdx.x];
                    while (first) {
                        s_par[__ffsll((long long)first) - 1] = uo;
                        first &= first - 1;
                    }
                }
                earlier |= all;
            }
            found |= earlier;
            __syncthreads();
        }
        // s_par is complete: the round's second barrier follows its writes
        if (threadIdx.x < 64 && (found >> threadIdx.x & 1)) {
            const int64_t slot = (int64_t)new2old[v] * 64 + threadIdx.x;
            __builtin_nontemporal_store((int32_t)s_par[threadIdx.x], &parent[slot]);
            __builtin_nontemporal_store(lv, &level[slot]);
        }
        if (threadIdx.x == 0) {
            next[v] = found;
            if (found) {
                visited[v] = vis | found;
                ++nnext;
                enext += e - b;
            }
            nbits += __popcll(found);
            nedges += min(base, e) - b;
            ++nverts;
        }
    }
    block_add(cnt_all + kUnitSlots * bin, {nbits, nedges, nverts, ngath, nnext, enext});
}

// ---- push (top-down) levels for sparse frontiers ----
// The frontier's vertices u (active bitmap -> sorted queue) push their lane
// bits along their edges: next[v] |= frontier[u] & ~visited[v].  The edge
// whose atomicOr sets a (v, root) bit first stores its source in the pair's
// parent slot; a second pass over the same edges lowers every newly reached
// pair's slot to the minimum frontier neighbour with atomicMin.  Work is balanced by edges: each block tile takes
// kPushTile consecutive edges of the frontier and finds each edge's source by
// binary search in the frontier's degree prefix.  The push kernels work on
// the lanes in lm (all lanes on push levels, the pushed lanes of a split
// level) and the targets below vend (all on push levels; on a split level
// the block and wave bins, whose kernels leave the pushed lanes out).
// a push level clears next entry by entry when the list is below n / kListedFrac
constexpr int64_t kListedFrac = 32;
constexpr int kPushTile = 1024;  // 20 KB of LDS per block (2048: 40 KB, 4 waves per SIMD)
constexpr int kPushGrid = 4096;
constexpr int kPushVertGrid = 2048;

__device__ __forceinline__ int64_t fq_src(const int64_t *__restrict__ qoff, int64_t lo, int64_t hi, int64_t e)
{
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (qoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// qoff[i] = degree of queue entry i; entries [queue length, bound) := 0, so
// that a scan over `bound` entries (a host-side upper bound of the queue's
// length) ends with the queue's edge total
__global__ void k_fq_degrees(const uint32_t *__restrict__ fq, const int64_t *__restrict__ cnt,
                             const int64_t *__restrict__ rowptr, int64_t *__restrict__ qoff, int64_t bound)
{
    const int64_t n = cnt[kFqLen];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < bound;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n) {
            const uint32_t u = fq[i];
            qoff[i] = rowptr[u + 1] - rowptr[u];
        } else {
            qoff[i] = 0;
        }
    }
}

// one pushed edge: u (relabelled id, uo its original id) -> v (vo).  mark (first
// pass only, when given): the edge that makes next[v] non-zero sets v's bit
// in mark, the next frontier's active bitmap.  The first pass writes the
// level bytes of the pairs it sets (every candidate parent of a pair is of
// the same level).
template <bool kMin>
__device__ __forceinline__ void push_edge(uint32_t u, uint32_t uo, uint32_t v, uint32_t vo,
                                          const uint64_t *__restrict__ visited,
                                          const uint64_t *__restrict__ frontier, uint64_t *__restrict__ next,
                                          int32_t *__restrict__ parent, uint8_t *__restrict__ level, uint8_t lv,
                                          uint64_t lm, uint64_t *__restrict__ mark)
{
    const uint64_t f = frontier[u] & lm & ~visited[v];
    if (!kMin) {
        // (a plain load of next[v] first, to skip atomics whose bits are set,
        // cost more than it saved: level 2 at scale 26 1.95 -> 1.7 ms without)
        if (!f) return;
        const uint64_t old = (uint64_t)atomicOr((unsigned long long *)&next[v], (unsigned long long)f);
        const uint64_t fresh = f & ~old;
        // the edge that sets a (v, root) bit first stores its source as the
        // pair's first parent candidate; the kMin pass lowers it
        if (fresh) write_slots(parent, level, vo, fresh, uo, lv);
        if (mark && old == 0) atomicOr((unsigned long long *)&mark[v >> 6], 1ull << (v & 63));
    } else if (f) {
        uint64_t bits = f;  // every bit of f is new at this level (visited excludes older ones)
        int32_t *rec = &parent[(int64_t)vo * 64];
        while (bits) {
            const int r = __ffsll((long long)bits) - 1;
            int32_t *slot = &rec[r];
            if (*slot > (int32_t)uo) atomicMin(slot, (int32_t)uo);
            bits &= bits - 1;
        }
    }
}

// The candidate pass of a split level (its targets, the block and wave
// bins, are few): every pushed edge lowers the candidate of each of its new
// (v, root) pairs, cand[v * npm + the lane's rank in lm], to its source's
// original id with atomicMin.  The candidates start at ~0u, so one pass
// gives the minimum-id parent; k_cand_commit then writes the pairs out.
__device__ __forceinline__ void push_cand(uint32_t u, uint32_t uo, uint32_t v, const uint64_t *__restrict__ visited,
                                          const uint64_t *__restrict__ frontier, uint32_t *__restrict__ cand,
                                          int npm, uint64_t lm)
{
    uint64_t bits = frontier[u] & lm & ~visited[v];
    while (bits) {
        const int r = __ffsll((long long)bits) - 1;
        bits &= bits - 1;
        uint32_t *slot = &cand[(int64_t)v * npm + __popcll(lm & ((1ull << r) - 1))];
        if (*slot > uo) atomicMin(slot, uo);
    }
}

// One thread per vertex below vend: the pushed lanes that found a candidate
// parent become the vertex's new pairs -- parent and level written at its
// original index, the bits ORed into next and visited (the pull kernels of
// the level, which own these words for the other lanes, have finished).
__global__ __launch_bounds__(kBlock) void k_cand_commit(const uint32_t *__restrict__ cand, int npm, uint64_t lm,
                                                        int64_t vend, const uint32_t *__restrict__ new2old,
                                                        uint64_t *__restrict__ next, uint64_t *__restrict__ visited,
                                                        int32_t *__restrict__ parent, uint8_t *__restrict__ level,
                                                        int64_t *__restrict__ cnt)
{
    const uint8_t lv = level_byte(cnt);
    int64_t nbits = 0, nv = 0;
    for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < vend; v += (int64_t)gridDim.x * kBlock) {
        const uint32_t *c = cand + v * npm;
        uint64_t bits = 0, b = lm;
        int64_t o = 0;
        for (int i = 0; b; ++i) {
            const int r = __ffsll((long long)b) - 1;
            b &= b - 1;
            const uint32_t x = c[i];
            if (x == ~0u) continue;
            if (!bits) o = (int64_t)new2old[v] * 64;
            bits |= 1ull << r;
            parent[o + r] = (int32_t)x;
            level[o + r] = lv;
        }
        if (bits) {
            next[v] |= bits;
            visited[v] |= bits;
            nbits += __popcll(bits);
            ++nv;
        }
    }
    block_add(cnt + kUnitSlots * kPushUnit, {nbits, (int64_t)0, nv});
}

// the push unit's next-frontier edges so far, summed over the stripes
__device__ __forceinline__ int64_t push_next_edges(const int64_t *__restrict__ cnt)
{
    int64_t t = 0;
    for (int st = 0; st < kStripes; ++st) t += cnt[st * kCntStride + kUnitSlots * kPushUnit + kNextEdges];
    return t;
}

// A tile's sources (at most kPushTile: every queue vertex has an edge) are
// staged in LDS first -- their first edge's offset within the tile, their
// relabelled and original ids and their row start less their queue offset --
// so that each edge finds its source by a binary search in LDS rather than
// in the global prefix.
template <bool kMin, bool kCand = false>
__global__ __launch_bounds__(kBlock) void k_push_edges(
    const uint32_t *__restrict__ fq, const int64_t *__restrict__ qoff, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, const uint32_t *__restrict__ cols_orig, const uint32_t *__restrict__ new2old,
    const uint64_t *__restrict__ visited, const uint64_t *__restrict__ frontier, uint64_t *__restrict__ next,
    int32_t *__restrict__ parent, uint8_t *__restrict__ level, int64_t *__restrict__ cnt, uint64_t lm,
    uint64_t *__restrict__ mark, int64_t vend, uint32_t *__restrict__ cand = nullptr, int npm = 0,
    bool rows_ok = false)
{
    __shared__ int64_t s_lo, s_hi;
    // the minimum pass of a pure push level whose new vertices have fewer
    // edges than its frontier runs by rows (k_min_rows)
    if (kMin && rows_ok && push_next_edges(cnt) < cnt[kFqEdges]) return;
    __shared__ int32_t s_off[kPushTile];
    __shared__ uint32_t s_u[kPushTile], s_uo[kPushTile];
    __shared__ int64_t s_rb[kPushTile];
    const int64_t qn = cnt[kFqLen], E = cnt[kFqEdges];
    const uint8_t lv = level_byte(cnt);
    for (int64_t t0 = blockIdx.x * (int64_t)kPushTile; t0 < E; t0 += (int64_t)gridDim.x * kPushTile) {
        const int64_t t1 = min(E, t0 + (int64_t)kPushTile);
        if (threadIdx.x == 0) {
            s_lo = fq_src(qoff, 0, qn - 1, t0);
            s_hi = fq_src(qoff, s_lo, qn - 1, t1 - 1);
        }
        __syncthreads();
        const int64_t lo = s_lo, hi = s_hi;
        const int ns = (int)min(hi - lo + 1, (int64_t)kPushTile + 1);
        if (ns <= kPushTile) {
            for (int j = threadIdx.x; j < ns; j += kBlock) {
                const uint32_t u = fq[lo + j];
                const int64_t q = qoff[lo + j];
                s_u[j] = u;
                s_uo[j] = new2old[u];
                s_rb[j] = rowptr[u] - q;
                s_off[j] = (int32_t)max(q - t0, (int64_t)0);
            }
            __syncthreads();
            int j = 0;  // the sources of a thread's edges ascend
            for (int64_t e = t0 + threadIdx.x; e < t1; e += kBlock) {
                const int32_t rel = (int32_t)(e - t0);
                int a = j, b = ns - 1;  // last j with s_off[j] <= rel
                while (a < b) {
                    const int mid = (a + b + 1) >> 1;
                    if (s_off[mid] <= rel) a = mid;
                    else b = mid - 1;
                }
                j = a;
                const int64_t pos = s_rb[j] + e;
                const uint32_t v = cols[pos];
                if (v >= vend) continue;
                if (kCand) push_cand(s_u[j], s_uo[j], v, visited, frontier, cand, npm, lm);
                else
                    push_edge<kMin>(s_u[j], s_uo[j], v, cols_orig[pos], visited, frontier, next, parent, level, lv,
                                    lm, mark);
            }
        } else {
            for (int64_t e = t0 + threadIdx.x; e < t1; e += kBlock) {
                const int64_t si = fq_src(qoff, lo, hi, e);
                const uint32_t u = fq[si];
                const int64_t pos = rowptr[u] + (e - qoff[si]);
                const uint32_t v = cols[pos];
                if (v >= vend) continue;
                if (kCand) push_cand(u, new2old[u], v, visited, frontier, cand, npm, lm);
                else
                    push_edge<kMin>(u, new2old[u], v, cols_orig[pos], visited, frontier, next, parent, level, lv,
                                    lm, mark);
            }
        }
        __syncthreads();
    }
    if (!kMin && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd((unsigned long long *)&cnt[kUnitSlots * kPushUnit + kEdgesRead], (unsigned long long)E);
}

// next[list[i]] := 0 for i < len (the entries a pure push level's frontier
// queue listed: the only non-zero words of the buffer two levels on)
__global__ void k_zero_listed(uint64_t *__restrict__ next, const uint32_t *__restrict__ list, int64_t len)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        next[list[i]] = 0;
}

// The minimum pass of a pure push level by rows, when its newly reached
// vertices have fewer edges than its frontier (levels past the peak: few new
// vertices of low degree): each new vertex scans its row in ascending
// original-id order, and the first frontier neighbour of each new lane is
// its parent.  The same minimum as the edge pass (k_push_edges<true>), which
// runs instead otherwise; each of the two exits at once when not chosen.
__global__ __launch_bounds__(kBlock) void k_min_rows(const uint32_t *__restrict__ nq,
                                                     const int64_t *__restrict__ rowptr,
                                                     const uint32_t *__restrict__ cols,
                                                     const uint32_t *__restrict__ cols_orig,
                                                     const uint32_t *__restrict__ new2old,
                                                     const uint64_t *__restrict__ frontier,
                                                     const uint64_t *__restrict__ next,
                                                     int32_t *__restrict__ parent, const int64_t *__restrict__ cnt)
{
    if (push_next_edges(cnt) >= cnt[kFqEdges]) return;
    const int64_t n = cnt[kNq];
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const uint32_t v = nq[i];
        const uint64_t w = next[v];
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        int32_t *rec = &parent[(int64_t)new2old[v] * 64];
        uint64_t found = 0;
        for (int64_t j = b; j < e && found != w; j += 4) {
            uint64_t fr[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) fr[k] = j + k < e ? frontier[cols[j + k]] : 0ull;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint64_t f = fr[k] & w & ~found;
                if (!f) continue;
                found |= f;
                const int32_t uo = (int32_t)cols_orig[j + k];
                for (; f; f &= f - 1) rec[__ffsll((long long)f) - 1] = uo;
            }
        }
    }
}

// A pure push level's next frontier from its queue nq (cnt[kNq] entries)
// instead of a pass over all n words: the push unit's next-frontier vertices
// and edges, the live lanes and each lane's frontier edges (exact).
__global__ __launch_bounds__(kBlock) void k_nq_counts(const uint32_t *__restrict__ nq,
                                                      const uint64_t *__restrict__ next,
                                                      const int64_t *__restrict__ rowptr, int64_t *__restrict__ cnt,
                                                      int64_t *__restrict__ lane_edges)
{
    __shared__ unsigned long long s_lane[64];
    __shared__ uint64_t s_live[kBlock / 64];
    const int64_t n = cnt[kNq];
    if (threadIdx.x < 64) s_lane[threadIdx.x] = 0;
    __syncthreads();
    int64_t nv = 0, ne = 0;
    uint64_t live = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const uint32_t v = nq[i];
        const uint64_t w = next[v];
        const int64_t d = rowptr[v + 1] - rowptr[v];
        ++nv;
        ne += d;
        live |= w;
        for (uint64_t b = w; b; b &= b - 1) atomicAdd(&s_lane[__ffsll((long long)b) - 1], (unsigned long long)d);
    }
    live = g5_wave_or(live);
    if (__lane_id() == 0) s_live[threadIdx.x >> 6] = live;
    block_add(cnt + kUnitSlots * kPushUnit + kNextVerts, {nv, ne});  // its barriers order s_lane, s_live
    if (threadIdx.x == 0) {
        uint64_t t = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t |= s_live[w];
        if (t) atomicOr((unsigned long long *)&cnt[(blockIdx.x % kStripes) * kCntStride + kLive], (unsigned long long)t);
    }
    if (threadIdx.x < 64 && s_lane[threadIdx.x])
        atomicAdd((unsigned long long *)&lane_edges[(blockIdx.x % kLaneStripes) * 64 + threadIdx.x],
                  s_lane[threadIdx.x]);
}

// visited |= next over the newly reached queue, and the push unit's counts
// (lanes lm only)
__global__ __launch_bounds__(kBlock) void k_push_commit(const uint32_t *__restrict__ nq,
                                                        const uint64_t *__restrict__ next,
                                                        uint64_t *__restrict__ visited, int64_t *__restrict__ cnt,
                                                        uint64_t lm)
{
    const int64_t n = cnt[kNq];
    int64_t nbits = 0, nv = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const uint32_t v = nq[i];
        const uint64_t w = next[v] & lm;
        visited[v] |= w;
        nbits += __popcll(w);
        ++nv;
    }
    block_add(cnt + kUnitSlots * kPushUnit, {nbits, (int64_t)0, nv});
}

// one workgroup of kBlock threads: clear every stripe's accumulated slots
__global__ void k_level_begin(int64_t *cnt, int64_t *lane_edges)
{
    for (int i = threadIdx.x; i < kStripes * kCntStride; i += blockDim.x)
        if (i % kCntStride < kAccum) cnt[i] = 0;
    for (int i = threadIdx.x; i < kLaneStripes * 64; i += blockDim.x) lane_edges[i] = 0;
    if (threadIdx.x == 0) cnt[kNq] = 0;
}

// active[w] bit i = ((next[64w + i] & lm) != 0).  A wave takes four
// 64-vertex words per step (four independent loads in flight per lane).
// kCount: also add the active vertices and their edges to the push unit's
// next-frontier slots (pull levels count them in the pull kernels).
// kOrLive: OR the words into the level's live-lanes slot (kLive).
template <bool kCount, bool kOrLive>
__global__ __launch_bounds__(kBlock) void k_build_active(const uint64_t *__restrict__ next, int64_t nwords,
                                                         const int64_t *__restrict__ rowptr,
                                                         uint64_t *__restrict__ active, int64_t *__restrict__ cnt,
                                                         uint64_t lm)
{
    __shared__ uint64_t s_live[kBlock / 64];
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
    int64_t nv = 0, ne = 0;
    uint64_t live = 0;
    for (int64_t w0 = (blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6)) * 4; w0 < nwords; w0 += waves * 4) {
        uint64_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = w0 + k < nwords ? next[(w0 + k) * 64 + lane] & lm : 0ull;
        live |= x[0] | x[1] | x[2] | x[3];
        uint64_t mine = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t bits = __ballot(x[k] != 0);
            if (lane == k) mine = bits;
            if (kCount && x[k]) {
                const int64_t v = (w0 + k) * 64 + lane;
                ++nv;
                ne += rowptr[v + 1] - rowptr[v];
            }
        }
        if (lane < 4 && w0 + lane < nwords) active[w0 + lane] = mine;
    }
    if (kOrLive) {
        live = g5_wave_or(live);
        if (lane == 0) s_live[threadIdx.x >> 6] = live;
    }
    if (kCount) block_add(cnt + kUnitSlots * kPushUnit + kNextVerts, {nv, ne});  // its barrier orders s_live
    else __syncthreads();
    if (kOrLive && threadIdx.x == 0) {
        uint64_t t = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t |= s_live[w];
        if (t) atomicOr((unsigned long long *)&cnt[(blockIdx.x % kStripes) * kCntStride + kLive], (unsigned long long)t);
    }
}

// One thread: the lanes the next block and wave kernels work on, those the
// next thread kernel works on, the lanes in use, the first word of level
// flags that the level just before wrote (its thread kernel's bin; INT64_MAX
// after a push level): the thread kernel trusts the flags from there on, and
// the wave kernel's first step (kWaveFirst neighbours, or 64 on split
// levels, whose rows are scanned to the end).
__global__ void k_set_pull_lanes(int64_t *cnt, uint64_t lanes, uint64_t thread_lanes, uint64_t used,
                                 int64_t flag_word, int64_t wave_step)
{
    cnt[kWaveStep] = wave_step;
    cnt[kPullLanes] = (int64_t)lanes;
    cnt[kThreadLanes] = (int64_t)thread_lanes;
    cnt[kUsed] = (int64_t)used;
    cnt[kFlagWord] = flag_word;
}

// one workgroup: fold the stripes into stripe 0, then advance the level
// ... and the per-lane frontier edges, summed over their stripes, into
// slots [kCntStride, kCntStride + 64) (stripes 1 and 2, free once folded),
// so that one copy of kHostSlots slots brings the level's counters home
__global__ void k_level_end(int64_t *cnt, const int64_t *lane_edges)
{
    for (int i = threadIdx.x; i < kAccum; i += blockDim.x) {
        int64_t t = 0;
        if (i == kLive)
            for (int st = 0; st < kStripes; ++st) t |= cnt[st * kCntStride + i];
        else
            for (int st = 0; st < kStripes; ++st) t += cnt[st * kCntStride + i];
        cnt[i] = t;
    }
    if (threadIdx.x == 0) cnt[kDepth] += 1;
    __syncthreads();
    for (int r = threadIdx.x; r < 64; r += blockDim.x) {
        int64_t t = 0;
        for (int st = 0; st < kLaneStripes; ++st) t += lane_edges[st * 64 + r];
        cnt[kCntStride + r] = t;
    }
}

// visited := unused lanes, both active bitmaps := 0 and, unless the caller
// clears them entry by entry (fronts false), both frontier buffers := 0
__global__ void k_ms_init(uint64_t *__restrict__ visited, uint64_t *__restrict__ fa, uint64_t *__restrict__ fb,
                          uint64_t *__restrict__ act, int64_t n, uint64_t unused_lanes, bool fronts)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        visited[i] = unused_lanes;
        if (fronts) {
            fa[i] = 0;
            fb[i] = 0;
        }
        if (i < 2 * (n / 64)) act[i] = 0;
    }
}

// reach[r] = the edges of root r's neighbours (its second level's frontier
// edges); one workgroup per root
__global__ __launch_bounds__(256) void k_root_reach(const int64_t *__restrict__ roots, const uint32_t *__restrict__ old2new,
                                                    const int64_t *__restrict__ rowptr,
                                                    const uint32_t *__restrict__ cols, int64_t *__restrict__ reach)
{
    const int64_t v = old2new[roots[blockIdx.x]];
    int64_t t = 0;
    for (int64_t j = rowptr[v] + threadIdx.x; j < rowptr[v + 1]; j += blockDim.x) {
        const uint32_t u = cols[j];
        t += rowptr[u + 1] - rowptr[u];
    }
    t = g5_block_sum256(t);
    if (threadIdx.x == 0) reach[blockIdx.x] = t;
}

// One thread per root (roots may share a vertex, hence the atomics).  Root r
// takes lane lane_of[r] = its rank by descending reach (the edges of its
// neighbours; ties by r): roots of similar reach then sit in neighbouring
// lanes, so the lanes a vertex gains
// at one level tend to share the 64 B sectors of its parent record, and a
// pull level writes fewer partial sectors.  A root without neighbours is
// reached and done: it never enters the frontier, so no frontier word of a
// vertex outside the degree bins (which the pull kernels never rewrite) is
// ever non-zero.
__global__ void k_ms_roots(const int64_t *__restrict__ roots, int nroots, const uint32_t *__restrict__ old2new,
                           const int64_t *__restrict__ rowptr, uint64_t *__restrict__ visited,
                           uint64_t *__restrict__ frontier, uint64_t *__restrict__ active,
                           int32_t *__restrict__ parent, uint8_t *__restrict__ level, int64_t *__restrict__ cnt,
                           int32_t *__restrict__ lane_of, const int64_t *__restrict__ reach)
{
    __shared__ int64_t s_key[64];
    const int r = threadIdx.x;
    int64_t v = 0, d = -1, key = -1;
    if (r < nroots) {
        v = old2new[roots[r]];
        d = rowptr[v + 1] - rowptr[v];
        key = reach[r];
    }
    s_key[r] = key;
    __syncthreads();
    if (r < nroots) {
        int lane = 0;
        for (int j = 0; j < nroots; ++j) lane += (s_key[j] > key || (s_key[j] == key && j < r)) ? 1 : 0;
        lane_of[r] = lane;
        atomicOr((unsigned long long *)&visited[v], 1ull << lane);
        parent[roots[r] * 64 + lane] = (int32_t)roots[r];
        level[roots[r] * 64 + lane] = 0;
        if (d != 0) {
            atomicOr((unsigned long long *)&frontier[v], 1ull << lane);
            atomicOr((unsigned long long *)&active[v >> 6], 1ull << (v & 63));
        }
    }
    if (r == 0) cnt[kDepth] = 1;
}

// Degree-class keys: (2^32-1 - class) << 32 | id, so an ascending sort gives
// descending degree class, ties by id.  Class 0 holds the vertices without
// neighbours, class c >= 1 the degrees d with 2^(c-2) < d <= 2^(c-1) (1: d = 1,
// 2: d = 2, 3: d in 3..4, 4: d in 5..8, ...).  Within a class the vertices keep
// their original order, so the 64 vertices of a wave have ascending original
// ids and the parent and level records they write lie close together.
__global__ void k_degree_keys(const int64_t *__restrict__ rowptr, int64_t n, uint64_t *__restrict__ keys)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = rowptr[v + 1] - rowptr[v];
        const uint32_t cls = d == 0 ? 0u : 1u + (uint32_t)(64 - __clzll((long long)(d - 1)));
        keys[v] = ((uint64_t)(0xFFFFFFFFu - cls) << 32) | (uint64_t)v;
    }
}

// vertices with degree > wave_max, in (thread_max, wave_max], in (0, thread_max]
__global__ void k_bin_counts(const int64_t *__restrict__ rowptr, int64_t n, int64_t thread_max, int64_t wave_max,
                             unsigned long long *__restrict__ bins)
{
    unsigned long long c[kBins] = {0, 0, 0};
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = rowptr[v + 1] - rowptr[v];
        if (d > wave_max) ++c[0];
        else if (d > thread_max) ++c[1];
        else if (d > 0) ++c[2];
    }
    for (int k = 0; k < kBins; ++k) {
        const unsigned long long t = (unsigned long long)g5_block_sum256((int64_t)c[k]);
        if (threadIdx.x == 0 && t) atomicAdd(&bins[k], t);
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
// original ids) mapped to new ids, and as it is into ncols_orig
__global__ __launch_bounds__(256) void k_perm_rows(const int64_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ cols,
                                                   const uint32_t *__restrict__ new2old,
                                                   const uint32_t *__restrict__ old2new, int64_t n,
                                                   const int64_t *__restrict__ nrowptr, uint32_t *__restrict__ ncols,
                                                   uint32_t *__restrict__ ncols_orig)
{
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t o = new2old[i];
        const int64_t b = rowptr[o], d = rowptr[o + 1] - b, dst = nrowptr[i];
        for (int64_t j = __lane_id(); j < d; j += 64) {
            const uint32_t c = cols[b + j];
            ncols[dst + j] = old2new[c];
            ncols_orig[dst + j] = c;
        }
    }
}

// head[v] = the first two neighbours of relabelled row v (0 past its end),
// head_orig[v] their original ids
__global__ void k_row_heads(const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
                            const uint32_t *__restrict__ cols_orig, int64_t n, uint64_t *__restrict__ head,
                            uint64_t *__restrict__ head_orig)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = rowptr[v], d = rowptr[v + 1] - b;
        const uint64_t lo = d > 0 ? cols[b] : 0u, hi = d > 1 ? cols[b + 1] : 0u;
        head[v] = lo | hi << 32;
        const uint64_t olo = d > 0 ? cols_orig[b] : 0u, ohi = d > 1 ? cols_orig[b + 1] : 0u;
        head_orig[v] = olo | ohi << 32;
    }
}

// Records of edgeless vertices that held a root (original ids; roots whose
// relabelled id is below bin_end have edges and are left alone): every slot
// back to "unreached".  One wavefront per listed vertex, lane = slot.
__global__ __launch_bounds__(256) void k_ms_clear_roots(const int64_t *__restrict__ list, int64_t len,
                                                        const uint32_t *__restrict__ old2new, int64_t bin_end,
                                                        int32_t *__restrict__ parent, uint8_t *__restrict__ level)
{
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t i = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6); i < len; i += waves) {
        const int64_t o = list[i];
        if ((int64_t)old2new[o] < bin_end) continue;
        parent[o * 64 + __lane_id()] = -1;
        level[o * 64 + __lane_id()] = 0xFF;
    }
}

// The same for every edgeless vertex (relabelled ids [from, to)).
__global__ __launch_bounds__(256) void k_ms_clear_range(const uint32_t *__restrict__ new2old, int64_t from,
                                                        int64_t to, int32_t *__restrict__ parent,
                                                        uint8_t *__restrict__ level)
{
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t v = from + blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6); v < to; v += waves) {
        const int64_t o = new2old[v];
        parent[o * 64 + __lane_id()] = -1;
        level[o * 64 + __lane_id()] = 0xFF;
    }
}

// The last pass of a batch run: slots of the used lanes that the search did
// not reach := unreached (-1, 255), for every vertex with edges (relabelled
// ids [0, bin_end)); edgeless vertices' records already read unreached.  A
// wave loads 64 vertices' visited words (lane = vertex), then writes the
// records of those with an unreached lane, one vertex per step (lane = slot).
__global__ __launch_bounds__(256) void k_ms_finalize(const uint64_t *__restrict__ visited,
                                                     const uint32_t *__restrict__ new2old, int64_t bin_end,
                                                     uint64_t used, int32_t *__restrict__ parent,
                                                     uint8_t *__restrict__ level)
{
    const int lane = __lane_id();
    for (int64_t base = blockIdx.x * (int64_t)kBlock + (threadIdx.x & ~63); base < bin_end;
         base += (int64_t)gridDim.x * kBlock) {
        const int64_t v = base + lane;
        const uint64_t miss = v < bin_end ? used & ~visited[v] : 0ull;
        const uint32_t o = miss ? new2old[v] : 0u;
        uint64_t todo = __ballot(miss != 0);
        while (todo) {
            const int src = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            if (lane_read64(miss, src) >> lane & 1) {
                const int64_t slot = (int64_t)lane_read(o, src) * 64 + lane;
                parent[slot] = -1;
                level[slot] = 0xFF;
            }
        }
    }
}

// Lane `lane`'s maps as the single-root search returns them, by original id:
// pred int64[n] (-1 unreached) and, when given, level int32[n] (-1
// unreached).  root < 0: the records were finalised by a batch run (they
// are the answer as they stand); otherwise (continuous mode, root = the
// lane's root) a slot counts only where the lane's visited bit is set, and
// levels are taken relative to the root's level byte (the engine's level
// counter when the root joined).
__global__ void k_ms_extract(const uint32_t *__restrict__ old2new, const int32_t *__restrict__ parent,
                             const uint8_t *__restrict__ level, const uint64_t *__restrict__ visited, int64_t n,
                             int lane, int64_t root, int64_t *__restrict__ pred, int32_t *__restrict__ lev)
{
    const bool final_maps = root < 0;
    const uint8_t root_byte = final_maps ? 0 : level[root * 64 + lane];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const bool reached = final_maps ? parent[i * 64 + lane] >= 0 : (visited[old2new[i]] >> lane & 1) != 0;
        pred[i] = reached ? (int64_t)parent[i * 64 + lane] : -1;
        if (lev) lev[i] = reached ? (int32_t)(uint8_t)(level[i * 64 + lane] - root_byte) : -1;
    }
}

// Levels from a parent tree by pointer jumping.  Each vertex holds one
// 64-bit word (distance << 32 | ancestor): the ancestor is `distance` tree
// edges above it, and the word is read and written whole, so a pass that
// mixes old and new words still sees consistent pairs.  A pass replaces
// (d, a) by (d + d_a, a_a) until every ancestor is the root (d_root = 0).
__global__ void k_jump_init(const int64_t *__restrict__ pred, int64_t n, uint64_t *__restrict__ w)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = pred[i];
        w[i] = p < 0 ? ~0ull : p == i ? (uint64_t)i : (1ull << 32) | (uint64_t)p;
    }
}

// *flags |= 1 when a word changed, 2 when a vertex's ancestor is unreached
// (a broken tree: its levels cannot be derived)
__global__ void k_jump(uint64_t *__restrict__ w, int64_t n, int *__restrict__ flags)
{
    int ch = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (x == ~0ull) continue;
        const uint32_t a = (uint32_t)x;
        const uint64_t y = __hip_atomic_load(&w[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == ~0ull) {
            ch |= 2;
            continue;
        }
        if ((y >> 32) == 0) continue;  // a is the root
        __hip_atomic_store(&w[i], ((x >> 32) + (y >> 32)) << 32 | (uint32_t)y, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        ch |= 1;
    }
    ch = (int)g5_wave_or32((uint32_t)ch);
    if (ch && __lane_id() == 0) atomicOr(flags, ch);
}

__global__ void k_jump_levels(const uint64_t *__restrict__ w, int64_t n, int32_t *__restrict__ lev)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = w[i];
        lev[i] = x == ~0ull ? -1 : (int32_t)(x >> 32);
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
    // the workgroup's 4 waves fold in LDS, then one atomic per lane slot
    __shared__ unsigned long long s_acc[4][64];
    s_acc[threadIdx.x >> 6][lane] = acc;
    __syncthreads();
    if (threadIdx.x < 64) {
        const unsigned long long t = s_acc[0][lane] + s_acc[1][lane] + s_acc[2][lane] + s_acc[3][lane];
        if (t) atomicAdd(&out[lane], t);
    }
}

// ---- continuous mode: lanes join and leave between levels ----

// visited[v] = (visited[v] | set) & ~clear
__global__ void k_lane_masks(uint64_t *__restrict__ visited, int64_t n, uint64_t set, uint64_t clear)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        visited[i] = (visited[i] | set) & ~clear;
}

// One thread: root jr[i] (original id) starts lane jr[64 + i] in the frontier
// the next level reads.  A root without neighbours is only marked visited:
// its search is complete (and vertices outside the degree bins are never
// rewritten by the pull kernels, so no frontier bit may be left there).
// added[0..1] = new frontier vertices and their edges.
__global__ void k_join(const int64_t *__restrict__ jr, int n, const uint32_t *__restrict__ old2new,
                       const int64_t *__restrict__ rowptr, uint64_t *__restrict__ visited,
                       uint64_t *__restrict__ front, uint64_t *__restrict__ act, int32_t *__restrict__ parent,
                       uint8_t *__restrict__ level, const int64_t *__restrict__ cnt, int64_t *__restrict__ added)
{
    // the next level finds the root's neighbours with byte cnt[kDepth]
    const uint8_t root_byte = (uint8_t)(cnt[kDepth] - 1);
    int64_t nv = 0, ne = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t root = jr[i];
        const int lane = (int)jr[64 + i];
        const int64_t v = old2new[root];
        const uint64_t bit = 1ull << lane;
        const int64_t deg = rowptr[v + 1] - rowptr[v];
        visited[v] |= bit;
        parent[root * 64 + lane] = (int32_t)root;
        level[root * 64 + lane] = root_byte;
        added[2 + i] = deg;
        if (deg == 0) continue;
        if (front[v] == 0) {
            ++nv;
            ne += deg;
        }
        front[v] |= bit;
        act[v >> 6] |= 1ull << (v & 63);
    }
    added[0] = nv;
    added[1] = ne;
}

// Lanes in `mask` leave: per lane the sum of generated degrees over its
// reached vertices (2 x nedge) and the checksum sum over reached v of
// mix64(orig(v) << 32 | (pred(v) + 1)) mod 2^64, in original ids; then the
// lanes are freed (visited set everywhere).  A wave takes 64 consecutive
// vertices (lane = vertex: visited words, ids and degrees in coalesced
// loads), then reads the parent records of the reached ones (by original
// id) 8 vertices at a time with lane = root, all 8 loads issued before any
// is used.
constexpr int kLeaveGroup = 8;

__device__ __forceinline__ uint64_t bcast64(uint64_t x, int l)
{
    return ((uint64_t)lane_read((uint32_t)(x >> 32), l) << 32) | lane_read((uint32_t)x, l);
}

__global__ __launch_bounds__(256) void k_leave(uint64_t *__restrict__ visited, const int32_t *__restrict__ parent,
                                               const uint32_t *__restrict__ new2old, const uint32_t *__restrict__ gdeg,
                                               int64_t n, uint64_t mask, unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long s_ck[64], s_dg[64];
    if (threadIdx.x < 64) {
        s_ck[threadIdx.x] = 0;
        s_dg[threadIdx.x] = 0;
    }
    __syncthreads();
    const int lane = __lane_id();
    unsigned long long ck = 0, dg = 0;  // root `lane`'s sums
    for (int64_t base = blockIdx.x * (int64_t)kBlock + (threadIdx.x & ~63); base < n;
         base += (int64_t)gridDim.x * kBlock) {
        const int64_t v = base + lane;
        uint64_t reach = 0;
        uint32_t o = 0, d = 0;
        if (v < n) {
            const uint64_t vis = visited[v];
            reach = vis & mask;
            if (reach) {
                o = new2old[v];
                d = gdeg[v];
            }
            if (~vis & mask) visited[v] = vis | mask;
        }
        const uint64_t todo = __ballot(reach != 0);
        for (int j0 = 0; j0 < 64; j0 += kLeaveGroup) {
            if (((todo >> j0) & ((1ull << kLeaveGroup) - 1)) == 0) continue;
            uint64_t rj[kLeaveGroup];
            uint32_t oj[kLeaveGroup];
            int32_t pj[kLeaveGroup];
#pragma unroll
            for (int k = 0; k < kLeaveGroup; ++k) {
                rj[k] = bcast64(reach, j0 + k);
                oj[k] = lane_read(o, j0 + k);
                pj[k] = (rj[k] >> lane & 1) ? parent[(int64_t)oj[k] * 64 + lane] : 0;
            }
#pragma unroll
            for (int k = 0; k < kLeaveGroup; ++k) {
                const uint32_t dj = lane_read(d, j0 + k);
                if (rj[k] >> lane & 1) {
                    ck += g5_mix64(((uint64_t)oj[k] << 32) | (uint32_t)(pj[k] + 1));
                    dg += dj;
                }
            }
        }
    }
    if (mask >> lane & 1) {
        atomicAdd(&s_ck[lane], ck);
        atomicAdd(&s_dg[lane], dg);
    }
    __syncthreads();
    if (threadIdx.x < 64 && (mask >> threadIdx.x & 1)) {
        if (s_ck[threadIdx.x]) atomicAdd(&out[threadIdx.x], s_ck[threadIdx.x]);
        if (s_dg[threadIdx.x]) atomicAdd(&out[64 + threadIdx.x], s_dg[threadIdx.x]);
    }
}

// The three pull kernels of parity p (act: they consult the active bitmap;
// prof: events between them).  They work on the lanes in cnt[kPullLanes]
// (the thread kernel: cnt[kThreadLanes]).
// They share one argument list; only the thread kernel reads `head` and
// `head_orig`.
void pull_bins(g5_msbfs *m, hipStream_t s, int p, bool act_check, bool prof, const int64_t *bf)
{
    const Relabeled &R = m->rl;
    const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    const uint64_t *act = m->d_act + (p == 0 ? 0 : nw);
    uint64_t *act_next = m->d_act + (p == 0 ? nw : 0);
    for (int k = 0; k < kBins; ++k) {
        if (prof) (void)hipEventRecord(m->ek[k], s);
        const int64_t first = bf[k], cnt = bf[k + 1] - first;
        if (!cnt) continue;
#define G5_PULL_ARGS first, cnt, (const int64_t *)R.rowptr, (const uint32_t *)R.cols, (const uint64_t *)R.head, \
                     (const uint64_t *)R.head_orig, (const uint32_t *)R.cols_orig, \
                     m->d_visited, front, act, next, m->d_parent, m->d_level, (const uint32_t *)R.new2old, m->d_cnt, \
                     k, act_next, m->d_lvflag
        const unsigned grid = k == 0 ? (unsigned)std::min<int64_t>(cnt, 4096)
                              : k == 1 ? g5::grid_for(cnt, kBlock / 64, 16384)
                                       : g5::grid_for(cnt, kBlock, 16384);
        auto kern = k == 0 ? (act_check ? k_pull_block<true> : k_pull_block<false>)
                    : k == 1 ? (act_check ? k_pull_wave<true> : k_pull_wave<false>)
                             : (act_check ? k_pull_thread<true> : k_pull_thread<false>);
        hipLaunchKernelGGL(kern, grid, kBlock, 0, s, G5_PULL_ARGS);
#undef G5_PULL_ARGS
    }
    if (prof) (void)hipEventRecord(m->ek[kBins], s);
}

// One pull level of parity p (the launches a hipGraph captures).
int level_launches(g5_msbfs *m, hipStream_t s, int p, bool act_check, bool prof = false)
{
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    uint64_t *act_next = m->d_act + (p == 0 ? nw : 0);
    hipLaunchKernelGGL(k_level_begin, 1, kBlock, 0, s, m->d_cnt, m->d_lane);
    pull_bins(m, s, p, act_check, prof, m->rl.bin_first);
    // the thread kernel wrote its words of the active bitmap; the block and
    // wave bins' words (and the live lanes among them) come from next
    const int64_t nw01 = (m->rl.bin_first[2] + 63) / 64;
    if (nw01)
        hipLaunchKernelGGL((k_build_active<false, true>), g5::grid_for(nw01, 16, 4096), kBlock, 0, s,
                           (const uint64_t *)next, nw01, (const int64_t *)m->rl.rowptr, act_next, m->d_cnt, ~0ull);
    hipLaunchKernelGGL(k_level_end, 1, kBlock, 0, s, m->d_cnt, (const int64_t *)m->d_lane);
    return hipGetLastError() == hipSuccess ? G5_OK : G5_E_HIP;
}

// The push kernels for the frontier lanes lm of parity p, after next has been
// cleared (push level) or written by the pull kernels for the other lanes
// (split level).  `act_front` holds the frontier's active bitmap restricted
// to lm, fq_bound bounds its vertex count and fq receives its queue.  Ends
// with next's active bitmap, the next frontier's counts and each lane's
// frontier edges.  A push level (lm = every lane) finds its newly reached
// vertices as it goes -- the edge that makes next[v] non-zero sets v's
// active bit, in act_next cleared beforehand -- and counts them from their
// queue, so it makes no pass over all n words; a split level builds the
// bitmaps from next, and its newly reached vertices are those of the block
// and wave bins.  When those bins' vertices times the pushed lanes fit in n
// words, a split level takes the candidate path instead of the two passes:
// one pass lowers per-pair candidates (push_cand) and k_cand_commit writes
// the pairs out.
int push_lanes(g5_msbfs *m, hipStream_t s, int p, const uint64_t *act_front, int64_t fq_bound, uint64_t lm,
               uint32_t *fq)
{
    const Relabeled &R = m->rl;
    const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    uint64_t *act_next = m->d_act + (p == 0 ? nw : 0);
    const bool whole = lm == ~0ull;
    // a split level pushes only into the block and wave bins: the thread
    // kernel pulls every lane of its vertices
    const int64_t vend = whole ? m->n : R.split_first[2];
    const int64_t nw_push = whole ? nw : (vend + 63) / 64;
    G5_HIP(g5::bitmap_to_queue(act_front, nw, fq, m->d_cnt + kFqLen, m->d_scratch, s));
    hipLaunchKernelGGL(k_fq_degrees, g5::grid_for(fq_bound, 256, 8192), 256, 0, s, (const uint32_t *)fq,
                       (const int64_t *)m->d_cnt, (const int64_t *)R.rowptr, m->d_qoff, fq_bound);
    G5_HIP(g5::exclusive_scan_i64(m->d_qoff, fq_bound, m->d_cnt + kFqEdges, m->d_scratch, s));
    if (whole) G5_HIP(g5::fill_u64(act_next, nw, 0, s));
    // a split level with few pushed lanes: one candidate pass into the newly
    // reached queue's buffer (n words; idle on split levels) and a commit
    const int npm = __builtin_popcountll(lm);
    if (!whole && vend * npm <= m->n) {
        uint32_t *cand = m->d_nq;
        if (vend) {
            G5_HIP(g5::fill_u32(cand, vend * npm, ~0u, s));
            hipLaunchKernelGGL((k_push_edges<false, true>), kPushGrid, kBlock, 0, s, (const uint32_t *)fq,
                               (const int64_t *)m->d_qoff, (const int64_t *)R.rowptr, (const uint32_t *)R.cols,
                               (const uint32_t *)R.cols_orig, (const uint32_t *)R.new2old,
                               (const uint64_t *)m->d_visited, front, next, m->d_parent, m->d_level, m->d_cnt, lm,
                               (uint64_t *)nullptr, vend, cand, npm);
            hipLaunchKernelGGL(k_cand_commit, g5::grid_for(vend, kBlock, 4096), kBlock, 0, s, (const uint32_t *)cand,
                               npm, lm, vend, (const uint32_t *)R.new2old, next, m->d_visited, m->d_parent,
                               m->d_level, m->d_cnt);
        }
        if (nw_push)
            hipLaunchKernelGGL((k_build_active<true, true>), g5::grid_for(nw_push, 16, 4096), kBlock, 0, s,
                               (const uint64_t *)next, nw_push, (const int64_t *)R.rowptr, act_next, m->d_cnt, ~0ull);
        return hipGetLastError() == hipSuccess ? G5_OK : G5_E_HIP;
    }
    hipLaunchKernelGGL(k_push_edges<false>, kPushGrid, kBlock, 0, s, (const uint32_t *)fq,
                       (const int64_t *)m->d_qoff, (const int64_t *)R.rowptr, (const uint32_t *)R.cols,
                       (const uint32_t *)R.cols_orig, (const uint32_t *)R.new2old, (const uint64_t *)m->d_visited,
                       front, next, m->d_parent,
                       m->d_level, m->d_cnt, lm, whole ? act_next : (uint64_t *)nullptr, vend);
    // the newly reached vertices of the lanes lm
    if (!whole && nw_push)
        hipLaunchKernelGGL((k_build_active<false, false>), g5::grid_for(nw_push, 16, 4096), kBlock, 0, s,
                           (const uint64_t *)next, nw_push, (const int64_t *)R.rowptr, act_next, m->d_cnt, lm);
    G5_HIP(g5::bitmap_to_queue(act_next, nw_push, m->d_nq, m->d_cnt + kNq, m->d_scratch, s));
    if (whole)
        hipLaunchKernelGGL(k_nq_counts, g5::grid_for(m->n, kBlock, 2048), kBlock, 0, s, (const uint32_t *)m->d_nq,
                           (const uint64_t *)next, (const int64_t *)R.rowptr, m->d_cnt, m->d_lane);
    hipLaunchKernelGGL(k_push_edges<true>, kPushGrid, kBlock, 0, s, (const uint32_t *)fq,
                       (const int64_t *)m->d_qoff, (const int64_t *)R.rowptr, (const uint32_t *)R.cols,
                       (const uint32_t *)R.cols_orig, (const uint32_t *)R.new2old, (const uint64_t *)m->d_visited,
                       front, next, m->d_parent,
                       m->d_level, m->d_cnt, lm, (uint64_t *)nullptr, vend, (uint32_t *)nullptr, 0, whole);
    if (whole)
        hipLaunchKernelGGL(k_min_rows, kPushVertGrid, kBlock, 0, s, (const uint32_t *)m->d_nq,
                           (const int64_t *)R.rowptr, (const uint32_t *)R.cols, (const uint32_t *)R.cols_orig,
                           (const uint32_t *)R.new2old, front, (const uint64_t *)next, m->d_parent,
                           (const int64_t *)m->d_cnt);
    hipLaunchKernelGGL(k_push_commit, kPushVertGrid, kBlock, 0, s, (const uint32_t *)m->d_nq,
                       (const uint64_t *)next, m->d_visited, m->d_cnt, lm);
    // the block and wave bins' part of the next frontier: active bitmap,
    // counts, live lanes (the thread kernel has done its own bin's)
    if (!whole && nw_push)
        hipLaunchKernelGGL((k_build_active<true, true>), g5::grid_for(nw_push, 16, 4096), kBlock, 0, s,
                           (const uint64_t *)next, nw_push, (const int64_t *)R.rowptr, act_next, m->d_cnt, ~0ull);
    return hipGetLastError() == hipSuccess ? G5_OK : G5_E_HIP;
}

// One push level on parity p: direct launches (the frontier queue's length
// sizes the degree scan on the host).  fqn = active vertices of the frontier.
// next is cleared entry by entry when the level before was a push level too
// (its frontier queue lists the only non-zero words of this level's next:
// listed_in) and the list is short; otherwise by a fill of all n words.  The two push
// queue buffers alternate, so that list survives into this level.
int push_level(g5_msbfs *m, hipStream_t s, int p, int64_t fqn, bool listed_in, bool prof)
{
    uint64_t *next = p == 0 ? m->d_fb : m->d_fa;
    const int64_t nw = m->n / 64;
    const uint64_t *act = m->d_act + (p == 0 ? 0 : nw);
    hipLaunchKernelGGL(k_level_begin, 1, kBlock, 0, s, m->d_cnt, m->d_lane);
    if (listed_in && m->listed_len * kListedFrac < m->n) {
        if (m->listed_len)
            hipLaunchKernelGGL(k_zero_listed, g5::grid_for(m->listed_len, 256, 4096), 256, 0, s, next,
                               (const uint32_t *)m->listed, m->listed_len);
    } else {
        G5_HIP(g5::fill_u64(next, m->n, 0, s));
    }
    if (prof) G5_HIP(hipEventRecord(m->ek[0], s));
    uint32_t *fq = m->listed == m->d_fq ? m->d_fq2 : m->d_fq;
    int rc = push_lanes(m, s, p, act, fqn, ~0ull, fq);
    if (rc) return rc;
    m->listed = fq;
    if (prof) G5_HIP(hipEventRecord(m->ek[1], s));
    hipLaunchKernelGGL(k_level_end, 1, kBlock, 0, s, m->d_cnt, (const int64_t *)m->d_lane);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

// One split level on parity p: the block and wave kernels on the lanes in
// cnt[kPullLanes] and the thread kernel on every frontier lane, then the
// push kernels on the lanes pm (disjoint from cnt[kPullLanes]) into the
// block and wave bins only.  fqn = active vertices of the whole frontier.
int split_level(g5_msbfs *m, hipStream_t s, int p, uint64_t pm, bool act_check, int64_t fqn, bool prof)
{
    const uint64_t *front = p == 0 ? m->d_fa : m->d_fb;
    const int64_t nw = m->n / 64;
    uint64_t *act_cur = m->d_act + (p == 0 ? 0 : nw);
    hipLaunchKernelGGL(k_level_begin, 1, kBlock, 0, s, m->d_cnt, m->d_lane);
    pull_bins(m, s, p, act_check, prof, m->rl.split_first);
    if (prof) G5_HIP(hipEventRecord(m->ek_push[0], s));
    // the frontier vertices with a pushed lane, in the frontier's own active
    // bitmap (the pull kernels are done with it; the thread kernel has
    // written its words of act_next)
    hipLaunchKernelGGL((k_build_active<false, false>), g5::grid_for(nw, 16, 4096), kBlock, 0, s, front, nw,
                       (const int64_t *)m->rl.rowptr, act_cur, m->d_cnt, pm);
    int rc = push_lanes(m, s, p, act_cur, fqn, pm, m->listed == m->d_fq ? m->d_fq2 : m->d_fq);
    if (rc) return rc;
    if (prof) G5_HIP(hipEventRecord(m->ek_push[1], s));
    hipLaunchKernelGGL(k_level_end, 1, kBlock, 0, s, m->d_cnt, (const int64_t *)m->d_lane);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

int capture_graphs(g5_msbfs *m)
{
    for (int p = 0; p < 2; ++p)
        for (int a = 0; a < 2; ++a) {
            hipGraph_t gr = nullptr;
            G5_HIP(hipStreamBeginCapture(m->s, hipStreamCaptureModeThreadLocal));
            int rc = level_launches(m, m->s, p, a != 0);
            hipError_t e = hipStreamEndCapture(m->s, &gr);
            if (rc) return rc;
            G5_HIP(e);
            e = hipGraphInstantiate(&m->graph[p][a], gr, nullptr, nullptr, 0);
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
    int64_t *ndeg = nullptr;
    auto cleanup = [&]() {
        if (keys) (void)hipFree(keys);
        if (tmp) (void)hipFree(tmp);
        if (ndeg) (void)hipFree(ndeg);
    };
    hipError_t e = hipMalloc(&keys, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&tmp, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&ndeg, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&R.new2old, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.old2new, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.gdeg, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&R.rowptr, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&R.cols, (size_t)std::max<int64_t>(g->nnz, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&R.cols_orig, (size_t)std::max<int64_t>(g->nnz, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&R.head, (size_t)std::max<int64_t>(n, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&R.head_orig, (size_t)std::max<int64_t>(n, 1) * 8);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_degree_keys, g5::grid_for(n, 256, 16384), 256, 0, s, (const int64_t *)g->d_rowptr, n,
                           keys);
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
                           (const uint32_t *)R.old2new, n, (const int64_t *)R.rowptr, R.cols, R.cols_orig);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_row_heads, g5::grid_for(n, 256, 16384), 256, 0, s, (const int64_t *)R.rowptr,
                           (const uint32_t *)R.cols, (const uint32_t *)R.cols_orig, n, R.head, R.head_orig);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return g5::fail_hip(e, "msbfs relabel", __FILE__, __LINE__);
    return G5_OK;
}

// The degree bins of the relabelled order for the current Tune bounds
// (contiguous id ranges, degrees descending).
// bf: the bins' first vertices for the thread-bin bound thread_max
int bins_for(g5_msbfs *m, int64_t thread_max, int64_t *bf)
{
    hipStream_t s = m->s;
    const Relabeled &R = m->rl;
    unsigned long long *d = (unsigned long long *)m->d_sums;  // transient scratch, free between calls
    unsigned long long hb[kBins] = {0, 0, 0};
    G5_HIP(g5::fill_u64((uint64_t *)d, kBins, 0, s));
    hipLaunchKernelGGL(k_bin_counts, g5::grid_for(m->n, 256, 16384), 256, 0, s, (const int64_t *)R.rowptr, m->n,
                       thread_max, m->tune.wave_max, d);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpyAsync(hb, d, kBins * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    bf[0] = 0;
    for (int k = 0; k < kBins; ++k) bf[k + 1] = bf[k] + (int64_t)hb[k];
    // the thread bin starts at a multiple of 64 (its waves write whole words
    // of the active bitmap and of the level flags): the first bins' ends
    // round up, moving at most 63 vertices each into the bin before (any bin
    // gives the same results)
    const int64_t end = bf[kBins];
    for (int k = 1; k < kBins; ++k) bf[k] = std::min(end, std::max(bf[k - 1], (bf[k] + 63) / 64 * 64));
    return G5_OK;
}

// the bins of pull levels and those of split levels
int compute_bins(g5_msbfs *m)
{
    int rc = bins_for(m, m->tune.thread_max, m->rl.bin_first);
    if (!rc) rc = bins_for(m, std::min(m->tune.split_max, m->tune.thread_max), m->rl.split_first);
    return rc;
}

void destroy_graphs(g5_msbfs *m)
{
    for (auto &gp : m->graph)
        for (hipGraphExec_t &ge : gp)
            if (ge) {
                (void)hipGraphExecDestroy(ge);
                ge = nullptr;
            }
}

// the lanes in use: a batch run's roots' lanes, continuous mode's busy lanes
uint64_t used_lanes(const g5_msbfs *m)
{
    if (m->cont) return m->busy;
    return m->nroots == 64 ? ~0ull : (1ull << m->nroots) - 1;
}

// The lanes a pull level pushes instead: those of the frontier with fewer
// than nnz / lane_ratio frontier edges, fewest edges first, as many as keep
// their edges together few enough for a push level (nnz / push_ratio), and
// not all of the frontier's lanes.  (Taking all such lanes or none made a
// level whose small lanes add up past the bound a full pull: 43 ms against
// 30 ms split, at scale 27.)  Only right after a push level or a join:
// after a pull level every lane has had a dense level, and the unvisited
// vertices left are few (a split there has measured slower).  0: no split.
uint64_t split_lanes(const g5_msbfs *m)
{
    const int64_t nnz = m->g->nnz;
    if (m->mode != 0 || m->tune.lane_ratio <= 0 || !m->split_ok) return 0;
    const uint64_t cand = m->front_lanes & m->lane_known;  // lanes of unknown size are pulled
    int order[64];
    int nc = 0;
    for (int r = 0; r < 64; ++r)
        if ((cand >> r & 1) && m->lane_edges[r] * m->tune.lane_ratio < nnz) order[nc++] = r;
    std::sort(order, order + nc, [&](int a, int b) {
        return m->lane_edges[a] != m->lane_edges[b] ? m->lane_edges[a] < m->lane_edges[b] : a < b;
    });
    uint64_t pm = 0;
    int64_t pe = 0;
    for (int i = 0; i < nc; ++i) {
        const int64_t e = m->lane_edges[order[i]];
        if ((pe + e) * m->tune.push_ratio >= nnz) break;
        pm |= 1ull << order[i];
        pe += e;
    }
    return pm == m->front_lanes ? 0 : pm;
}

// One level at parity m->p: push or pull by the frontier's edges (a pull
// level may push some lanes, split_lanes), then the level's counters on the
// host; sets *new_bits (new (vertex, root) pairs) and *live (OR of the next
// frontier), and moves the frontier state on.
int run_level(g5_msbfs *m, int64_t *new_bits_out, uint64_t *live_out)
{
    hipStream_t s = m->s;
    const int p = m->p;
    // until this level completes, the buffers' state is not known
    const bool listed_in = m->listed_ok;
    m->listed_ok = false;
    m->last_empty = false;
    const int64_t act_verts = m->act_verts, act_edges = m->act_edges;
    const int64_t nnz = m->g->nnz;
    const bool push = m->mode == 2 || (m->mode == 0 && (act_edges < 0 || act_edges * m->tune.push_ratio < nnz));
    const bool act_check = act_edges < 0 || act_edges * m->tune.active_ratio < nnz;
    const uint64_t pm = push ? 0 : split_lanes(m);
    if (push) {
        int rc = push_level(m, s, p, act_verts, listed_in, m->prof);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(k_set_pull_lanes, 1, 1, 0, s, m->d_cnt, m->front_lanes & ~pm, m->front_lanes, used_lanes(m),
                           m->flag_word, pm ? (int64_t)64 : (int64_t)kWaveFirst);
        if (pm) {
            int rc = split_level(m, s, p, pm, act_check, act_verts, m->prof);
            if (rc) return rc;
        } else if (m->prof) {
            int rc = level_launches(m, s, p, act_check, true);
            if (rc) return rc;
        } else {
            G5_HIP(hipGraphLaunch(m->graph[p][act_check ? 1 : 0], s));
        }
    }
    G5_HIP(hipMemcpyAsync(m->h_cnt, m->d_cnt, kHostSlots * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    int64_t new_bits = 0;
    const int64_t *h = m->h_cnt;
    int64_t next_verts = 0, next_edges = 0;
    for (int c = 0; c < kUnits; ++c) {
        new_bits += h[kUnitSlots * c + kNewBits];
        // a split level counts the block and wave bins' next frontier in the
        // push unit (after the push), the thread bin's in its own
        if (pm && c != kPushUnit && c != 2) continue;
        next_verts += h[kUnitSlots * c + kNextVerts];
        next_edges += h[kUnitSlots * c + kNextEdges];
    }
    m->lane_known = push ? ~0ull : 0ull;
    if (push)
        for (int r = 0; r < 64; ++r) m->lane_edges[r] = h[kCntStride + r];
    m->listed_ok = push;
    m->listed_len = push ? h[kFqLen] : 0;
    m->last_empty = new_bits == 0;
    m->split_ok = push;
    if (pm) ++m->split_levels;
    if (m->prof) {
        double rec[kLevelRec] = {0};
        rec[8] = (double)new_bits;
        rec[9] = push ? 1.0 : pm ? 2.0 : 0.0;
        rec[10] = (double)act_verts;
        rec[11] = (double)act_edges;
        if (push || pm) {
            // 16 B row offsets per frontier vertex; per frontier edge and
            // pass (two passes) 4 B column + 8 B frontier word + 8 B visited
            // word + 8 B next word; 16 B visited read-modify-write per newly
            // reached vertex; 4 B first parent candidate + 4 B parent + 1 B
            // level per new (vertex, root) pair
            float kms = 0;
            if (pm) G5_HIP(hipEventElapsedTime(&kms, m->ek_push[0], m->ek_push[1]));
            else G5_HIP(hipEventElapsedTime(&kms, m->ek[0], m->ek[1]));
            const double E = (double)h[kFqEdges];
            m->k_ms[kPushUnit] += kms;
            m->k_launches[kPushUnit] += 1;
            m->k_bytes[kPushUnit] += 16.0 * h[kFqLen] + 56.0 * E + 16.0 * h[kUnitSlots * kPushUnit + kVisitedVerts] +
                                     9.0 * h[kUnitSlots * kPushUnit + kNewBits];
            rec[kPushUnit] = kms;
            rec[4 + kPushUnit] = E;
        }
        if (!push) {
            // per bin: 16 B row offsets + 8 B visited read + 8 B next write
            // per vertex with an unvisited lane; 4 B column (+ 1 bit of the
            // active bitmap when checked) per edge read; 8 B per frontier
            // word gathered; 4 B parent + 1 B level per new (vertex, root)
            // pair; per newly reached vertex its original id and at least
            // one parent's (4 + 4 B)
            for (int k = 0; k < kBins; ++k) {
                const int64_t *bf = pm ? m->rl.split_first : m->rl.bin_first;
                if (bf[k + 1] == bf[k]) continue;
                float kms = 0;
                G5_HIP(hipEventElapsedTime(&kms, m->ek[k], m->ek[k + 1]));
                m->k_ms[k] += kms;
                m->k_launches[k] += 1;
                m->k_bytes[k] += 32.0 * h[kUnitSlots * k + kVisitedVerts] +
                                 (act_check ? 4.125 : 4.0) * h[kUnitSlots * k + kEdgesRead] +
                                 8.0 * h[kUnitSlots * k + kGathers] + 5.0 * h[kUnitSlots * k + kNewBits] +
                                 8.0 * h[kUnitSlots * k + kNextVerts];
                rec[k] = kms;
                rec[4 + k] = (double)h[kUnitSlots * k + kEdgesRead];
            }
        }
        m->lvl.insert(m->lvl.end(), rec, rec + kLevelRec);
    }
    *new_bits_out = new_bits;
    *live_out = (uint64_t)h[kLive];
    m->front_lanes = (uint64_t)h[kLive];
    m->act_verts = next_verts;
    m->act_edges = next_edges;
    m->p ^= 1;
    // the level flags this level's thread kernel wrote (none: a push level,
    // or an empty thread bin)
    const int64_t *tf = pm ? m->rl.split_first : m->rl.bin_first;
    m->flag_word = push || tf[2] == tf[3] ? INT64_MAX : tf[2] / 64;
    return G5_OK;
}

}  // namespace

extern "C" int g5_msbfs_free(g5_msbfs *m);

extern "C" int g5_msbfs_create(g5_graph *g, g5_msbfs **out)
{
    G5_REQUIRE(g && out, G5_E_ARG, "g5_msbfs_create: null argument");
    G5_REQUIRE(g->n_rows == g->n_cols && g->row_start == g->col_start, G5_E_ARG,
               "g5_msbfs_create: needs a diagonal (1x1) block");
    G5_REQUIRE(g->n_rows <= INT32_MAX, G5_E_ARG, "g5_msbfs_create: %lld vertices, at most 2^31 - 1 (int32 parents)",
               (long long)g->n_rows);
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
    if (e == hipSuccess) e = hipMalloc(&m->d_lvflag, (size_t)std::max<int64_t>(n / 64, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_fq, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_fq2, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_nq, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_qoff, (size_t)(n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_parent, (size_t)n * 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&m->d_level, (size_t)n * 64);
    if (e == hipSuccess) e = hipMalloc(&m->d_dirty, (size_t)kDirtyCap * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_cnt, kStripes * kCntStride * 8);
    if (e == hipSuccess) e = hipHostMalloc(&m->h_cnt, kHostSlots * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&m->d_lane, kLaneStripes * 64 * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_scratch, (g5::scan_scratch_words(n + 1) + 16) * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_roots, 64 * 8 + 64 * 4);  // roots, then lane_of
    if (e == hipSuccess) e = hipMalloc(&m->d_join, 194 * 8);
    if (e == hipSuccess) e = hipMalloc(&m->d_sums, 128 * 8);
    if (e == hipSuccess) e = hipEventCreate(&m->ev0);
    if (e == hipSuccess) e = hipEventCreate(&m->ev1);
    if (e == hipSuccess) e = hipEventCreate(&m->ev_fin);
    for (int k = 0; k <= kBins && e == hipSuccess; ++k) e = hipEventCreate(&m->ek[k]);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&m->ek_push[k]);
    // every record starts unreached; a batch run keeps it so wherever it does
    // not reach
    if (e == hipSuccess) e = g5::fill_i32(m->d_parent, n * 64, -1, m->s);
    if (e == hipSuccess) e = g5::fill_u32((uint32_t *)m->d_level, n * 16, 0xFFFFFFFFu, m->s);
    if (e != hipSuccess) return fail(g5::fail_hip(e, "g5_msbfs_create allocation", __FILE__, __LINE__));
    int rc = build_relabel(m);
    if (rc) return fail(rc);
    rc = compute_bins(m);
    if (rc) return fail(rc);
    rc = capture_graphs(m);
    if (rc) return fail(rc);
    *out = m;
    return G5_OK;
}

extern "C" int g5_msbfs_footprint(const g5_graph *g, int64_t *graph_bytes, int64_t *fixed_bytes,
                                  int64_t *lane_bytes)
{
    G5_REQUIRE(g, G5_E_ARG, "g5_msbfs_footprint: null graph");
    const int64_t n = std::max<int64_t>(g->n_rows, 64), nnz = std::max<int64_t>(g->nnz, 1);
    // the records: parent int32 and level byte per (vertex, lane); one bit
    // per vertex and lane in visited and the two frontier buffers
    const int64_t per_lane = 5 * n + (3 * n + 7) / 8;
    const int64_t all = 3 * 8 * n                      // visited, frontier buffers
                        + 3 * (n / 64) * 8             // active bitmaps, level flags
                        + 3 * 4 * n + 8 * (n + 1)      // push queues, degree prefix
                        + 64 * 5 * n                   // parent and level records
                        + kDirtyCap * 8 + kStripes * kCntStride * 8 + kLaneStripes * 64 * 8 +
                        ((int64_t)g5::scan_scratch_words(n + 1) + 16) * 8 + 64 * 12 + 194 * 8 + 128 * 8 +
                        3 * 4 * n + 8 * (n + 1) + 2 * 4 * nnz  // relabelled copy: ids, degrees, rows
                        + 2 * 8 * n;                       // head, head_orig
    if (graph_bytes) *graph_bytes = 8 * (g->n_rows + 1) + 4 * g->nnz + 4 * g->n_rows;
    if (fixed_bytes) *fixed_bytes = all - 64 * per_lane;
    if (lane_bytes) *lane_bytes = per_lane;
    return G5_OK;
}

extern "C" int g5_msbfs_free(g5_msbfs *m)
{
    if (!m) return G5_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    for (auto &gp : m->graph)
        for (hipGraphExec_t ge : gp)
            if (ge) (void)hipGraphExecDestroy(ge);
    void *dev[] = {m->d_fq,      m->d_nq,      m->d_qoff,     m->d_act,   m->d_visited, m->d_fa,
                   m->d_fb,      m->d_parent,  m->d_cnt,      m->d_scratch, m->d_roots, m->d_join,
                   m->d_sums,    m->rl.new2old, m->rl.old2new, m->rl.rowptr, m->rl.cols, m->rl.gdeg,
                   m->d_lane,    m->d_fq2,     m->rl.head,    m->d_level,   m->d_dirty,  m->rl.head_orig,
                   m->rl.cols_orig, m->d_lvflag};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (m->h_cnt) (void)hipHostFree(m->h_cnt);
    hipEvent_t evs[] = {m->ev0, m->ev1, m->ev_fin};
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : m->ek)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : m->ek_push)
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
    m->cont = false;
    m->busy = 0;
    m->nroots = nroots;
    std::copy(roots, roots + nroots, m->roots);
    int64_t *d_roots = m->d_roots;
    G5_HIP(hipMemcpyAsync(d_roots, roots, (size_t)nroots * 8, hipMemcpyHostToDevice, s));
    const bool dirty_listed = (int64_t)m->dirty.size() <= kDirtyCap;
    if (dirty_listed && !m->dirty.empty())
        G5_HIP(hipMemcpyAsync(m->d_dirty, m->dirty.data(), m->dirty.size() * 8, hipMemcpyHostToDevice, s));
    G5_HIP(hipStreamSynchronize(s));
    G5_HIP(hipEventRecord(m->ev0, s));
    // records of edgeless vertices that held an earlier root read unreached
    // again (every edgeless vertex's, when too many roots came since)
    const int64_t bin_end = m->rl.bin_first[kBins];
    if (!dirty_listed) {
        if (m->n > bin_end)
            hipLaunchKernelGGL(k_ms_clear_range, g5::grid_for(m->n - bin_end, 4, 16384), 256, 0, s,
                               (const uint32_t *)m->rl.new2old, bin_end, m->n, m->d_parent, m->d_level);
    } else if (!m->dirty.empty()) {
        hipLaunchKernelGGL(k_ms_clear_roots, g5::grid_for((int64_t)m->dirty.size(), 4, 4096), 256, 0, s,
                           (const int64_t *)m->d_dirty, (int64_t)m->dirty.size(), (const uint32_t *)m->rl.old2new,
                           bin_end, m->d_parent, m->d_level);
    }
    m->dirty.assign(roots, roots + nroots);
    const uint64_t unused = nroots == 64 ? 0ull : ~((1ull << nroots) - 1);
    // when the last level run was a push level that found nothing, its next
    // buffer is clear and its frontier buffer is non-zero only at the entries
    // of its frontier queue: clear those instead of both buffers
    const bool listed = m->listed_ok && m->last_empty && m->listed_len * kListedFrac < m->n;
    hipLaunchKernelGGL(k_ms_init, g5::grid_for(m->n, 256, 16384), 256, 0, s, m->d_visited, m->d_fa, m->d_fb,
                       m->d_act, m->n, unused, !listed);
    if (listed && m->listed_len) {
        uint64_t *last_front = m->p == 0 ? m->d_fb : m->d_fa;  // parity p ^ 1 read it
        hipLaunchKernelGGL(k_zero_listed, g5::grid_for(m->listed_len, 256, 4096), 256, 0, s, last_front,
                           (const uint32_t *)m->listed, m->listed_len);
    }
    hipLaunchKernelGGL(k_root_reach, nroots, 256, 0, s, (const int64_t *)d_roots, (const uint32_t *)m->rl.old2new,
                       (const int64_t *)m->rl.rowptr, (const uint32_t *)m->rl.cols, m->d_join);
    hipLaunchKernelGGL(k_ms_roots, 1, 64, 0, s, (const int64_t *)d_roots, nroots, (const uint32_t *)m->rl.old2new,
                       (const int64_t *)m->rl.rowptr, m->d_visited, m->d_fa, m->d_act, m->d_parent, m->d_level,
                       m->d_cnt, (int32_t *)(m->d_roots + 64), (const int64_t *)m->d_join);
    G5_HIP(hipGetLastError());
    int levels = 1;
    m->lvl.clear();
    // the first frontier is the set of distinct roots; its edge count is not
    // known on the host, and a handful of roots always takes the push path
    m->p = 0;
    m->flag_word = INT64_MAX;
    m->act_edges = -1;
    m->front_lanes = nroots == 64 ? ~0ull : (1ull << nroots) - 1;
    m->lane_known = 0;
    m->split_levels = 0;
    m->listed_ok = true;  // k_ms_init cleared both frontier buffers
    m->listed_len = 0;
    {
        int64_t tmp[64];
        std::copy(roots, roots + nroots, tmp);
        std::sort(tmp, tmp + nroots);
        m->act_verts = std::unique(tmp, tmp + nroots) - tmp;
    }
    for (;;) {
        int64_t new_bits = 0;
        uint64_t live = 0;
        int rc = run_level(m, &new_bits, &live);
        if (rc) return rc;
        if (new_bits == 0) break;
        ++levels;
    }
    // the maps are final once the used lanes' unreached slots say so
    G5_HIP(hipEventRecord(m->ev_fin, s));
    if (bin_end)
        hipLaunchKernelGGL(k_ms_finalize, g5::grid_for(bin_end, kBlock, 8192), kBlock, 0, s,
                           (const uint64_t *)m->d_visited, (const uint32_t *)m->rl.new2old, bin_end, ~unused,
                           m->d_parent, m->d_level);
    G5_HIP(hipGetLastError());
    G5_HIP(hipEventRecord(m->ev1, s));
    G5_HIP(hipMemcpyAsync(m->lane_of, m->d_roots + 64, 64 * 4, hipMemcpyDeviceToHost, s));
    G5_HIP(hipEventSynchronize(m->ev1));
    G5_HIP(hipStreamSynchronize(s));
    G5_HIP(hipEventElapsedTime(&m->search_ms, m->ev0, m->ev_fin));
    G5_HIP(hipEventElapsedTime(&m->finalize_ms, m->ev_fin, m->ev1));
    // level L is stored as the byte L: final up to level 254 (255: unreached)
    m->run_levels = levels;
    m->levels_final = levels <= 255;
    if (ms) *ms = m->search_ms + m->finalize_ms;
    if (nlevels) *nlevels = levels;
    return G5_OK;
}

extern "C" int g5_msbfs_open(g5_msbfs *m)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_open: null handle");
    hipStream_t s = m->s;
    // every lane free: free lanes count as visited everywhere, so they never
    // keep a vertex from the fully-visited early exit
    hipLaunchKernelGGL(k_ms_init, g5::grid_for(m->n, 256, 16384), 256, 0, s, m->d_visited, m->d_fa, m->d_fb,
                       m->d_act, m->n, ~0ull, true);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(s));
    m->cont = true;
    m->busy = 0;
    m->nroots = 0;
    m->p = 0;
    m->flag_word = INT64_MAX;
    m->act_verts = 0;
    m->act_edges = 0;
    m->front_lanes = 0;
    m->listed_ok = true;  // k_ms_init cleared both frontier buffers
    m->listed_len = 0;
    m->lane_known = 0;  // no lane has a frontier; joins add theirs
    std::fill(m->lane_edges, m->lane_edges + 64, 0);
    m->split_levels = 0;
    m->lvl.clear();
    return G5_OK;
}

extern "C" int g5_msbfs_join(g5_msbfs *m, const int64_t *roots, const int *lanes, int n)
{
    G5_REQUIRE(m && n >= 0 && n <= 64 && (n == 0 || (roots && lanes)), G5_E_ARG, "g5_msbfs_join: bad arguments");
    G5_REQUIRE(m->cont, G5_E_STATE, "g5_msbfs_join: g5_msbfs_open not called");
    if (n == 0) return G5_OK;
    uint64_t mask = 0;
    int64_t h[128];
    for (int i = 0; i < n; ++i) {
        G5_REQUIRE(lanes[i] >= 0 && lanes[i] < 64, G5_E_ARG, "g5_msbfs_join: lane %d out of range", lanes[i]);
        G5_REQUIRE(!((m->busy | mask) >> lanes[i] & 1), G5_E_STATE, "g5_msbfs_join: lane %d is not free", lanes[i]);
        G5_REQUIRE(roots[i] >= 0 && roots[i] < m->n, G5_E_ARG, "g5_msbfs_join: root %lld out of range",
                   (long long)roots[i]);
        mask |= 1ull << lanes[i];
        h[i] = roots[i];
        h[64 + i] = lanes[i];
    }
    hipStream_t s = m->s;
    uint64_t *front = m->p == 0 ? m->d_fa : m->d_fb;
    uint64_t *act = m->d_act + (m->p == 0 ? 0 : m->n / 64);
    G5_HIP(hipMemcpyAsync(m->d_join, h, sizeof(h), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_lane_masks, g5::grid_for(m->n, 256, 16384), 256, 0, s, m->d_visited, m->n, 0ull, mask);
    hipLaunchKernelGGL(k_join, 1, 1, 0, s, (const int64_t *)m->d_join, n, (const uint32_t *)m->rl.old2new,
                       (const int64_t *)m->rl.rowptr, m->d_visited, front, act, m->d_parent, m->d_level,
                       (const int64_t *)m->d_cnt, m->d_join + 128);
    G5_HIP(hipGetLastError());
    int64_t added[66] = {};
    G5_HIP(hipMemcpyAsync(added, m->d_join + 128, (size_t)(2 + n) * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    m->busy |= mask;
    for (int i = 0; i < n; ++i) {
        m->lane_root[lanes[i]] = roots[i];
        m->lane_steps[lanes[i]] = 0;
        if ((int64_t)m->dirty.size() <= kDirtyCap) m->dirty.push_back(roots[i]);
    }
    m->act_verts += added[0];
    m->act_edges += added[1];
    // a joined lane's frontier is its root alone (none if the root has no
    // neighbours): its frontier edges are the root's degree
    for (int i = 0; i < n; ++i) {
        m->lane_edges[lanes[i]] = added[2 + i];
        m->lane_known |= 1ull << lanes[i];
        m->last_empty = false;  // the roots sit in the buffer the next level reads
        if (added[2 + i]) m->front_lanes |= 1ull << lanes[i];
    }
    m->split_ok = true;
    m->flag_word = INT64_MAX;
    return G5_OK;
}

extern "C" int g5_msbfs_step(g5_msbfs *m, uint64_t *live, float *ms)
{
    G5_REQUIRE(m && live, G5_E_ARG, "g5_msbfs_step: null argument");
    G5_REQUIRE(m->cont, G5_E_STATE, "g5_msbfs_step: g5_msbfs_open not called");
    *live = 0;
    if (ms) *ms = 0;
    if (!m->busy) return G5_OK;
    G5_HIP(hipEventRecord(m->ev0, m->s));
    int64_t new_bits = 0;
    uint64_t l = 0;
    int rc = run_level(m, &new_bits, &l);
    if (rc) return rc;
    for (int r = 0; r < 64; ++r)
        if (m->busy >> r & 1) ++m->lane_steps[r];
    G5_HIP(hipEventRecord(m->ev1, m->s));
    G5_HIP(hipEventSynchronize(m->ev1));
    if (ms) G5_HIP(hipEventElapsedTime(ms, m->ev0, m->ev1));
    *live = l & m->busy;
    return G5_OK;
}

extern "C" int g5_msbfs_leave(g5_msbfs *m, uint64_t lanes, int64_t *nedge, uint64_t *checksum)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_leave: null handle");
    G5_REQUIRE(m->cont, G5_E_STATE, "g5_msbfs_leave: g5_msbfs_open not called");
    G5_REQUIRE((lanes & ~m->busy) == 0, G5_E_STATE, "g5_msbfs_leave: lane not in use");
    if (!lanes) return G5_OK;
    hipStream_t s = m->s;
    G5_HIP(g5::fill_u64((uint64_t *)m->d_sums, 128, 0, s));
    hipLaunchKernelGGL(k_leave, g5::grid_for(m->n, kBlock, 4096), kBlock, 0, s, m->d_visited,
                       (const int32_t *)m->d_parent, (const uint32_t *)m->rl.new2old, (const uint32_t *)m->rl.gdeg,
                       m->n, lanes, m->d_sums);
    G5_HIP(hipGetLastError());
    unsigned long long h[128];
    G5_HIP(hipMemcpyAsync(h, m->d_sums, sizeof(h), hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < 64; ++r)
        if (lanes >> r & 1) {
            if (checksum) checksum[r] = h[r];
            if (nedge) nedge[r] = (int64_t)(h[64 + r] / 2);
        }
    m->busy &= ~lanes;
    // a freed lane is visited everywhere: no pull or push finds anything on it
    m->front_lanes &= ~lanes;
    return G5_OK;
}

namespace {
// the lane that holds the caller's lane `lane` (batch runs rank the roots)
int internal_lane(const g5_msbfs *m, int lane) { return m->cont ? lane : m->lane_of[lane]; }
}  // namespace

namespace {
// the lane's level bytes hold its levels: a finalised batch run of depth
// below 255, or a continuous-mode lane that has run fewer than 255 levels
bool lane_levels_exact(const g5_msbfs *m, int lane)
{
    return m->cont ? m->lane_steps[lane] < 255 : m->levels_final;
}

// extraction of internal lane il into device arrays (lev may be null)
int extract_lane(g5_msbfs *m, int il, int64_t *d_pred, int32_t *d_lev)
{
    hipLaunchKernelGGL(k_ms_extract, g5::grid_for(m->n, 256, 16384), 256, 0, m->s, (const uint32_t *)m->rl.old2new,
                       (const int32_t *)m->d_parent, (const uint8_t *)m->d_level, (const uint64_t *)m->d_visited,
                       m->n, il, m->cont ? m->lane_root[il] : (int64_t)-1, d_pred, d_lev);
    G5_HIP(hipGetLastError());
    return G5_OK;
}
}  // namespace

extern "C" int g5_msbfs_result(g5_msbfs *m, int lane, int64_t *pred, int32_t *level)
{
    G5_REQUIRE(m && lane >= 0 && lane < 64 && (m->cont ? (m->busy >> lane & 1) != 0 : lane < m->nroots), G5_E_ARG,
               "g5_msbfs_result: lane %d not in use", lane);
    const int64_t n = m->n;
    hipStream_t s = m->s;
    const int il = internal_lane(m, lane);
    // levels from the level bytes, or (a search deeper than they hold) from
    // the parent tree by pointer jumping
    const bool jump = level && !lane_levels_exact(m, il);
    int64_t *dp = nullptr;
    uint64_t *dw = nullptr;
    int32_t *dl = nullptr;
    int *dflag = nullptr;
    auto cleanup = [&]() {
        if (dp) (void)hipFree(dp);
        if (dw) (void)hipFree(dw);
        if (dl) (void)hipFree(dl);
        if (dflag) (void)hipFree(dflag);
    };
    hipError_t e = hipMalloc(&dp, (size_t)n * 8);
    if (e == hipSuccess && level) e = hipMalloc(&dl, (size_t)n * 4);
    if (e == hipSuccess && jump) e = hipMalloc(&dw, (size_t)n * 8);
    if (e == hipSuccess && jump) e = hipMalloc(&dflag, sizeof(int));
    if (e != hipSuccess) {
        cleanup();
        G5_HIP(e);
    }
    int rc = extract_lane(m, il, dp, jump ? nullptr : dl);
    if (rc) {
        cleanup();
        return rc;
    }
    if (pred) e = hipMemcpyAsync(pred, dp, (size_t)n * 8, hipMemcpyDeviceToHost, s);
    int flags = 0;
    if (e == hipSuccess && jump) {
        hipLaunchKernelGGL(k_jump_init, g5::grid_for(n, 256, 16384), 256, 0, s, (const int64_t *)dp, n, dw);
        e = hipGetLastError();
        // each pass at least doubles every vertex's jump: 40 passes cover
        // any depth below 2^32; a parent cycle never converges
        bool converged = false;
        for (int pass = 0; pass < 40 && e == hipSuccess && !converged; ++pass) {
            e = g5::fill_i32(dflag, 1, 0, s);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_jump, g5::grid_for(n, 256, 16384), 256, 0, s, dw, n, dflag);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipMemcpyAsync(&flags, dflag, sizeof(int), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            converged = (flags & 1) == 0;
            if (flags & 2) break;
        }
        if (e == hipSuccess && !converged) flags |= 4;
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_jump_levels, g5::grid_for(n, 256, 16384), 256, 0, s, (const uint64_t *)dw, n, dl);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess && level) e = hipMemcpyAsync(level, dl, (size_t)n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    G5_HIP(e);
    G5_REQUIRE(!(flags & 2), G5_E_STATE, "g5_msbfs_result: lane %d has a vertex whose parent is unreached", lane);
    G5_REQUIRE(!(flags & 4), G5_E_STATE, "g5_msbfs_result: parent tree of lane %d has a cycle", lane);
    return G5_OK;
}

extern "C" int g5_msbfs_pred_device(g5_msbfs *m, int lane, int64_t *d_pred)
{
    G5_REQUIRE(m && d_pred, G5_E_ARG, "g5_msbfs_pred_device: null argument");
    G5_REQUIRE(lane >= 0 && lane < 64 && (m->cont ? (m->busy >> lane & 1) != 0 : lane < m->nroots), G5_E_ARG,
               "g5_msbfs_pred_device: lane %d not in use", lane);
    int rc = extract_lane(m, internal_lane(m, lane), d_pred, nullptr);
    if (rc) return rc;
    G5_HIP(hipStreamSynchronize(m->s));
    return G5_OK;
}

extern "C" int g5_msbfs_level_device(g5_msbfs *m, int lane, int32_t *d_level)
{
    G5_REQUIRE(m && d_level, G5_E_ARG, "g5_msbfs_level_device: null argument");
    G5_REQUIRE(lane >= 0 && lane < 64 && (m->cont ? (m->busy >> lane & 1) != 0 : lane < m->nroots), G5_E_ARG,
               "g5_msbfs_level_device: lane %d not in use", lane);
    const int il = internal_lane(m, lane);
    G5_REQUIRE(lane_levels_exact(m, il), G5_E_STATE,
               "g5_msbfs_level_device: lane %d searched 255 levels or more; use g5_msbfs_result", lane);
    int64_t *dp = nullptr;
    G5_HIP(hipMalloc(&dp, (size_t)m->n * 8));
    int rc = extract_lane(m, il, dp, d_level);
    hipError_t e = hipStreamSynchronize(m->s);
    (void)hipFree(dp);
    if (rc) return rc;
    G5_HIP(e);
    return G5_OK;
}

extern "C" int g5_msbfs_maps(g5_msbfs *m, const int32_t **d_pred, const uint8_t **d_level, int32_t *lane_of,
                             int *levels_final)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_maps: null handle");
    G5_REQUIRE(!m->cont && m->nroots > 0, G5_E_STATE, "g5_msbfs_maps: no batch run");
    if (d_pred) *d_pred = m->d_parent;
    if (d_level) *d_level = m->d_level;
    if (lane_of) std::copy(m->lane_of, m->lane_of + m->nroots, lane_of);
    if (levels_final) *levels_final = m->levels_final ? 1 : 0;
    return G5_OK;
}

extern "C" int g5_msbfs_times(g5_msbfs *m, float *search_ms, float *finalize_ms)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_times: null handle");
    G5_REQUIRE(!m->cont && m->nroots > 0, G5_E_STATE, "g5_msbfs_times: no batch run");
    if (search_ms) *search_ms = m->search_ms;
    if (finalize_ms) *finalize_ms = m->finalize_ms;
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
    for (int r = 0; r < m->nroots; ++r) nedge[r] = (int64_t)(h[internal_lane(m, r)] / 2);
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
    if (cap > kBins) bins[kBins] = m->split_levels;
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

namespace {

int64_t *tune_slot(g5_msbfs *m, int key)
{
    switch (key) {
    case G5_TUNE_THREAD_MAX: return &m->tune.thread_max;
    case G5_TUNE_WAVE_MAX: return &m->tune.wave_max;
    case G5_TUNE_PUSH_RATIO: return &m->tune.push_ratio;
    case G5_TUNE_ACTIVE_RATIO: return &m->tune.active_ratio;
    case G5_TUNE_LANE_RATIO: return &m->tune.lane_ratio;
    case G5_TUNE_SPLIT_MAX: return &m->tune.split_max;
    default: return nullptr;
    }
}

// Installs t: new degree bins and re-captured level graphs when the bounds move.
int apply_tune(g5_msbfs *m, const Tune &t)
{
    const bool rebin = t.thread_max != m->tune.thread_max || t.wave_max != m->tune.wave_max ||
                       t.split_max != m->tune.split_max;
    m->tune = t;
    if (!rebin) return G5_OK;
    G5_HIP(hipStreamSynchronize(m->s));
    int rc = compute_bins(m);
    if (rc) return rc;
    destroy_graphs(m);
    return capture_graphs(m);
}

}  // namespace

extern "C" int g5_msbfs_tune(g5_msbfs *m, int key, int64_t value)
{
    G5_REQUIRE(m, G5_E_ARG, "g5_msbfs_tune: null handle");
    int64_t *slot = tune_slot(m, key);
    G5_REQUIRE(slot, G5_E_ARG, "g5_msbfs_tune: unknown key %d", key);
    Tune t = m->tune;
    switch (key) {
    case G5_TUNE_THREAD_MAX: t.thread_max = value; break;
    case G5_TUNE_WAVE_MAX: t.wave_max = value; break;
    case G5_TUNE_PUSH_RATIO: t.push_ratio = value; break;
    case G5_TUNE_ACTIVE_RATIO: t.active_ratio = value; break;
    case G5_TUNE_LANE_RATIO: t.lane_ratio = value; break;
    default: t.split_max = value; break;
    }
    G5_REQUIRE(t.thread_max >= 0 && t.wave_max >= t.thread_max && t.wave_max <= (int64_t)1 << 40, G5_E_ARG,
               "g5_msbfs_tune: need 0 <= thread_max <= wave_max");
    G5_REQUIRE(t.push_ratio >= 1 && t.active_ratio >= 1, G5_E_ARG, "g5_msbfs_tune: ratios must be >= 1");
    G5_REQUIRE(t.lane_ratio >= 0 && t.lane_ratio <= (int64_t)1 << 40, G5_E_ARG,
               "g5_msbfs_tune: lane_ratio must be >= 0");
    G5_REQUIRE(t.split_max >= 0 && t.split_max <= (int64_t)1 << 40, G5_E_ARG,
               "g5_msbfs_tune: split_max must be in [0, 2^40]");
    return apply_tune(m, t);
}

extern "C" int g5_msbfs_get_tune(g5_msbfs *m, int key, int64_t *value)
{
    G5_REQUIRE(m && value, G5_E_ARG, "g5_msbfs_get_tune: null argument");
    int64_t *slot = tune_slot(m, key);
    G5_REQUIRE(slot, G5_E_ARG, "g5_msbfs_get_tune: unknown key %d", key);
    *value = *slot;
    return G5_OK;
}

extern "C" int g5_msbfs_autotune(g5_msbfs *m, const int64_t *roots, int nroots, int passes, double *best_ms,
                                 int *trials)
{
    G5_REQUIRE(m && roots && nroots >= 1 && nroots <= 64 && passes >= 1 && passes <= 4, G5_E_ARG,
               "g5_msbfs_autotune: bad arguments");
    G5_REQUIRE(!(m->cont && m->busy), G5_E_STATE, "g5_msbfs_autotune: lanes in use");
    // candidates per key (0 ends a list); -1 stands for 0 (lane_ratio: never
    // split; split_max: no thread bin on split levels).  The two thread-bin
    // bounds are separate keys, so that the descent can raise the pull
    // levels' bound while the split levels keep theirs.
    static const int64_t cand[6][7] = {{8, 16, 32, 64, 128, 256, 512},
                                       {512, 1024, 2048, 4096, 8192, 16384, 0},
                                       {4, 8, 16, 32, 64, 0, 0},
                                       {1, 2, 4, 8, 0, 0, 0},
                                       {-1, 16, 32, 64, 128, 256, 0},
                                       {-1, 8, 16, 32, 64, 128, 0}};
    const bool prof = m->prof;
    m->prof = false;
    int n_trials = 0;
    // a configuration's time: the better of two searches
    auto measure = [&](const Tune &t, double *out) -> int {
        int rc = apply_tune(m, t);
        if (rc) return rc;
        double best = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            int32_t nl = 0;
            float ms = 0;
            rc = g5_msbfs_run(m, roots, nroots, &nl, &ms);
            if (rc) return rc;
            best = std::min(best, (double)ms);
        }
        ++n_trials;
        *out = best;
        return G5_OK;
    };
    Tune best = m->tune;
    double best_t = 0;
    int rc = measure(best, &best_t);
    for (int pass = 0; pass < passes && !rc; ++pass)
        for (int key = 0; key < 6 && !rc; ++key)
            for (int i = 0; i < 7 && !rc; ++i) {
                if (cand[key][i] == 0) break;
                const int64_t v = cand[key][i] < 0 ? 0 : cand[key][i];
                Tune t = best;
                int64_t *f = key == 0   ? &t.thread_max
                             : key == 1 ? &t.wave_max
                             : key == 2 ? &t.push_ratio
                             : key == 3 ? &t.active_ratio
                             : key == 4 ? &t.lane_ratio
                                        : &t.split_max;
                if (*f == v) continue;
                *f = v;
                if (t.wave_max < t.thread_max) continue;
                double ms = 0;
                rc = measure(t, &ms);
                if (!rc && ms < best_t * 0.99) {
                    best = t;
                    best_t = ms;
                }
            }
    if (!rc) rc = apply_tune(m, best);
    m->prof = prof;
    if (rc) return rc;
    if (best_ms) *best_ms = best_t;
    if (trials) *trials = n_trials;
    return G5_OK;
}

// Placement of the parent records.  The batched search runs at one of two
// speeds (about 12 % apart at scale 26) depending on the physical memory
// behind the 256 B parent records, which take scattered whole and partial
// record stores: moving only the records to a fresh allocation switches
// between the two, at the same virtual address as well, while moving any
// other buffer changes nothing.  g5_msbfs_place tries `trials` fresh
// allocations (each a copy of the current records, so the records stay
// valid) and keeps the one whose search of `roots` is fastest (best of two
// runs; a move must gain 2 %).  It needs the records' size again in free
// memory while it runs and stops trying when an allocation fails.
extern "C" int g5_msbfs_place(g5_msbfs *m, const int64_t *roots, int nroots, int trials, float *first_ms,
                              float *best_ms, int *moves)
{
    G5_REQUIRE(m && roots && nroots >= 1 && nroots <= 64 && trials >= 0 && trials <= 16, G5_E_ARG,
               "g5_msbfs_place: bad arguments");
    G5_REQUIRE(!(m->cont && m->busy), G5_E_STATE, "g5_msbfs_place: lanes in use");
    const size_t bytes = (size_t)m->n * 64 * 4;
    hipStream_t s = m->s;
    const bool prof = m->prof;
    m->prof = false;
    auto time_runs = [&](float *out) -> int {
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            int32_t nl = 0;
            float ms = 0;
            int rc = g5_msbfs_run(m, roots, nroots, &nl, &ms);
            if (rc) return rc;
            best = std::min(best, ms);
        }
        *out = best;
        return G5_OK;
    };
    // the captured pull levels hold the records' address
    auto swap_records = [&](int32_t *to) -> int {
        m->d_parent = to;
        destroy_graphs(m);
        return capture_graphs(m);
    };
    float cur = 0;
    int kept = 0;
    int rc = time_runs(&cur);
    if (first_ms) *first_ms = cur;
    for (int t = 0; t < trials && !rc; ++t) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) {
            (void)hipGetLastError();
            break;
        }
        int32_t *old = m->d_parent;
        hipError_t e = hipMemcpyAsync(q, old, bytes, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            (void)hipFree(q);
            rc = g5::fail_hip(e, "g5_msbfs_place: copy", __FILE__, __LINE__);
            break;
        }
        rc = swap_records((int32_t *)q);
        float ms = 0;
        if (!rc) rc = time_runs(&ms);
        if (!rc && ms < cur * 0.98f) {
            // the copy ran the same roots as the original: both hold their maps
            (void)hipFree(old);
            cur = ms;
            ++kept;
        } else {
            const int rc2 = swap_records(old);
            (void)hipFree(q);
            if (!rc) rc = rc2;
        }
    }
    m->prof = prof;
    if (rc) return rc;
    if (best_ms) *best_ms = cur;
    if (moves) *moves = kept;
    return G5_OK;
}

extern "C" int g5_msbfs_reach(g5_msbfs *m, const int64_t *roots, int n, int64_t *reach)
{
    G5_REQUIRE(m && n >= 0 && n <= 64 && (n == 0 || (roots && reach)), G5_E_ARG, "g5_msbfs_reach: bad arguments");
    if (n == 0) return G5_OK;
    for (int i = 0; i < n; ++i)
        G5_REQUIRE(roots[i] >= 0 && roots[i] < m->n, G5_E_ARG, "g5_msbfs_reach: root %lld out of range",
                   (long long)roots[i]);
    hipStream_t s = m->s;
    G5_HIP(hipStreamSynchronize(s));  // d_roots and d_sums are free between calls
    G5_HIP(hipMemcpyAsync(m->d_roots, roots, (size_t)n * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_root_reach, n, 256, 0, s, (const int64_t *)m->d_roots, (const uint32_t *)m->rl.old2new,
                       (const int64_t *)m->rl.rowptr, (const uint32_t *)m->rl.cols, (int64_t *)m->d_sums);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpyAsync(reach, m->d_sums, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    return G5_OK;
}
