// g5_validate.hip -- Graph500 result validation on the device
// (validate_bfs_result, cpu_2d/validate/validate.h:249-491) for a whole (1x1)
// graph.
//
// The reference checks a predecessor map against the generated edge list:
// value ranges (:69-105), the root is its own parent (:285-295), no other
// vertex is its own parent (:300-342), a depth map built from the parents
// (build_bfs_depth_map, :110-242: 16-bit depths, 65535 = unreached), every
// edge's endpoint depths differ by at most one (:425-433), and every vertex
// with a parent other than itself has an edge to it (:436-471).  It counts
// the directed edge copies whose source has a parent; half that is nedge.
//
// Here the edge list is the CSR (sorted, deduplicated, self-loops dropped).
// A self-loop never fails the depth check and never proves a parent edge (a
// vertex that is its own parent is exempt from that check), and a duplicate
// repeats a check, so every flag comes out the same; the count uses the
// generated degree gdeg, which holds every directed copy, duplicates and
// self-loops included.
//
// Depth map: pointer doubling over words (dist << 32 | ancestor).  The root
// is a fixed point at distance 0 and a vertex without a parent is the
// terminal kNone; a jump whose distance reaches 65535 also becomes kNone, so
// a vertex is reached exactly when the reference's relaxation gives it a
// depth below 65535 (a parent cycle away from the root saturates within 17
// rounds).  The reference's "inconsistent depth" failure cannot occur with a
// set-once depth per vertex; it is never raised.
//
// Edge checks: one thread per row for rows of at most kLight neighbours
// (scanning the row also finds the parent), rows above that are split into
// pieces of kPiece neighbours, one wavefront per piece, with the parent found
// by binary search in the sorted row.  When the deepest level is below 254
// the depths are narrowed to bytes first, halving the gathered footprint.
#include "g5_internal.hpp"

struct g5_validator {
    const g5_graph *g = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t *d_pred = nullptr;
    uint64_t *d_j[2] = {nullptr, nullptr};
    uint16_t *d_d16 = nullptr;
    uint8_t *d_d8 = nullptr;
    int32_t *d_depth = nullptr;
    uint64_t *d_pieces = nullptr;
    int64_t n_pieces = 0;
    unsigned long long *d_acc = nullptr;
    unsigned long long *h_acc = nullptr;
};

namespace {

constexpr uint64_t kNone = ~0ull;
constexpr int64_t kLight = 32;
constexpr int64_t kPiece = 2048;
constexpr int kMaxRounds = 24;
enum Acc { kFlags = 0, kEdgeSum = 1, kMaxDepth = 2, kPieceCount = 3, kChanged = 4, kNAcc = 8 };

__device__ __forceinline__ uint64_t pack(uint64_t dist, uint64_t anc) { return dist << 32 | anc; }

__device__ __forceinline__ void wave_flags(uint32_t f, unsigned long long *acc)
{
    f = g5_wave_or32(f);
    if (f && g5_lane() == 0) atomicOr(acc + kFlags, (unsigned long long)f);
}

// the workgroup's total of x added to *dst with one atomic (every thread of
// the 256-thread workgroup calls it; same-address atomics serialise, so one
// per workgroup rather than per wave)
__device__ __forceinline__ void block_sum_add(int64_t x, unsigned long long *dst)
{
    __shared__ int64_t s_part[4];
    x = g5_wave_sum(x);
    if (g5_lane() == 0) s_part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int64_t t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
        if (t) atomicAdd(dst, (unsigned long long)t);
    }
}

// ranges, root and self-parent checks, the jump words, and sum(gdeg) over
// vertices with a parent
__global__ void k_val_init(const int64_t *__restrict__ pred, int64_t n, int64_t root, int64_t nglob,
                           const uint32_t *__restrict__ gdeg, uint64_t *__restrict__ J,
                           unsigned long long *__restrict__ acc)
{
    uint32_t f = 0;
    int64_t sum = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = pred[v];
        if (p < -1 || p >= nglob || p >= n) f |= G5_BAD_RANGE;
        uint64_t w;
        if (v == root) {
            if (p != root) f |= G5_BAD_ROOT;
            w = pack(0, (uint64_t)root);
        } else {
            if (p == v) f |= G5_BAD_SELF_PARENT;
            w = (p < 0 || p >= n) ? kNone : pack(1, (uint64_t)p);
        }
        J[v] = w;
        if (p != -1) sum += gdeg[v];
    }
    wave_flags(f, acc);
    block_sum_add(sum, acc + kEdgeSum);
}

// one doubling round: J'(v) = J(anc(v)) after J(v), distances added and
// saturated to kNone at 65535
__global__ void k_val_jump(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t n, uint32_t root,
                           unsigned long long *__restrict__ acc)
{
    uint32_t changed = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t w = in[v];
        uint64_t nw = w;
        const uint32_t a = (uint32_t)w;
        if (w != kNone && a != root) {
            const uint64_t w2 = in[a];
            if (w2 == kNone) {
                nw = kNone;
            } else {
                const uint64_t d = (w >> 32) + (w2 >> 32);
                nw = d >= 65535 ? kNone : pack(d, (uint32_t)w2);
            }
        }
        out[v] = nw;
        changed |= nw != w;
    }
    if (__any(changed) && g5_lane() == 0) atomicOr(acc + kChanged, 1ull);
}

// 16-bit depths (65535 = unreached), the optional int32 copy (-1 unreached)
// and the deepest level
__global__ void k_val_depth(const uint64_t *__restrict__ J, int64_t n, uint32_t root, uint16_t *__restrict__ d16,
                            int32_t *__restrict__ d32, unsigned long long *__restrict__ acc)
{
    uint32_t mx = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t w = J[v];
        const bool reached = w != kNone && (uint32_t)w == root;
        const uint32_t d = reached ? (uint32_t)(w >> 32) : 65535u;
        d16[v] = (uint16_t)d;
        if (d32) d32[v] = reached ? (int32_t)d : -1;
        if (reached) mx = max(mx, d);
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, s, 64));
    if (g5_lane() == 0) atomicMax(acc + kMaxDepth, (unsigned long long)mx);
}

__global__ void k_val_narrow(const uint16_t *__restrict__ d16, int64_t n, uint8_t *__restrict__ d8)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint16_t d = d16[v];
        d8[v] = d == 65535 ? (uint8_t)255 : (uint8_t)d;
    }
}

template <typename D>
__device__ __forceinline__ bool far_apart(D a, D b)
{
    return (int)a - (int)b > 1 || (int)b - (int)a > 1;
}

// pieces of the rows above kLight neighbours (sizing pass, once per graph)
__global__ void k_val_count_pieces(const int64_t *__restrict__ rowptr, int64_t n, unsigned long long *cnt)
{
    int64_t c = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t deg = rowptr[v + 1] - rowptr[v];
        if (deg > kLight) c += (deg + kPiece - 1) / kPiece;
    }
    block_sum_add(c, cnt);
}

// thread per row: light rows are checked whole; heavy rows get a binary
// search for the parent and their pieces listed
template <typename D>
__global__ void k_val_rows(const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
                           const int64_t *__restrict__ pred, const D *__restrict__ depth, int64_t n,
                           uint64_t *__restrict__ pieces, int64_t piece_cap, unsigned long long *__restrict__ acc)
{
    uint32_t f = 0;
    // n is a multiple of 64, so every lane of a wave runs the same iterations
    // and the wave scan below sees all 64 lanes
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = rowptr[v], e = rowptr[v + 1];
        const int64_t p = pred[v];
        const bool need_parent = p != -1 && p != v;
        const D dv = depth[v];
        int64_t np = 0;  // pieces of a heavy row
        if (e - b <= kLight) {
            bool found = false;
            for (int64_t k = b; k < e; ++k) {
                const uint32_t u = cols[k];
                if (far_apart(dv, depth[u])) f |= G5_BAD_EDGE_LEVEL;
                found |= (int64_t)u == p;
            }
            if (need_parent && !found) f |= G5_BAD_PARENT_EDGE;
        } else {
            if (need_parent) {
                int64_t lo = b, hi = e;  // first col >= p
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if ((int64_t)cols[mid] < p) lo = mid + 1;
                    else hi = mid;
                }
                if (lo == e || (int64_t)cols[lo] != p) f |= G5_BAD_PARENT_EDGE;
            }
            np = (e - b + kPiece - 1) / kPiece;
        }
        // the wave's pieces take one contiguous range of the list
        const int64_t incl = g5_wave_incl_scan(np);
        const int64_t total = __shfl(incl, 63, 64);
        int64_t base = 0;
        if (total) {
            if (g5_lane() == 63) base = (int64_t)atomicAdd(acc + kPieceCount, (unsigned long long)total);
            base = __shfl(base, 63, 64) + incl - np;
        }
        for (int64_t j = 0; j < np && base + j < piece_cap; ++j) pieces[base + j] = (uint64_t)v << 32 | (uint64_t)j;
    }
    wave_flags(f, acc);
}

// one wavefront per piece of a heavy row: the depth check of kPiece neighbours
template <typename D>
__global__ void k_val_pieces(const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
                             const D *__restrict__ depth, const uint64_t *__restrict__ pieces,
                             const unsigned long long *__restrict__ count, int64_t piece_cap,
                             unsigned long long *__restrict__ acc)
{
    const int64_t np = min((int64_t)*count, piece_cap);
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t lane = g5_lane();
    uint32_t f = 0;
    for (int64_t i = wave; i < np; i += nwaves) {
        const uint64_t pc = pieces[i];
        const int64_t v = (int64_t)(pc >> 32);
        const int64_t b = rowptr[v] + (int64_t)(uint32_t)pc * kPiece;
        const int64_t e = min(rowptr[v + 1], b + kPiece);
        const D dv = depth[v];
        for (int64_t k = b + lane; k < e; k += 64)
            if (far_apart(dv, depth[cols[k]])) f |= G5_BAD_EDGE_LEVEL;
    }
    wave_flags(f, acc);
}

template <typename D>
hipError_t edge_checks(g5_validator *v, const D *depth)
{
    const g5_graph *g = v->g;
    hipLaunchKernelGGL(k_val_rows<D>, g5::grid_for(g->n_rows, 256, 32768), 256, 0, v->s,
                       (const int64_t *)g->d_rowptr, (const uint32_t *)g->d_cols, (const int64_t *)v->d_pred, depth,
                       g->n_rows, v->d_pieces, v->n_pieces, v->d_acc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || v->n_pieces == 0) return e;
    hipLaunchKernelGGL(k_val_pieces<D>, g5::grid_for(v->n_pieces * 64, 256, 32768), 256, 0, v->s,
                       (const int64_t *)g->d_rowptr, (const uint32_t *)g->d_cols, depth,
                       (const uint64_t *)v->d_pieces, (const unsigned long long *)(v->d_acc + kPieceCount),
                       v->n_pieces, v->d_acc);
    return hipGetLastError();
}

int read_acc(g5_validator *v)
{
    G5_HIP(hipMemcpyAsync(v->h_acc, v->d_acc, kNAcc * 8, hipMemcpyDeviceToHost, v->s));
    G5_HIP(hipStreamSynchronize(v->s));
    return G5_OK;
}

void release(g5_validator *v)
{
    void *dev[] = {v->d_pred, v->d_j[0], v->d_j[1], v->d_d16, v->d_d8, v->d_depth, v->d_pieces, v->d_acc};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (v->h_acc) (void)hipHostFree(v->h_acc);
    if (v->ev0) (void)hipEventDestroy(v->ev0);
    if (v->ev1) (void)hipEventDestroy(v->ev1);
    if (v->s) (void)hipStreamDestroy(v->s);
    delete v;
}

int validate(g5_validator *v, int64_t root, int64_t nglob, const int64_t *pred, bool pred_on_device, int *bad,
             int64_t *nedge, int32_t *depth, float *ms)
{
    G5_REQUIRE(v && pred && bad && nedge, G5_E_ARG, "g5_validate: null argument");
    const g5_graph *g = v->g;
    const int64_t n = g->n_rows;
    hipStream_t s = v->s;
    *bad = 0;
    *nedge = 0;
    G5_HIP(hipEventRecord(v->ev0, s));
    if (root < 0 || root >= nglob || root >= n) {  // validate.h:258-264
        *bad = G5_BAD_RANGE;
        if (ms) *ms = 0.f;
        return G5_OK;
    }
    if (pred_on_device) G5_HIP(hipMemcpyAsync(v->d_pred, pred, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    else G5_HIP(hipMemcpyAsync(v->d_pred, pred, (size_t)n * 8, hipMemcpyHostToDevice, s));
    G5_HIP(g5::fill_u64((uint64_t *)v->d_acc, kNAcc, 0, s));
    const int grid = g5::grid_for(n, 256, 32768);
    hipLaunchKernelGGL(k_val_init, grid, 256, 0, s, (const int64_t *)v->d_pred, n, root, nglob,
                       (const uint32_t *)g->d_gdeg, v->d_j[0], v->d_acc);
    G5_HIP(hipGetLastError());
    int rc = read_acc(v);
    if (rc) return rc;
    if (v->h_acc[kFlags] & G5_BAD_RANGE) {  // the reference stops at a range error
        *bad = G5_BAD_RANGE;
        G5_HIP(hipEventRecord(v->ev1, s));
        G5_HIP(hipEventSynchronize(v->ev1));
        if (ms) G5_HIP(hipEventElapsedTime(ms, v->ev0, v->ev1));
        return G5_OK;
    }
    int cur = 0;
    bool converged = false;
    for (int r = 0; r < kMaxRounds && !converged; ++r) {
        G5_HIP(g5::fill_u64((uint64_t *)v->d_acc + kChanged, 1, 0, s));
        hipLaunchKernelGGL(k_val_jump, grid, 256, 0, s, (const uint64_t *)v->d_j[cur], v->d_j[cur ^ 1], n,
                           (uint32_t)root, v->d_acc);
        G5_HIP(hipGetLastError());
        cur ^= 1;
        rc = read_acc(v);
        if (rc) return rc;
        converged = v->h_acc[kChanged] == 0;
    }
    G5_REQUIRE(converged, G5_E_STATE, "g5_validate: depth map did not converge in %d rounds", kMaxRounds);
    hipLaunchKernelGGL(k_val_depth, grid, 256, 0, s, (const uint64_t *)v->d_j[cur], n, (uint32_t)root, v->d_d16,
                       depth ? v->d_depth : nullptr, v->d_acc);
    G5_HIP(hipGetLastError());
    rc = read_acc(v);
    if (rc) return rc;
    if (v->h_acc[kMaxDepth] <= 253) {
        hipLaunchKernelGGL(k_val_narrow, grid, 256, 0, s, (const uint16_t *)v->d_d16, n, v->d_d8);
        G5_HIP(hipGetLastError());
        G5_HIP(edge_checks<uint8_t>(v, v->d_d8));
    } else {
        G5_HIP(edge_checks<uint16_t>(v, v->d_d16));
    }
    G5_HIP(hipEventRecord(v->ev1, s));
    rc = read_acc(v);
    if (rc) return rc;
    G5_REQUIRE((int64_t)v->h_acc[kPieceCount] <= v->n_pieces, G5_E_STATE,
               "g5_validate: %lld row pieces exceed the %lld sized for the graph",
               (long long)v->h_acc[kPieceCount], (long long)v->n_pieces);
    if (depth) {
        G5_HIP(hipMemcpyAsync(depth, v->d_depth, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        G5_HIP(hipStreamSynchronize(s));
    }
    *bad = (int)v->h_acc[kFlags];
    *nedge = (int64_t)(v->h_acc[kEdgeSum] / 2);
    if (ms) G5_HIP(hipEventElapsedTime(ms, v->ev0, v->ev1));
    return G5_OK;
}

}  // namespace

extern "C" int g5_validator_create(const g5_graph *g, g5_validator **out)
{
    G5_REQUIRE(g && out, G5_E_ARG, "g5_validator_create: null argument");
    G5_REQUIRE(g->n_rows == g->n_cols && g->row_start == 0 && g->col_start == 0, G5_E_STATE,
               "g5_validator_create: needs a whole (1x1) graph");
    G5_REQUIRE(g->n_rows < (int64_t)0xFFFFFFFF, G5_E_ARG, "g5_validator_create: %lld vertices exceed 32-bit ids",
               (long long)g->n_rows);
    G5_REQUIRE(g->n_rows % 64 == 0, G5_E_ARG, "g5_validator_create: rows must be a multiple of 64");
    G5_HIP(hipSetDevice(g->device));
    g5_validator *v = new g5_validator;
    v->g = g;
    const int64_t n = g->n_rows ? g->n_rows : 1;
    hipError_t e = hipStreamCreate(&v->s);
    if (e == hipSuccess) e = hipEventCreate(&v->ev0);
    if (e == hipSuccess) e = hipEventCreate(&v->ev1);
    if (e == hipSuccess) e = hipMalloc(&v->d_pred, (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&v->d_j[0], (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&v->d_j[1], (size_t)n * 8);
    if (e == hipSuccess) e = hipMalloc(&v->d_d16, (size_t)n * 2);
    if (e == hipSuccess) e = hipMalloc(&v->d_d8, (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&v->d_depth, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&v->d_acc, kNAcc * 8);
    if (e == hipSuccess) e = hipHostMalloc(&v->h_acc, kNAcc * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = g5::fill_u64((uint64_t *)v->d_acc, kNAcc, 0, v->s);
    if (e == hipSuccess && g->n_rows > 0) {
        hipLaunchKernelGGL(k_val_count_pieces, g5::grid_for(g->n_rows, 256, 32768), 256, 0, v->s,
                           (const int64_t *)g->d_rowptr, g->n_rows, v->d_acc + kPieceCount);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(v->h_acc, v->d_acc, kNAcc * 8, hipMemcpyDeviceToHost, v->s);
    if (e == hipSuccess) e = hipStreamSynchronize(v->s);
    if (e == hipSuccess) {
        v->n_pieces = (int64_t)v->h_acc[kPieceCount];
        e = hipMalloc(&v->d_pieces, (size_t)(v->n_pieces ? v->n_pieces : 1) * 8);
    }
    if (e != hipSuccess) {
        release(v);
        return g5::fail_hip(e, "g5_validator_create", __FILE__, __LINE__);
    }
    *out = v;
    return G5_OK;
}

extern "C" int g5_validator_free(g5_validator *v)
{
    if (v) {
        (void)hipStreamSynchronize(v->s);
        release(v);
    }
    return G5_OK;
}

extern "C" int g5_validate(g5_validator *v, int64_t root, int64_t nglobalverts, const int64_t *pred, int *bad,
                           int64_t *nedge, int32_t *depth, float *ms)
{
    return validate(v, root, nglobalverts, pred, false, bad, nedge, depth, ms);
}

extern "C" int g5_validate_device(g5_validator *v, int64_t root, int64_t nglobalverts, const int64_t *d_pred,
                                  int *bad, int64_t *nedge, int32_t *depth, float *ms)
{
    return validate(v, root, nglobalverts, d_pred, true, bad, nedge, depth, ms);
}
