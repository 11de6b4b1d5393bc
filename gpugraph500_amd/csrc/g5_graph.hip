// g5_graph.hip -- the rank's DistMatrix2d block in HBM: upload of a host CSR,
// and device-side Graph500 Kronecker generation + CSR construction.
//
// Generation follows generate_kronecker_range (cpu_2d/generator/
// graph_generator.c:157-187): every edge index ei owns an MRG stream skipped
// ahead by ei * 2^64 steps (splittable_mrg.c:105-125), so edges are generated
// independently, one thread per edge.  The skip matrices A^(v * 2^(8b)) are
// computed on the host by squaring the one-step matrix modulo 2^31-1.
//
// CSR construction follows setupMatrix2 (cpu_2d/distmatrix2d.hh:745-1166) at a
// 1x1 grid: both directions of every edge, sorted by (row, col), duplicates
// merged, self-loops dropped, rows padded to a multiple of 64.  The sort and
// the unique/select passes are rocPRIM device primitives (construction is not
// on the BFS hot path).
#include <cstring>
#include <random>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "g5_internal.hpp"

namespace {

constexpr uint32_t kP = 0x7FFFFFFFu;
constexpr uint32_t kX = 107374182u;
constexpr uint32_t kY = 104480u;

struct Mat5 {
    uint32_t m[25];
};

uint32_t h_mulmod(uint32_t a, uint32_t b) { return (uint32_t)((uint64_t)a * b % kP); }

Mat5 h_matmul(const Mat5 &a, const Mat5 &b)
{
    Mat5 r;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            uint64_t s = 0;
            for (int k = 0; k < 5; ++k) s += h_mulmod(a.m[i * 5 + k], b.m[k * 5 + j]);
            r.m[i * 5 + j] = (uint32_t)(s % kP);
        }
    return r;
}

void h_apply(const Mat5 &a, uint32_t z[5])
{
    uint32_t o[5];
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
        for (int k = 0; k < 5; ++k) s += h_mulmod(a.m[i * 5 + k], z[k]);
        o[i] = (uint32_t)(s % kP);
    }
    std::memcpy(z, o, sizeof(o));
}

uint32_t h_step(uint32_t z[5])
{
    const uint32_t n = (uint32_t)(((uint64_t)h_mulmod(z[0], kX) + h_mulmod(z[4], kY)) % kP);
    z[4] = z[3];
    z[3] = z[2];
    z[2] = z[1];
    z[1] = z[0];
    z[0] = n;
    return n;
}

// Host tables: pow2[b] = A^(2^(8b)), b = 0..23.
struct MrgTables {
    Mat5 pow2[24];
    MrgTables()
    {
        Mat5 a{};
        a.m[0] = kX;
        a.m[4] = kY;
        for (int i = 1; i < 5; ++i) a.m[i * 5 + i - 1] = 1;
        for (int b = 0; b < 24; ++b) {
            pow2[b] = a;
            for (int s = 0; s < 8; ++s) a = h_matmul(a, a);
        }
    }
    Mat5 power(int b, unsigned v) const  // A^(v * 2^(8b))
    {
        Mat5 r{};
        for (int i = 0; i < 5; ++i) r.m[i * 6] = 1;
        Mat5 p = pow2[b];
        while (v) {
            if (v & 1) r = h_matmul(r, p);
            p = h_matmul(p, p);
            v >>= 1;
        }
        return r;
    }
};

const MrgTables &tables()
{
    static MrgTables t;
    return t;
}

struct KronSetup {
    uint32_t base[5];
    uint64_t val0, val1;
};

KronSetup kron_setup(uint64_t s1, uint64_t s2)
{
    KronSetup k;
    // make_mrg_seed, cpu_2d/generator/utils.c:61-68
    k.base[0] = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
    k.base[1] = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
    k.base[2] = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
    k.base[3] = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
    k.base[4] = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
    // scramble constants: skip by 50 * 2^128 + 7 * 2^64 (graph_generator.c:169-179)
    uint32_t z[5];
    std::memcpy(z, k.base, sizeof(z));
    h_apply(tables().power(8, 7), z);
    h_apply(tables().power(16, 50), z);
    k.val0 = h_step(z);
    k.val0 *= 0xFFFFFFFFull;
    k.val0 += h_step(z);
    k.val1 = h_step(z);
    k.val1 *= 0xFFFFFFFFull;
    k.val1 += h_step(z);
    return k;
}

// Device table: skip[k][v] = A^(v * 2^(64 + 8k)), k = 0..7 (the bytes of ei).
struct DevSkip {
    uint32_t *d = nullptr;
    int init(int device)
    {
        if (d) return G5_OK;
        std::vector<uint32_t> h(8 * 256 * 25);
        const MrgTables &t = tables();
        for (int k = 0; k < 8; ++k) {
            Mat5 p = t.pow2[8 + k], acc{};
            for (int i = 0; i < 5; ++i) acc.m[i * 6] = 1;
            for (int v = 0; v < 256; ++v) {
                std::memcpy(&h[(k * 256 + v) * 25], acc.m, sizeof(acc.m));
                acc = h_matmul(acc, p);
            }
        }
        (void)device;
        G5_HIP(hipMalloc(&d, h.size() * 4));
        G5_HIP(hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        return G5_OK;
    }
};

DevSkip g_skip;

__device__ __forceinline__ uint32_t d_red(uint64_t x)  // x < 2^62 -> x mod (2^31-1)
{
    x = (x & kP) + (x >> 31);
    x = (x & kP) + (x >> 31);
    return (uint32_t)(x >= kP ? x - kP : x);
}

__device__ __forceinline__ uint32_t d_step(uint32_t z[5])
{
    const uint32_t n = d_red((uint64_t)z[0] * kX + (uint64_t)z[4] * kY);
    z[4] = z[3];
    z[3] = z[2];
    z[2] = z[1];
    z[1] = z[0];
    z[0] = n;
    return n;
}

__device__ __forceinline__ void d_apply(const uint32_t *__restrict__ m, uint32_t z[5])
{
    uint32_t o[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) s += d_red((uint64_t)m[i * 5 + k] * z[k]);
        o[i] = d_red(s);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) z[i] = o[i];
}

__device__ __forceinline__ uint64_t d_scramble(uint64_t v, int lgN, uint64_t val0, uint64_t val1)
{
    v += val0 + val1;
    v *= (val0 | 0x4519840211493211ull);
    v = __brevll(v) >> (64 - lgN);
    v *= (val1 | 0x3050852102C843A5ull);
    v = __brevll(v) >> (64 - lgN);
    return v;
}

// One Kronecker edge (make_one_edge, graph_generator.c:122-151 with the
// initiator draw of :45-80).
__device__ __forceinline__ void d_edge(int64_t ei, const uint32_t *__restrict__ skip,
                                       const KronSetup &ks, int lgN, uint64_t &src, uint64_t &dst)
{
    uint32_t z[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) z[i] = ks.base[i];
    uint64_t e = (uint64_t)ei;
    for (int k = 0; e; ++k, e >>= 8)
        if (e & 0xFF) d_apply(skip + (k * 256 + (e & 0xFF)) * 25, z);
    uint64_t bs = 0, bt = 0;
    const uint32_t limit = 0xFFFFFFFFu % 10000u;
    for (int lev = lgN - 1; lev >= 0; --lev) {
        uint32_t val = d_step(z);
        while (val < limit) val = d_step(z);
        val %= 10000u;
        int sq;
        if (val < 1900u) sq = 1;
        else if (val < 3800u) sq = 2;
        else if (val < 9500u) sq = 0;
        else sq = 3;
        uint64_t so = (uint64_t)(sq >> 1), to = (uint64_t)(sq & 1);
        if (bs == bt && so > to) {  // clip-and-flip
            so = 0;
            to = 1;
        }
        bs += so << lev;
        bt += to << lev;
    }
    src = d_scramble(bs, lgN, ks.val0, ks.val1);
    dst = d_scramble(bt, lgN, ks.val0, ks.val1);
}

__global__ void k_gen_edges(int64_t start, int64_t end, const uint32_t *skip, KronSetup ks, int lgN,
                            int64_t *out)
{
    for (int64_t i = start + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < end;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t a, b;
        d_edge(i, skip, ks, lgN, a, b);
        out[2 * (i - start)] = (int64_t)a;
        out[2 * (i - start) + 1] = (int64_t)b;
    }
}

// Both directed copies of every edge as (row << sh | col) keys; self-loops
// are kept here (they count in the generated degree) and dropped later.
__global__ void k_gen_keys(int64_t m, const uint32_t *skip, KronSetup ks, int lgN, int sh,
                           uint64_t *keys, unsigned long long *d_maxid)
{
    uint64_t mx = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t a, b;
        d_edge(i, skip, ks, lgN, a, b);
        keys[2 * i] = (a << sh) | b;
        keys[2 * i + 1] = (b << sh) | a;
        mx = max(mx, max(a, b));
    }
    mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, 32, 64));
#pragma unroll
    for (int d = 16; d >= 1; d >>= 1) mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(d_maxid, (unsigned long long)mx);
}

// pos[r] = first index i with row(keys[i]) >= r, for r in [0, n_rows].
__global__ void k_row_bounds(const uint64_t *keys, int64_t n, int sh, int64_t n_rows, int64_t *pos)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = (i < n) ? (int64_t)(keys[i] >> sh) : n_rows;
        const int64_t rp = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> sh);
        for (int64_t rr = rp + 1; rr <= r; ++rr) pos[rr] = i;
    }
}

__global__ void k_gdeg(const int64_t *pos, int64_t n_rows, uint32_t *gdeg)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x)
        gdeg[r] = (uint32_t)(pos[r + 1] - pos[r]);
}

__global__ void k_key_cols(const uint64_t *keys, int64_t n, uint64_t mask, uint32_t *cols)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        cols[i] = (uint32_t)(keys[i] & mask);
}

struct NotSelfLoop {
    int sh;
    uint64_t mask;
    __host__ __device__ bool operator()(uint64_t k) const { return (k >> sh) != (k & mask); }
};

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

int build_kronecker(int scale, int edgefactor, uint64_t s1, uint64_t s2, g5_graph *g)
{
    hipStream_t s = nullptr;  // construction runs on the null stream
    int dev = 0;
    G5_HIP(hipGetDevice(&dev));
    int rc = g_skip.init(dev);
    if (rc) return rc;
    const KronSetup ks = kron_setup(s1, s2);
    const int64_t m = (int64_t)edgefactor << scale;
    const int64_t n2 = 2 * m;
    const int sh = scale < 1 ? 1 : scale;
    DevBuf ka, kb, tmp, maxid, cnt;
    G5_HIP(hipMalloc(&ka.p, (size_t)(n2 ? n2 : 1) * 8));
    G5_HIP(hipMalloc(&kb.p, (size_t)(n2 ? n2 : 1) * 8));
    G5_HIP(hipMalloc(&maxid.p, 16));
    G5_HIP(hipMemset(maxid.p, 0, 16));
    uint64_t *A = (uint64_t *)ka.p, *B = (uint64_t *)kb.p;
    if (m > 0)
        hipLaunchKernelGGL(k_gen_keys, g5::grid_for(m, 256, 65536), 256, 0, s, m,
                           (const uint32_t *)g_skip.d, ks, scale, sh, A,
                           (unsigned long long *)maxid.p);
    G5_HIP(hipGetLastError());
    unsigned long long h_max = 0;
    G5_HIP(hipMemcpy(&h_max, maxid.p, 8, hipMemcpyDeviceToHost));
    g->nglobal = m > 0 ? (int64_t)h_max + 1 : 0;
    g->n_rows = g->n_cols = (g->nglobal + 63) / 64 * 64;
    g->n_gen_edges = m;
    g->scale = scale;

    // sort all directed keys by (row, col)
    rocprim::double_buffer<uint64_t> db(A, B);
    size_t tb = 0;
    G5_HIP(rocprim::radix_sort_keys(nullptr, tb, db, (size_t)n2, 0u, (unsigned)(2 * sh), s));
    G5_HIP(hipMalloc(&tmp.p, tb ? tb : 1));
    G5_HIP(rocprim::radix_sort_keys(tmp.p, tb, db, (size_t)n2, 0u, (unsigned)(2 * sh), s));
    uint64_t *sorted = db.current(), *other = db.alternate();

    // generated directed degree from the row ranges of the full sorted list
    G5_HIP(hipMalloc(&g->d_rowptr, (size_t)(g->n_rows + 1) * 8));
    G5_HIP(hipMalloc(&g->d_gdeg, (size_t)(g->n_rows ? g->n_rows : 1) * 4));
    hipLaunchKernelGGL(k_row_bounds, g5::grid_for(n2 + 1, 256, 65536), 256, 0, s,
                       (const uint64_t *)sorted, n2, sh, g->n_rows, g->d_rowptr);
    hipLaunchKernelGGL(k_gdeg, g5::grid_for(g->n_rows, 256, 65536), 256, 0, s,
                       (const int64_t *)g->d_rowptr, g->n_rows, g->d_gdeg);
    G5_HIP(hipGetLastError());

    // unique, then drop self-loops (distmatrix2d.hh:1033-1165)
    G5_HIP(hipMalloc(&cnt.p, 16));
    size_t tu = 0, tsel = 0;
    const uint64_t mask = (sh >= 64) ? ~0ull : ((1ull << sh) - 1);
    G5_HIP(rocprim::unique(nullptr, tu, sorted, other, (size_t *)cnt.p, (size_t)n2,
                           rocprim::equal_to<uint64_t>(), s));
    G5_HIP(rocprim::select(nullptr, tsel, other, sorted, (size_t *)cnt.p + 1, (size_t)n2,
                           NotSelfLoop{sh, mask}, s));
    if (tu > tb || tsel > tb) {
        (void)hipFree(tmp.p);
        tmp.p = nullptr;
        G5_HIP(hipMalloc(&tmp.p, std::max(tu, tsel)));
    }
    G5_HIP(rocprim::unique(tmp.p, tu, sorted, other, (size_t *)cnt.p, (size_t)n2,
                           rocprim::equal_to<uint64_t>(), s));
    size_t h_cnt[2] = {0, 0};
    G5_HIP(hipMemcpy(h_cnt, cnt.p, 8, hipMemcpyDeviceToHost));
    G5_HIP(rocprim::select(tmp.p, tsel, other, sorted, (size_t *)cnt.p + 1, h_cnt[0],
                           NotSelfLoop{sh, mask}, s));
    G5_HIP(hipMemcpy(h_cnt, cnt.p, 16, hipMemcpyDeviceToHost));
    g->nnz = (int64_t)h_cnt[1];
    G5_HIP(hipMalloc(&g->d_cols, (size_t)(g->nnz ? g->nnz : 1) * 4));
    hipLaunchKernelGGL(k_key_cols, g5::grid_for(g->nnz, 256, 65536), 256, 0, s,
                       (const uint64_t *)sorted, g->nnz, mask, g->d_cols);
    hipLaunchKernelGGL(k_row_bounds, g5::grid_for(g->nnz + 1, 256, 65536), 256, 0, s,
                       (const uint64_t *)sorted, g->nnz, sh, g->n_rows, g->d_rowptr);
    G5_HIP(hipGetLastError());
    G5_HIP(hipDeviceSynchronize());
    return G5_OK;
}

// ---- the R x C blocks of the same graph ----

// Bounds of part i of `parts` over n_pad vertices in 64-vertex units
// (DistMatrix2d::setupMatrix2, cpu_2d/distmatrix2d.hh:811-815, with ALG = 64).
void part_range(int64_t n_pad, int parts, int i, int64_t *start, int64_t *len)
{
    const int64_t units = (n_pad + 63) / 64;
    *start = ((int64_t)i * (units / parts) + std::min<int64_t>(i, units % parts)) * 64;
    *len = (units / parts + (i < units % parts ? 1 : 0)) * 64;
}

// first index in [b, e) whose column (key & mask) is >= c (columns ascend in a row)
__device__ __forceinline__ int64_t col_lower_bound(const uint64_t *__restrict__ keys, int64_t b, int64_t e,
                                                   uint64_t mask, uint64_t c)
{
    while (b < e) {
        const int64_t mid = (b + e) >> 1;
        if ((keys[mid] & mask) < c) b = mid + 1;
        else e = mid;
    }
    return b;
}

// Per local row i: the entries of global row rs + i with a column in
// [cs, ce).  Counts go to cnt64 (nullable) and/or cnt32 (nullable).
__global__ void k_block_counts(const uint64_t *__restrict__ keys, const int64_t *__restrict__ pos, int64_t rs,
                               int64_t rl, uint64_t cs, uint64_t ce, uint64_t mask, int64_t *__restrict__ cnt64,
                               uint32_t *__restrict__ cnt32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rl;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = pos[rs + i], e = pos[rs + i + 1];
        const int64_t n = col_lower_bound(keys, b, e, mask, ce) - col_lower_bound(keys, b, e, mask, cs);
        if (cnt64) cnt64[i] = n;
        if (cnt32) cnt32[i] = (uint32_t)n;
    }
}

// one wavefront per local row: copy the row's in-block columns as local ids
__global__ __launch_bounds__(256) void k_block_fill(const uint64_t *__restrict__ keys, const int64_t *__restrict__ pos,
                                                    int64_t rs, int64_t rl, uint64_t cs, uint64_t ce, uint64_t mask,
                                                    const int64_t *__restrict__ rowptr, uint32_t *__restrict__ cols)
{
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < rl; i += waves) {
        const int64_t b = pos[rs + i], e = pos[rs + i + 1];
        const int64_t lo = col_lower_bound(keys, b, e, mask, cs);
        const int64_t n = rowptr[i + 1] - rowptr[i];
        for (int64_t k = __lane_id(); k < n; k += 64) cols[rowptr[i] + k] = (uint32_t)((keys[lo + k] & mask) - cs);
    }
}

// All R x C blocks of the Kronecker graph, as the R x C ranks of the
// reference hold them: block (r, c) keeps the symmetrised entries with row in
// row part r and column in column part c, deduplicated without self-loops
// (distmatrix2d.hh:1033-1165), local column ids (WOLO); gdeg counts the
// block's generated entries (duplicates and self-loops included), so a row's
// generated degree is the sum over its row of blocks.  The vertex count is
// padded to a multiple of R (the CUDA path's PAD, cuda/cuda_bfs.h:40).
int build_kronecker_2d(int scale, int edgefactor, uint64_t s1, uint64_t s2, int R, int C, g5_graph **blocks)
{
    hipStream_t s = nullptr;
    int dev = 0;
    G5_HIP(hipGetDevice(&dev));
    int rc = g_skip.init(dev);
    if (rc) return rc;
    const KronSetup ks = kron_setup(s1, s2);
    const int64_t m = (int64_t)edgefactor << scale;
    const int64_t n2 = 2 * m;
    const int sh = scale < 1 ? 1 : scale;
    const uint64_t mask = (1ull << sh) - 1;
    DevBuf ka, kb, tmp, maxid, cnt, pos, deg;
    G5_HIP(hipMalloc(&ka.p, (size_t)n2 * 8));
    G5_HIP(hipMalloc(&kb.p, (size_t)n2 * 8));
    G5_HIP(hipMalloc(&maxid.p, 16));
    G5_HIP(hipMemset(maxid.p, 0, 16));
    uint64_t *A = (uint64_t *)ka.p, *B = (uint64_t *)kb.p;
    hipLaunchKernelGGL(k_gen_keys, g5::grid_for(m, 256, 65536), 256, 0, s, m, (const uint32_t *)g_skip.d, ks, scale,
                       sh, A, (unsigned long long *)maxid.p);
    G5_HIP(hipGetLastError());
    unsigned long long h_max = 0;
    G5_HIP(hipMemcpy(&h_max, maxid.p, 8, hipMemcpyDeviceToHost));
    const int64_t nglobal = (int64_t)h_max + 1;
    const int64_t n_pad = nglobal + (nglobal % R ? R - nglobal % R : 0);
    int64_t rs[64], rl[64], cs[64], cl[64];
    for (int r = 0; r < R; ++r) part_range(n_pad, R, r, &rs[r], &rl[r]);
    for (int c = 0; c < C; ++c) part_range(n_pad, C, c, &cs[c], &cl[c]);
    const int64_t rows_total = rs[R - 1] + rl[R - 1];

    rocprim::double_buffer<uint64_t> db(A, B);
    size_t tb = 0;
    G5_HIP(rocprim::radix_sort_keys(nullptr, tb, db, (size_t)n2, 0u, (unsigned)(2 * sh), s));
    G5_HIP(hipMalloc(&tmp.p, tb ? tb : 1));
    G5_HIP(rocprim::radix_sort_keys(tmp.p, tb, db, (size_t)n2, 0u, (unsigned)(2 * sh), s));
    uint64_t *sorted = db.current(), *other = db.alternate();
    G5_HIP(hipMalloc(&pos.p, (size_t)(rows_total + 1) * 8));
    int64_t *d_pos = (int64_t *)pos.p;

    for (int i = 0; i < R * C; ++i) {
        const int r = i / C, c = i % C;
        g5_graph *g = new g5_graph();
        blocks[i] = g;
        g->device = dev;
        g->scale = scale;
        g->n_rows = rl[r];
        g->n_cols = cl[c];
        g->row_start = rs[r];
        g->col_start = cs[c];
        g->nglobal = nglobal;
        g->n_gen_edges = m;
        G5_HIP(hipMalloc(&g->d_rowptr, (size_t)(g->n_rows + 1) * 8));
        G5_HIP(hipMalloc(&g->d_gdeg, (size_t)std::max<int64_t>(g->n_rows, 1) * 4));
    }
    // generated degrees per block from the full list
    hipLaunchKernelGGL(k_row_bounds, g5::grid_for(n2 + 1, 256, 65536), 256, 0, s, (const uint64_t *)sorted, n2, sh,
                       rows_total, d_pos);
    for (int i = 0; i < R * C; ++i) {
        g5_graph *g = blocks[i];
        hipLaunchKernelGGL(k_block_counts, g5::grid_for(g->n_rows, 256, 65536), 256, 0, s, (const uint64_t *)sorted,
                           (const int64_t *)d_pos, g->row_start, g->n_rows, (uint64_t)g->col_start,
                           (uint64_t)(g->col_start + g->n_cols), mask, (int64_t *)nullptr, g->d_gdeg);
    }
    G5_HIP(hipGetLastError());

    // unique, then drop self-loops
    G5_HIP(hipMalloc(&cnt.p, 16));
    size_t tu = 0, tsel = 0;
    G5_HIP(rocprim::unique(nullptr, tu, sorted, other, (size_t *)cnt.p, (size_t)n2, rocprim::equal_to<uint64_t>(), s));
    G5_HIP(rocprim::select(nullptr, tsel, other, sorted, (size_t *)cnt.p + 1, (size_t)n2, NotSelfLoop{sh, mask}, s));
    if (tu > tb || tsel > tb) {
        (void)hipFree(tmp.p);
        tmp.p = nullptr;
        G5_HIP(hipMalloc(&tmp.p, std::max(tu, tsel)));
    }
    G5_HIP(rocprim::unique(tmp.p, tu, sorted, other, (size_t *)cnt.p, (size_t)n2, rocprim::equal_to<uint64_t>(), s));
    size_t h_cnt[2] = {0, 0};
    G5_HIP(hipMemcpy(h_cnt, cnt.p, 8, hipMemcpyDeviceToHost));
    G5_HIP(rocprim::select(tmp.p, tsel, other, sorted, (size_t *)cnt.p + 1, h_cnt[0], NotSelfLoop{sh, mask}, s));
    G5_HIP(hipMemcpy(h_cnt, cnt.p, 16, hipMemcpyDeviceToHost));
    const int64_t nnz_all = (int64_t)h_cnt[1];
    hipLaunchKernelGGL(k_row_bounds, g5::grid_for(nnz_all + 1, 256, 65536), 256, 0, s, (const uint64_t *)sorted,
                       nnz_all, sh, rows_total, d_pos);
    G5_HIP(hipGetLastError());

    // per block: row lengths -> row offsets -> columns
    int64_t max_rows = 0;
    for (int r = 0; r < R; ++r) max_rows = std::max(max_rows, rl[r]);
    G5_HIP(hipMalloc(&deg.p, (size_t)(max_rows + 1) * 8));
    DevBuf scan;
    G5_HIP(hipMalloc(&scan.p, (g5::scan_scratch_words(max_rows + 1) + 16) * 8));
    for (int i = 0; i < R * C; ++i) {
        g5_graph *g = blocks[i];
        hipLaunchKernelGGL(k_block_counts, g5::grid_for(g->n_rows, 256, 65536), 256, 0, s, (const uint64_t *)sorted,
                           (const int64_t *)d_pos, g->row_start, g->n_rows, (uint64_t)g->col_start,
                           (uint64_t)(g->col_start + g->n_cols), mask, (int64_t *)deg.p, (uint32_t *)nullptr);
        G5_HIP(hipGetLastError());
        G5_HIP(g5::exclusive_scan_i64((int64_t *)deg.p, g->n_rows, (int64_t *)deg.p + g->n_rows, (int64_t *)scan.p, s));
        G5_HIP(hipMemcpyAsync(g->d_rowptr, deg.p, (size_t)(g->n_rows + 1) * 8, hipMemcpyDeviceToDevice, s));
        int64_t nnz = 0;
        G5_HIP(hipMemcpy(&nnz, (int64_t *)deg.p + g->n_rows, 8, hipMemcpyDeviceToHost));
        g->nnz = nnz;
        G5_HIP(hipMalloc(&g->d_cols, (size_t)std::max<int64_t>(nnz, 1) * 4));
        hipLaunchKernelGGL(k_block_fill, g5::grid_for((g->n_rows + 3) / 4, 1, 65536), 256, 0, s,
                           (const uint64_t *)sorted, (const int64_t *)d_pos, g->row_start, g->n_rows,
                           (uint64_t)g->col_start, (uint64_t)(g->col_start + g->n_cols), mask,
                           (const int64_t *)g->d_rowptr, g->d_cols);
        G5_HIP(hipGetLastError());
    }
    G5_HIP(hipDeviceSynchronize());
    return G5_OK;
}

}  // namespace

namespace g5 {

hipError_t sort_u64(uint64_t *keys, uint64_t *tmp, int64_t n, int end_bit, hipStream_t s)
{
    if (n <= 1) return hipSuccess;
    rocprim::double_buffer<uint64_t> db(keys, tmp);
    size_t tb = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, tb, db, (size_t)n, 0u, (unsigned)end_bit, s);
    if (e != hipSuccess) return e;
    void *t = nullptr;
    e = hipMallocAsync(&t, tb ? tb : 1, s);
    if (e != hipSuccess) return e;
    e = rocprim::radix_sort_keys(t, tb, db, (size_t)n, 0u, (unsigned)end_bit, s);
    hipError_t e2 = hipFreeAsync(t, s);
    if (e != hipSuccess) return e;
    if (e2 != hipSuccess) return e2;
    if (db.current() != keys) return hipMemcpyAsync(keys, db.current(), (size_t)n * 8, hipMemcpyDeviceToDevice, s);
    return hipSuccess;
}

}  // namespace g5

extern "C" int g5_graph_upload(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows,
                               int64_t n_cols, const int64_t *gdeg, int64_t row_start,
                               int64_t col_start, g5_graph **out)
{
    G5_REQUIRE(rowptr && out && n_rows >= 0 && n_cols >= 0, G5_E_ARG, "g5_graph_upload: bad arguments");
    G5_REQUIRE(n_rows % 64 == 0 && n_cols % 64 == 0, G5_E_ARG,
               "g5_graph_upload: n_rows and n_cols must be multiples of 64");
    const int64_t nnz = rowptr[n_rows];
    G5_REQUIRE(nnz >= 0 && (nnz == 0 || cols), G5_E_ARG, "g5_graph_upload: bad nnz/cols");
    for (int64_t i = 0; i < n_rows; ++i)
        G5_REQUIRE(rowptr[i] <= rowptr[i + 1], G5_E_ARG, "g5_graph_upload: rowptr not monotone at %lld", (long long)i);
    for (int64_t j = 0; j < nnz; ++j)
        G5_REQUIRE(cols[j] < (uint64_t)n_cols, G5_E_ARG, "g5_graph_upload: column id out of range at %lld", (long long)j);
    g5_graph *g = new g5_graph();
    G5_HIP(hipGetDevice(&g->device));
    g->n_rows = n_rows;
    g->n_cols = n_cols;
    g->nnz = nnz;
    g->nglobal = std::max(n_rows + row_start, n_cols + col_start);
    g->row_start = row_start;
    g->col_start = col_start;
    int rc = G5_OK;
    auto fail = [&](int r) {
        g5_graph_free(g);
        return r;
    };
    hipError_t e;
    if ((e = hipMalloc(&g->d_rowptr, (size_t)(n_rows + 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&g->d_cols, (size_t)(nnz ? nnz : 1) * 4)) != hipSuccess ||
        (e = hipMalloc(&g->d_gdeg, (size_t)(n_rows ? n_rows : 1) * 4)) != hipSuccess)
        return fail(g5::fail_hip(e, "hipMalloc", __FILE__, __LINE__));
    if ((e = hipMemcpy(g->d_rowptr, rowptr, (size_t)(n_rows + 1) * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (nnz && (e = hipMemcpy(g->d_cols, cols, (size_t)nnz * 4, hipMemcpyHostToDevice)) != hipSuccess))
        return fail(g5::fail_hip(e, "hipMemcpy", __FILE__, __LINE__));
    std::vector<uint32_t> gd(n_rows);
    int64_t gen2 = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        const int64_t v = gdeg ? gdeg[i] : rowptr[i + 1] - rowptr[i];
        gd[i] = (uint32_t)v;
        gen2 += v;
    }
    g->n_gen_edges = gen2 / 2;
    if (n_rows && (e = hipMemcpy(g->d_gdeg, gd.data(), (size_t)n_rows * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(g5::fail_hip(e, "hipMemcpy", __FILE__, __LINE__));
    *out = g;
    return rc;
}

extern "C" int g5_graph_kronecker(int scale, int edgefactor, uint64_t seed1, uint64_t seed2,
                                  g5_graph **out)
{
    G5_REQUIRE(out && scale >= 1 && scale <= 30 && edgefactor >= 1 && edgefactor <= 1024, G5_E_ARG,
               "g5_graph_kronecker: scale must be in [1,30] and edgefactor in [1,1024]");
    g5_graph *g = new g5_graph();
    G5_HIP(hipGetDevice(&g->device));
    int rc = build_kronecker(scale, edgefactor, seed1, seed2, g);
    if (rc) {
        g5_graph_free(g);
        return rc;
    }
    *out = g;
    return G5_OK;
}

extern "C" int g5_block_bounds(int64_t nglobal, int R, int C, int r, int c, int64_t *row_start, int64_t *n_rows,
                               int64_t *col_start, int64_t *n_cols)
{
    G5_REQUIRE(nglobal >= 0 && R >= 1 && C >= 1 && R <= 64 && C <= 64 && r >= 0 && r < R && c >= 0 && c < C &&
                   row_start && n_rows && col_start && n_cols,
               G5_E_ARG, "g5_block_bounds: bad arguments");
    const int64_t n_pad = nglobal + (nglobal % R ? R - nglobal % R : 0);
    part_range(n_pad, R, r, row_start, n_rows);
    part_range(n_pad, C, c, col_start, n_cols);
    return G5_OK;
}

extern "C" int g5_graph_kronecker_2d(int scale, int edgefactor, uint64_t seed1, uint64_t seed2, int R, int C,
                                     g5_graph **blocks)
{
    G5_REQUIRE(blocks && scale >= 1 && scale <= 30 && edgefactor >= 1 && edgefactor <= 1024 && R >= 1 && C >= 1 &&
                   R <= 64 && C <= 64,
               G5_E_ARG, "g5_graph_kronecker_2d: scale in [1,30], edgefactor in [1,1024], 1 <= R, C <= 64");
    for (int i = 0; i < R * C; ++i) blocks[i] = nullptr;
    int rc = build_kronecker_2d(scale, edgefactor, seed1, seed2, R, C, blocks);
    if (rc) {
        for (int i = 0; i < R * C; ++i) {
            g5_graph_free(blocks[i]);
            blocks[i] = nullptr;
        }
    }
    return rc;
}

extern "C" int g5_graph_gdeg(const g5_graph *g, int64_t *gdeg)
{
    G5_REQUIRE(g && gdeg, G5_E_ARG, "g5_graph_gdeg: null argument");
    std::vector<uint32_t> t((size_t)g->n_rows);
    if (g->n_rows) G5_HIP(hipMemcpy(t.data(), g->d_gdeg, (size_t)g->n_rows * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < g->n_rows; ++i) gdeg[i] = t[i];
    return G5_OK;
}

namespace {
// hist[0] = rows without neighbours; hist[k] = rows with 2^(k-1) <= degree
// < 2^k, the last bin open-ended.  Per-workgroup counts in LDS, then one
// atomic per non-empty bin.
__global__ __launch_bounds__(256) void k_degree_hist(const int64_t *__restrict__ rowptr, int64_t n, int nbins,
                                                     unsigned long long *__restrict__ hist)
{
    __shared__ unsigned long long s_h[64];
    if (threadIdx.x < 64) s_h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = rowptr[v + 1] - rowptr[v];
        const int k = d == 0 ? 0 : min(nbins - 1, 64 - __clzll((long long)d));
        atomicAdd(&s_h[k], 1ull);
    }
    __syncthreads();
    if (threadIdx.x < nbins && s_h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_h[threadIdx.x]);
}
}  // namespace

extern "C" int g5_graph_degree_hist(const g5_graph *g, int64_t *hist, int nbins)
{
    G5_REQUIRE(g && hist && nbins >= 2 && nbins <= 64, G5_E_ARG, "g5_graph_degree_hist: bad arguments");
    unsigned long long *d = nullptr;
    G5_HIP(hipMalloc(&d, 64 * 8));
    hipError_t e = g5::fill_u64((uint64_t *)d, 64, 0, nullptr);
    if (e == hipSuccess && g->n_rows) {
        hipLaunchKernelGGL(k_degree_hist, g5::grid_for(g->n_rows, 256, 4096), 256, 0, nullptr,
                           (const int64_t *)g->d_rowptr, g->n_rows, nbins, d);
        e = hipGetLastError();
    }
    unsigned long long h[64];
    if (e == hipSuccess) e = hipMemcpy(h, d, 64 * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    G5_HIP(e);
    for (int k = 0; k < nbins; ++k) hist[k] = (int64_t)h[k];
    return G5_OK;
}

extern "C" int g5_graph_bounds(const g5_graph *g, int64_t *row_start, int64_t *col_start)
{
    G5_REQUIRE(g && row_start && col_start, G5_E_ARG, "g5_graph_bounds: null argument");
    *row_start = g->row_start;
    *col_start = g->col_start;
    return G5_OK;
}

extern "C" int g5_graph_info(const g5_graph *g, int64_t *n_rows, int64_t *n_cols, int64_t *nnz,
                             int64_t *nglobal, int64_t *n_gen_edges)
{
    G5_REQUIRE(g, G5_E_ARG, "g5_graph_info: null graph");
    if (n_rows) *n_rows = g->n_rows;
    if (n_cols) *n_cols = g->n_cols;
    if (nnz) *nnz = g->nnz;
    if (nglobal) *nglobal = g->nglobal;
    if (n_gen_edges) *n_gen_edges = g->n_gen_edges;
    return G5_OK;
}

extern "C" int g5_graph_download(const g5_graph *g, int64_t *rowptr, uint32_t *cols, int64_t *gdeg)
{
    G5_REQUIRE(g, G5_E_ARG, "g5_graph_download: null graph");
    if (rowptr) G5_HIP(hipMemcpy(rowptr, g->d_rowptr, (size_t)(g->n_rows + 1) * 8, hipMemcpyDeviceToHost));
    if (cols && g->nnz) G5_HIP(hipMemcpy(cols, g->d_cols, (size_t)g->nnz * 4, hipMemcpyDeviceToHost));
    if (gdeg && g->n_rows) {
        std::vector<uint32_t> gd(g->n_rows);
        G5_HIP(hipMemcpy(gd.data(), g->d_gdeg, (size_t)g->n_rows * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < g->n_rows; ++i) gdeg[i] = gd[i];
    }
    return G5_OK;
}

extern "C" int g5_graph_free(g5_graph *g)
{
    if (!g) return G5_OK;
    if (g->d_rowptr) (void)hipFree(g->d_rowptr);
    if (g->d_cols) (void)hipFree(g->d_cols);
    if (g->d_gdeg) (void)hipFree(g->d_gdeg);
    delete g;
    return G5_OK;
}

extern "C" int g5_kronecker_edges(int scale, uint64_t seed1, uint64_t seed2, int64_t start,
                                  int64_t end, int64_t *edges)
{
    G5_REQUIRE(scale >= 1 && scale <= 32 && start >= 0 && end >= start && (edges || end == start),
               G5_E_ARG, "g5_kronecker_edges: bad arguments");
    if (end == start) return G5_OK;
    int dev = 0;
    G5_HIP(hipGetDevice(&dev));
    int rc = g_skip.init(dev);
    if (rc) return rc;
    const KronSetup ks = kron_setup(seed1, seed2);
    DevBuf buf;
    const int64_t n = end - start;
    G5_HIP(hipMalloc(&buf.p, (size_t)n * 16));
    hipLaunchKernelGGL(k_gen_edges, g5::grid_for(n, 256, 65536), 256, 0, nullptr, start, end,
                       (const uint32_t *)g_skip.d, ks, scale, (int64_t *)buf.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpy(edges, buf.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    return G5_OK;
}

namespace {

__global__ void k_cand_degrees(const int64_t *__restrict__ rowptr, int64_t row_start, int64_t n_rows,
                               const int64_t *__restrict__ cand, int n, int64_t *__restrict__ deg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = cand[i] - row_start;
    deg[i] = (r >= 0 && r < n_rows) ? rowptr[r + 1] - rowptr[r] : 0;
}

}  // namespace

// Root selection of cpu_2d/main.cpp:351-411: draws from a default-seeded
// std::knuth_b through uniform_int_distribution<int64>(0, vertices-1); a draw
// is kept if it was not drawn before and its CSR degree is nonzero; at most
// 32 draws per requested root.  The draw sequence does not depend on which
// draws are kept, so all candidate degrees are fetched in one launch.
extern "C" int g5_select_roots(const g5_graph *g, int64_t vertices, int nroots, int64_t *roots, int *found)
{
    G5_REQUIRE(g && roots && found && nroots >= 0 && vertices >= 1, G5_E_ARG, "g5_select_roots: bad arguments");
    const int maxtries = nroots * 32;
    std::knuth_b gen;
    std::uniform_int_distribution<int64_t> dist(0, vertices - 1);
    std::vector<int64_t> cand(maxtries), deg(maxtries);
    for (int i = 0; i < maxtries; ++i) cand[i] = dist(gen);
    *found = 0;
    if (maxtries == 0) return G5_OK;
    DevBuf dc, dd;
    G5_HIP(hipMalloc(&dc.p, (size_t)maxtries * 8));
    G5_HIP(hipMalloc(&dd.p, (size_t)maxtries * 8));
    G5_HIP(hipMemcpy(dc.p, cand.data(), (size_t)maxtries * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cand_degrees, (maxtries + 255) / 256, 256, 0, nullptr, (const int64_t *)g->d_rowptr,
                       g->row_start, g->n_rows, (const int64_t *)dc.p, maxtries, (int64_t *)dd.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpy(deg.data(), dd.p, (size_t)maxtries * 8, hipMemcpyDeviceToHost));
    int n = 0;
    for (int t = 0; t < maxtries && n < nroots; ++t) {
        bool dup = false;
        for (int i = 0; i < n; ++i) dup |= roots[i] == cand[t];
        if (dup) continue;
        if (deg[t] > 0) roots[n++] = cand[t];
    }
    *found = n;
    return G5_OK;
}
