/*
 * g5oracle.c -- CPU oracle (restatement of the reference algorithms).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked into, or called
 * by, the product library.  See g5oracle.h for what pins each function.
 *
 * Plain C99, single-threaded, no SIMD: the point is to be obviously the same
 * algorithm as the reference, not to be fast.
 */
#include "g5oracle.h"

#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* Graph500 Kronecker generator                                             */
/* ======================================================================== */

#define MRG_P 0x7FFFFFFFu /* modulus 2^31 - 1 */
#define MRG_X 107374182u  /* cpu_2d/generator/splittable_mrg.c:33 */
#define MRG_Y 104480u

typedef struct { uint32_t m[5][5]; } mat5;

static uint32_t mulmod(uint32_t a, uint32_t b) { return (uint32_t)((uint64_t)a * b % MRG_P); }

static mat5 mat_mul(const mat5 *a, const mat5 *b)
{
    mat5 r;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            uint64_t s = 0;
            for (int k = 0; k < 5; ++k) s += mulmod(a->m[i][k], b->m[k][j]);
            r.m[i][j] = (uint32_t)(s % MRG_P);
        }
    return r;
}

static void mat_apply(const mat5 *a, uint32_t z[5])
{
    uint32_t o[5];
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
        for (int k = 0; k < 5; ++k) s += mulmod(a->m[i][k], z[k]);
        o[i] = (uint32_t)(s % MRG_P);
    }
    memcpy(z, o, sizeof(o));
}

/* skip[b][v] = A^(v * 2^(8b)), A the one-step MRG matrix
 * (cpu_2d/generator/splittable_mrg.c:25-57, 105-125).  The reference reads
 * these from a precomputed table; here they are computed by squaring, which
 * fixes the same matrices because they are exact powers mod 2^31-1. */
static mat5 *g_skip = NULL;

static void init_skip_tables(void)
{
    if (g_skip) return;
    mat5 *t = (mat5 *)malloc(sizeof(mat5) * 24 * 256);
    mat5 a;
    memset(&a, 0, sizeof(a));
    a.m[0][0] = MRG_X;
    a.m[0][4] = MRG_Y;
    for (int i = 1; i < 5; ++i) a.m[i][i - 1] = 1;
    mat5 id;
    memset(&id, 0, sizeof(id));
    for (int i = 0; i < 5; ++i) id.m[i][i] = 1;
    mat5 p = a; /* A^(2^(8b)) */
    for (int b = 0; b < 24; ++b) {
        t[b * 256] = id;
        for (int v = 1; v < 256; ++v) t[b * 256 + v] = mat_mul(&t[b * 256 + v - 1], &p);
        for (int s = 0; s < 8; ++s) p = mat_mul(&p, &p);
    }
    g_skip = t;
}

/* splittable_mrg.c:105-125 */
static void mrg_skip(uint32_t z[5], uint64_t hi, uint64_t mid, uint64_t lo)
{
    init_skip_tables();
    for (int b = 0; lo; ++b, lo >>= 8)
        if (lo & 0xFF) mat_apply(&g_skip[b * 256 + (lo & 0xFF)], z);
    for (int b = 8; mid; ++b, mid >>= 8)
        if (mid & 0xFF) mat_apply(&g_skip[b * 256 + (mid & 0xFF)], z);
    for (int b = 16; hi; ++b, hi >>= 8)
        if (hi & 0xFF) mat_apply(&g_skip[b * 256 + (hi & 0xFF)], z);
}

/* splittable_mrg.c:92-100,129-133: one step of the original recurrence */
static uint32_t mrg_get_uint_orig(uint32_t z[5])
{
    uint32_t n = (uint32_t)(((uint64_t)mulmod(z[0], MRG_X) + mulmod(z[4], MRG_Y)) % MRG_P);
    z[4] = z[3];
    z[3] = z[2];
    z[2] = z[1];
    z[1] = z[0];
    z[0] = n;
    return n;
}

/* cpu_2d/generator/utils.c:61-68 */
void orc_make_mrg_seed(uint64_t s1, uint64_t s2, uint32_t seed[5])
{
    seed[0] = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
    seed[1] = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
    seed[2] = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
    seed[3] = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
    seed[4] = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
}

/* graph_generator.c:45-80 (SPK_NOISE_LEVEL == 0): initiator 0.57/0.19/0.19/0.05 */
static int bernoulli4(uint32_t z[5])
{
    const uint32_t limit = 0xFFFFFFFFu % 10000u;
    uint32_t val = mrg_get_uint_orig(z);
    while (val < limit) val = mrg_get_uint_orig(z);
    val %= 10000u;
    if (val < 1900u) return 1;
    val -= 1900u;
    if (val < 1900u) return 2;
    val -= 1900u;
    if (val < 5700u) return 0;
    return 3;
}

static uint64_t bitrev64(uint64_t x)
{
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i) r |= ((x >> i) & 1u) << (63 - i);
    return r;
}

/* graph_generator.c:108-119 */
static int64_t scramble(int64_t v0, int lgN, uint64_t val0, uint64_t val1)
{
    uint64_t v = (uint64_t)v0;
    v += val0 + val1;
    v *= (val0 | UINT64_C(0x4519840211493211));
    v = bitrev64(v) >> (64 - lgN);
    v *= (val1 | UINT64_C(0x3050852102C843A5));
    v = bitrev64(v) >> (64 - lgN);
    return (int64_t)v;
}

/* graph_generator.c:157-187 with make_one_edge at :122-151 */
void orc_generate_kronecker_range(const uint32_t seed[5], int logN, int64_t start_edge,
                                  int64_t end_edge, int64_t *edges)
{
    uint32_t base[5];
    memcpy(base, seed, sizeof(base));
    uint64_t val0, val1;
    {
        uint32_t z[5];
        memcpy(z, base, sizeof(z));
        mrg_skip(z, 50, 7, 0);
        val0 = mrg_get_uint_orig(z);
        val0 *= UINT64_C(0xFFFFFFFF);
        val0 += mrg_get_uint_orig(z);
        val1 = mrg_get_uint_orig(z);
        val1 *= UINT64_C(0xFFFFFFFF);
        val1 += mrg_get_uint_orig(z);
    }
    for (int64_t ei = start_edge; ei < end_edge; ++ei) {
        uint32_t z[5];
        memcpy(z, base, sizeof(z));
        mrg_skip(z, 0, (uint64_t)ei, 0);
        int64_t nverts = (int64_t)1 << logN;
        int64_t bs = 0, bt = 0;
        while (nverts > 1) {
            int sq = bernoulli4(z);
            int so = sq / 2, to = sq % 2;
            if (bs == bt && so > to) { /* clip-and-flip */
                int t = so;
                so = to;
                to = t;
            }
            nverts /= 2;
            bs += nverts * so;
            bt += nverts * to;
        }
        edges[2 * (ei - start_edge)] = scramble(bs, logN, val0, val1);
        edges[2 * (ei - start_edge) + 1] = scramble(bt, logN, val0, val1);
    }
}

/* ======================================================================== */
/* DistMatrix2d at 1x1                                                      */
/* ======================================================================== */

/* distmatrix2d.hh:760-802: global vertex count = max id + 1 */
int64_t orc_num_vertices(const int64_t *edges, int64_t m)
{
    int64_t mx = -1;
    for (int64_t i = 0; i < 2 * m; ++i)
        if (edges[i] > mx) mx = edges[i];
    return mx + 1;
}

/* distmatrix2d.hh:811-817 with R = C = 1, ALG = 64 */
int64_t orc_padded_rows(int64_t nglobal) { return (nglobal + 63) / 64 * 64; }

static void radix_sort_u64(uint64_t *a, uint64_t *tmp, int64_t n, int bits)
{
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * 65536);
    for (int sh = 0; sh < bits; sh += 16) {
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (int64_t i = 0; i < n; ++i) cnt[(a[i] >> sh) & 0xFFFF]++;
        size_t s = 0;
        for (int d = 0; d < 65536; ++d) {
            size_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (int64_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
        memcpy(a, tmp, sizeof(uint64_t) * (size_t)n);
    }
    free(cnt);
}

/* distmatrix2d.hh:745-775 (symmetrise), :981-985 (sort by (v0,v1)),
 * :1033-1165 (per-row dedup, self-loops dropped, WOLO local column ids). */
int64_t orc_build_csr(const int64_t *edges, int64_t m, int64_t n_rows, int64_t *rowptr,
                      uint32_t *cols, int64_t *gdeg)
{
    int sh = 1;
    while (((int64_t)1 << sh) < n_rows) ++sh;
    const int64_t n2 = 2 * m;
    uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n2 ? n2 : 1));
    uint64_t *t = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n2 ? n2 : 1));
    if (gdeg) memset(gdeg, 0, sizeof(int64_t) * (size_t)n_rows);
    for (int64_t i = 0; i < m; ++i) {
        uint64_t a = (uint64_t)edges[2 * i], b = (uint64_t)edges[2 * i + 1];
        k[2 * i] = (a << sh) | b;
        k[2 * i + 1] = (b << sh) | a;
        if (gdeg) {
            gdeg[a]++;
            gdeg[b]++;
        }
    }
    radix_sort_u64(k, t, n2, 2 * sh);
    free(t);
    const uint64_t lo = ((uint64_t)1 << sh) - 1;
    int64_t nnz = 0, j = 0;
    rowptr[0] = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        int64_t last_valid = -1;
        while (j < n2 && (int64_t)(k[j] >> sh) == i) {
            int64_t v1 = (int64_t)(k[j] & lo);
            ++j;
            if (v1 != i) {
                last_valid = v1;
                cols[nnz++] = (uint32_t)v1;
                break;
            }
        }
        while (j < n2 && (int64_t)(k[j] >> sh) == i) {
            int64_t v1 = (int64_t)(k[j] & lo);
            if (v1 != i && v1 != last_valid) cols[nnz++] = (uint32_t)v1;
            last_valid = v1;
            ++j;
        }
        rowptr[i + 1] = nnz;
    }
    free(k);
    return nnz;
}

/* ======================================================================== */
/* BFS                                                                      */
/* ======================================================================== */

/* CPUBFS_bin inside the GlobalBFS loop at a 1x1 grid:
 *   setStartVertex     cpu_2d/cpubfs_bin.cpp:116-141
 *   runLocalBFS        cpu_2d/cpubfs_bin.cpp:143-180 (ascending word/bit scan,
 *                      first discoverer wins => minimum-id parent)
 *   istheresomethingnew cpu_2d/cpubfs_bin.cpp:102-114, checked for depth > 0
 *                      (cpu_2d/globalbfs.hh:716-739)
 *   vreduce at one rank = identity; setModOutgoingFQ (visited |= out)
 *                      cpu_2d/cpubfs_bin.cpp:74-87
 *   row fold           setIncommingFQ copies out -> in, cpubfs_bin.cpp:95-100
 * level[] records the loop depth at which each vertex was discovered. */
int orc_bfs_bin(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows, int64_t root,
                int64_t *pred, int32_t *level, int32_t *nlevels)
{
    if (n_rows % 64 != 0 || root < 0 || root >= n_rows) return -1;
    const int64_t w = n_rows / 64;
    uint64_t *fq_in = (uint64_t *)calloc((size_t)w, 8);
    uint64_t *fq_out = (uint64_t *)calloc((size_t)w, 8);
    uint64_t *visited = (uint64_t *)calloc((size_t)w, 8);
    for (int64_t i = 0; i < n_rows; ++i) {
        pred[i] = -1;
        level[i] = -1;
    }
    fq_in[root / 64] = UINT64_C(1) << (root & 63);
    visited[root / 64] = UINT64_C(1) << (root & 63);
    pred[root] = root;
    level[root] = 0;
    int depth = 0;
    for (;;) {
        memset(fq_out, 0, (size_t)w * 8);
        for (int64_t i = 0; i < w; ++i) {
            if (!fq_in[i]) continue;
            for (int ii = 0; ii < 64; ++ii) {
                if (!(fq_in[i] >> ii & 1u)) continue;
                const int64_t u = i * 64 + ii;
                for (int64_t j = rowptr[u]; j < rowptr[u + 1]; ++j) {
                    const uint32_t v = cols[j];
                    const uint64_t bit = UINT64_C(1) << (v & 63);
                    if (!(visited[v >> 6] & bit) && !(fq_out[v >> 6] & bit)) {
                        fq_out[v >> 6] |= bit;
                        pred[v] = u;
                        level[v] = depth + 1;
                    }
                }
            }
        }
        if (depth > 0) {
            int any = 0;
            for (int64_t i = 0; i < w && !any; ++i) any = fq_out[i] != 0;
            if (!any) break;
        }
        for (int64_t i = 0; i < w; ++i) visited[i] |= fq_out[i];
        memcpy(fq_in, fq_out, (size_t)w * 8);
        ++depth;
    }
    /* depth = number of expansions that found something, +1 for the root level */
    int32_t mx = 0;
    for (int64_t i = 0; i < n_rows; ++i)
        if (level[i] > mx) mx = level[i];
    *nlevels = mx + 1;
    free(fq_in);
    free(fq_out);
    free(visited);
    return 0;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* SimpleCPUBFS at 1x1 (cpu_2d/simplecpubfs.cpp:35-58 and its setStartVertex):
 * sorted in-queue popped from the back, so the LAST (maximum-id) frontier
 * neighbour becomes the parent.  Levels equal orc_bfs_bin's; parents do not. */
int orc_bfs_simple(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows, int64_t root,
                   int64_t *pred, int32_t *level, int32_t *nlevels)
{
    if (root < 0 || root >= n_rows) return -1;
    unsigned char *visited = (unsigned char *)calloc((size_t)n_rows, 1);
    int64_t *fq_in = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_rows);
    int64_t *fq_out = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_rows);
    int64_t nin = 0, nout;
    for (int64_t i = 0; i < n_rows; ++i) {
        pred[i] = -1;
        level[i] = -1;
    }
    visited[root] = 1;
    pred[root] = root;
    level[root] = 0;
    fq_in[nin++] = root;
    int depth = 0;
    for (;;) {
        nout = 0;
        while (nin > 0) {
            const int64_t u = fq_in[--nin];
            for (int64_t j = rowptr[u]; j < rowptr[u + 1]; ++j) {
                const int64_t v = cols[j];
                if (!visited[v]) {
                    fq_out[nout++] = v;
                    visited[v] = 1;
                    pred[v] = u;
                    level[v] = depth + 1;
                }
            }
        }
        qsort(fq_out, (size_t)nout, sizeof(int64_t), cmp_i64);
        if (depth > 0 && nout == 0) break;
        memcpy(fq_in, fq_out, sizeof(int64_t) * (size_t)nout);
        nin = nout;
        ++depth;
    }
    int32_t mx = 0;
    for (int64_t i = 0; i < n_rows; ++i)
        if (level[i] > mx) mx = level[i];
    *nlevels = mx + 1;
    free(visited);
    free(fq_in);
    free(fq_out);
    return 0;
}

/* ======================================================================== */
/* Validator (cpu_2d/validate/validate.h:249-491, at a 1x1 grid)            */
/* ======================================================================== */

int orc_validate(const int64_t *edges, int64_t m, int64_t n_rows, int64_t nglobalverts,
                 int64_t root, const int64_t *pred, int32_t *depth_out, int64_t *nedge)
{
    int bad = 0;
    *nedge = 0;
    /* check_value_ranges, validate.h:69-105; root range :258-264 */
    for (int64_t i = 0; i < n_rows; ++i)
        if (pred[i] < -1 || pred[i] >= nglobalverts || pred[i] >= n_rows) bad |= ORC_BAD_RANGE;
    if (root < 0 || root >= nglobalverts || root >= n_rows) bad |= ORC_BAD_RANGE;
    if (bad) return bad;
    /* root is its own parent, validate.h:285-295 */
    if (pred[root] != root) bad |= ORC_BAD_ROOT;
    /* nothing else is its own parent, validate.h:300-342 */
    for (int64_t i = 0; i < n_rows; ++i)
        if (i != root && pred[i] == i) bad |= ORC_BAD_SELF_PARENT;
    /* build_bfs_depth_map, validate.h:110-242: depth := UINT16_MAX, root := 0,
     * then repeat depth[v] = depth[pred v] + 1 until nothing changes; an
     * inconsistent already-set depth fails. */
    uint16_t *d = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n_rows);
    for (int64_t i = 0; i < n_rows; ++i) d[i] = UINT16_MAX;
    d[root] = 0;
    for (;;) {
        int any = 0;
        for (int64_t i = 0; i < n_rows; ++i) {
            if (i == root || pred[i] == -1) continue;
            const uint16_t dp = d[pred[i]];
            if (dp == UINT16_MAX) continue;
            if (d[i] != UINT16_MAX && d[i] != (uint16_t)(dp + 1)) {
                bad |= ORC_BAD_DEPTH_MAP;
            } else if (d[i] != (uint16_t)(dp + 1)) {
                d[i] = (uint16_t)(dp + 1);
                any = 1;
            }
        }
        if (!any) break;
    }
    /* edge checks over the directed (symmetrised) list, validate.h:401-471 */
    uint64_t *mark = (uint64_t *)calloc((size_t)(n_rows + 63) / 64, 8);
    int64_t visit = 0;
    for (int64_t e = 0; e < 2 * m; ++e) {
        const int64_t a = edges[e], b = edges[e ^ 1]; /* (v0,v1) then (v1,v0) */
        const int d0 = d[a], d1 = d[b];
        if (d0 - d1 > 1 || d1 - d0 > 1) bad |= ORC_BAD_EDGE_LEVEL;
        if (a == pred[b]) mark[b >> 6] |= UINT64_C(1) << (b & 63);
        if (pred[a] != -1) ++visit;
    }
    for (int64_t i = 0; i < n_rows; ++i)
        if (!(mark[i >> 6] >> (i & 63) & 1u) && pred[i] != -1 && pred[i] != i)
            bad |= ORC_BAD_PARENT_EDGE;
    *nedge = visit / 2;
    if (depth_out)
        for (int64_t i = 0; i < n_rows; ++i) depth_out[i] = d[i] == UINT16_MAX ? -1 : d[i];
    free(d);
    free(mark);
    return bad;
}

int64_t orc_nedge(const int64_t *gdeg, const int32_t *level, int64_t n_rows)
{
    int64_t s = 0;
    for (int64_t i = 0; i < n_rows; ++i)
        if (level[i] >= 0) s += gdeg[i];
    return s / 2;
}

/* ======================================================================== */
/* s4-bp128-d4 = CompositeCodec<SIMDBinaryPacking<SIMDIntegratedBlockPacker<  */
/*               CoarseDelta4SIMD,true>>, VariableByte<true>>                 */
/* (cpu_2d/compression/cpusimd/include/codecfactory.h:103-105)              */
/* ======================================================================== */

#define COOKIE 123456u /* simdbinarypacking.h:167 */

static uint32_t bits32(uint32_t v)
{
    uint32_t b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

size_t orc_s4bp128d4_max_words(size_t n)
{
    const size_t L = n / 128 * 128;
    return 4 + 4 * ((L + 2047) / 2048) + L + (5 * (n - L) + 3) / 4;
}

/* simdbinarypacking.h:174-235 (block part), compositecodec.h:31-51 (split at
 * a multiple of 128), variablebyte.h:30-45,100-161 (tail).  Deltas are
 * 4-strided, carried across miniblocks (deltatemplates.h:67-84); widths from
 * simdbinarypacking.h:118-132 / util.h:38-41; vertical packing as in
 * src/simdintegratedbitpacking.cpp (e.g. b=3 :875); b=32 stores raw values
 * (:16381), b=0 stores nothing (:37-40).  The output is taken to be 16-byte
 * aligned, so exactly three cookie words follow the length (:182-183). */
int orc_s4bp128d4_encode(const uint32_t *in, size_t n, uint32_t *out, size_t *nwords)
{
    const size_t L = n / 128 * 128;
    size_t p = 0;
    out[p++] = (uint32_t)L;
    out[p++] = COOKIE;
    out[p++] = COOKIE;
    out[p++] = COOKIE;
    uint32_t prev[4] = {0, 0, 0, 0};
    for (size_t g = 0; g < L; g += 2048) {
        const size_t nb = (L - g) / 128 < 16 ? (L - g) / 128 : 16;
        uint32_t B[16] = {0};
        uint32_t pr[4];
        memcpy(pr, prev, sizeof(pr));
        for (size_t i = 0; i < nb; ++i) {
            const uint32_t *x = in + g + 128 * i;
            uint32_t acc = 0;
            for (int k = 0; k < 128; ++k) {
                const uint32_t ref = k < 4 ? pr[k] : x[k - 4];
                acc |= x[k] - ref;
            }
            memcpy(pr, x + 124, sizeof(pr));
            B[i] = bits32(acc);
        }
        for (int w = 0; w < 4; ++w)
            out[p++] = B[4 * w] << 24 | B[4 * w + 1] << 16 | B[4 * w + 2] << 8 | B[4 * w + 3];
        for (size_t i = 0; i < nb; ++i) {
            const uint32_t *x = in + g + 128 * i;
            const uint32_t b = B[i];
            if (b == 32) {
                memcpy(out + p, x, 128 * 4);
            } else if (b > 0) {
                memset(out + p, 0, 4 * b * 4);
                for (int k = 0; k < 128; ++k) {
                    const uint32_t ref = k < 4 ? prev[k] : x[k - 4];
                    const uint32_t dlt = x[k] - ref;
                    const uint32_t lane = (uint32_t)(k & 3), idx = (uint32_t)(k >> 2);
                    const uint32_t bitpos = idx * b, word = bitpos >> 5, sh = bitpos & 31;
                    out[p + 4 * word + lane] |= dlt << sh;
                    if (sh + b > 32) out[p + 4 * (word + 1) + lane] |= dlt >> (32 - sh);
                }
            }
            p += 4 * b;
            memcpy(prev, x + 124, sizeof(prev));
        }
    }
    if (L < n) {
        unsigned char *bo = (unsigned char *)(out + p);
        size_t nb = 0;
        uint32_t pv = 0;
        for (size_t k = L; k < n; ++k) {
            uint32_t v = in[k] - pv;
            pv = in[k];
            while (v >= 128u) {
                bo[nb++] = (unsigned char)(v & 127u);
                v >>= 7;
            }
            bo[nb++] = (unsigned char)(v | 128u);
        }
        while (nb & 3) bo[nb++] = 0;
        p += nb / 4;
    }
    *nwords = p;
    return 0;
}

/* simdbinarypacking.h:237-289, compositecodec.h:52-71, variablebyte.h:217-341 */
int orc_s4bp128d4_decode(const uint32_t *in, size_t nwords, uint32_t *out, size_t *n)
{
    if (nwords < 4) return -1;
    const size_t L = in[0];
    if (in[1] != COOKIE || in[2] != COOKIE || in[3] != COOKIE) return -2;
    size_t p = 4, o = 0;
    uint32_t prev[4] = {0, 0, 0, 0};
    while (o < L) {
        const size_t nb = (L - o) / 128 < 16 ? (L - o) / 128 : 16;
        if (p + 4 > nwords) return -3;
        uint32_t B[16];
        for (int w = 0; w < 4; ++w) {
            B[4 * w] = (in[p + w] >> 24) & 0xFF;
            B[4 * w + 1] = (in[p + w] >> 16) & 0xFF;
            B[4 * w + 2] = (in[p + w] >> 8) & 0xFF;
            B[4 * w + 3] = in[p + w] & 0xFF;
        }
        p += 4;
        for (size_t i = 0; i < nb; ++i) {
            const uint32_t b = B[i];
            if (b > 32 || p + 4 * b > nwords) return -4;
            uint32_t *x = out + o;
            for (int k = 0; k < 128; ++k) {
                uint32_t dlt = 0;
                if (b == 32) {
                    x[k] = in[p + k];
                    continue;
                }
                if (b > 0) {
                    const uint32_t lane = (uint32_t)(k & 3), idx = (uint32_t)(k >> 2);
                    const uint32_t bitpos = idx * b, word = bitpos >> 5, sh = bitpos & 31;
                    uint64_t v = in[p + 4 * word + lane] >> sh;
                    if (sh + b > 32) v |= (uint64_t)in[p + 4 * (word + 1) + lane] << (32 - sh);
                    dlt = (uint32_t)(v & ((UINT64_C(1) << b) - 1));
                }
                x[k] = dlt + (k < 4 ? prev[k] : x[k - 4]);
            }
            memcpy(prev, x + 124, sizeof(prev));
            p += 4 * b;
            o += 128;
        }
    }
    /* varint tail; zero padding bytes never set the stop bit and emit nothing */
    const unsigned char *bi = (const unsigned char *)(in + p);
    const size_t nbytes = (nwords - p) * 4;
    uint32_t pv = 0;
    size_t q = 0;
    while (q < nbytes) {
        uint32_t v = 0;
        unsigned sh = 0;
        int done = 0;
        while (q < nbytes) {
            const unsigned char c = bi[q++];
            v += (uint32_t)(c & 127u) << sh;
            sh += 7;
            if (c & 128u) {
                done = 1;
                break;
            }
        }
        if (done) {
            pv += v;
            out[o++] = pv;
        }
    }
    *n = o;
    return 0;
}

/* CpuSimd::compress, cpu_2d/compression/cpusimd.hh:71-118: n > threshold
 * => narrow to uint32 and encode, otherwise a raw uint32 copy. */
int orc_cpusimd_compress(const int64_t *in, size_t n, size_t threshold, uint32_t *out,
                         size_t *csize)
{
    for (size_t i = 0; i < n; ++i)
        if (in[i] < 0 || in[i] > (int64_t)UINT32_MAX) return -1; /* :82-91 */
    if (n > threshold) {
        uint32_t *t = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (size_t i = 0; i < n; ++i) t[i] = (uint32_t)in[i];
        orc_s4bp128d4_encode(t, n, out, csize);
        free(t);
    } else {
        for (size_t i = 0; i < n; ++i) out[i] = (uint32_t)in[i];
        *csize = n;
    }
    return 0;
}

/* CpuSimd::decompress / isCompressed, cpusimd.hh:121-157,170-173: the stream
 * is decoded iff n > threshold && csize != n (n = original size, in/out). */
int orc_cpusimd_decompress(const uint32_t *in, size_t csize, size_t threshold, int64_t *out,
                           size_t *n)
{
    if (*n > threshold && csize != *n) {
        uint32_t *t = (uint32_t *)malloc(sizeof(uint32_t) * (*n ? *n : 1));
        size_t got = 0;
        int rc = orc_s4bp128d4_decode(in, csize, t, &got);
        if (rc == 0)
            for (size_t i = 0; i < got; ++i) out[i] = t[i];
        *n = got;
        free(t);
        return rc;
    }
    for (size_t i = 0; i < csize; ++i) out[i] = in[i];
    *n = csize;
    return 0;
}

/* ======================================================================== */
/* Root selection, cpu_2d/main.cpp:351-411                                  */
/* ======================================================================== */

/* std::knuth_b = shuffle_order_engine<minstd_rand0, 256> as libstdc++ defines
 * it: table filled from the base engine, then y; each draw picks slot
 * j = floor(256 * (y - min) / (range + 1)) in double arithmetic. */
typedef struct {
    uint64_t x;
    uint64_t v[256];
    uint64_t y;
} knuth_b;

static uint64_t minstd0(knuth_b *g)
{
    g->x = g->x * 16807u % 2147483647u;
    return g->x;
}

static void knuth_init(knuth_b *g)
{
    g->x = 1;
    for (int i = 0; i < 256; ++i) g->v[i] = minstd0(g);
    g->y = minstd0(g);
}

static uint64_t knuth_next(knuth_b *g)
{
    const double range = 2147483646.0 - 1.0;
    const size_t j = (size_t)(256.0 * ((double)(g->y - 1) / (range + 1.0)));
    g->y = g->v[j];
    g->v[j] = minstd0(g);
    return g->y;
}

/* uniform_int_distribution<int64>(0, vertices-1) downscaling branch */
static int64_t uniform_draw(knuth_b *g, int64_t vertices)
{
    const uint64_t urngrange = 2147483646u - 1u;
    const uint64_t uerange = (uint64_t)vertices;
    const uint64_t scaling = urngrange / uerange;
    const uint64_t past = uerange * scaling;
    uint64_t r;
    do r = knuth_next(g) - 1; while (r >= past);
    return (int64_t)(r / scaling);
}

void orc_knuth_b_draws(int64_t vertices, int count, int64_t *out)
{
    knuth_b g;
    knuth_init(&g);
    for (int i = 0; i < count; ++i) out[i] = uniform_draw(&g, vertices);
}

int orc_select_roots(const int64_t *rowptr, int64_t n_rows, int64_t vertices, int nroots,
                     int64_t *roots)
{
    if (vertices < 1 || vertices - 1 > (int64_t)(2147483646u - 1u)) return -1;
    knuth_b g;
    knuth_init(&g);
    int found = 0, tries = 0;
    const int maxtries = nroots * 32;
    while (found < nroots && tries < maxtries) {
        ++tries;
        const int64_t s = uniform_draw(&g, vertices);
        int dup = 0;
        for (int i = 0; i < found; ++i) dup |= roots[i] == s;
        if (dup) continue;
        const int64_t deg = (s < n_rows) ? rowptr[s + 1] - rowptr[s] : 0;
        if (deg > 0) roots[found++] = s;
    }
    return found;
}

/* ======================================================================== */
/* vreduce combines: bitmap OR (cpubfs_bin.cpp:55-65) and the canonical     */
/* minimum-id predecessor combine replacing globalbfs.hh:235-524's rank     */
/* precedence (-1 = absent)                                                 */
/* ======================================================================== */

void orc_vreduce_or(int k, const uint64_t *maps, int64_t nwords, uint64_t *out)
{
    for (int64_t i = 0; i < nwords; ++i) {
        uint64_t v = 0;
        for (int j = 0; j < k; ++j) v |= maps[(int64_t)j * nwords + i];
        out[i] = v;
    }
}

void orc_vreduce_min_pred(int k, const int64_t *preds, int64_t n, int64_t *out)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t best = -1;
        for (int j = 0; j < k; ++j) {
            const int64_t p = preds[(int64_t)j * n + i];
            if (p >= 0 && (best < 0 || p < best)) best = p;
        }
        out[i] = best;
    }
}
