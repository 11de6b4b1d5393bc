/*
 * g5hip.h -- C ABI of the MI355X (gfx950) gpugraph500 hot path.
 *
 * Drop-in boundary for one rank's share of the reference's 2D BFS:
 *   - the local expansion plug-in (the GPU "Derived" class of GlobalBFS,
 *     cpu_2d/cuda/cuda_bfs.h:39-84, cpu_2d/globalbfs.hh:51-126),
 *   - the vreduce combines (cpu_2d/vreduce.hpp:41-544,
 *     cpu_2d/globalbfs.hh:235-524),
 *   - the s4-bp128-d4 frontier-queue codec behind Compression<T,T_C>
 *     (cpu_2d/compression/compression.hh:11-28, cpusimd.hh:20-184).
 *
 * Conventions: every function returns 0 on success and a negative G5_E* code
 * on failure; g5_last_error() describes the last failure of the calling
 * thread.  No exception crosses this boundary.  Handles are opaque.  Pointers
 * named d_* are device (HBM) pointers; all others are host pointers.
 * Vertex ids in queues are uint32 local ids (the reference narrows its int64
 * queues to uint32 before coding, cpusimd.hh:93-96).
 */
#ifndef G5HIP_H
#define G5HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    G5_OK = 0,
    G5_E_ARG = -1,      /* invalid argument (sizes, ids, null handles) */
    G5_E_HIP = -2,      /* a HIP runtime call failed */
    G5_E_NOMEM = -3,    /* device allocation failed */
    G5_E_OVERFLOW = -4, /* output buffer too small (reference: "increase
                           queue-sizing factor", enactor_multi_gpu_2d.cuh:706-729) */
    G5_E_STREAM = -5,   /* malformed codec stream (simdbinarypacking.h:240-247) */
    G5_E_STATE = -6     /* call out of order (e.g. level before start) */
};

typedef struct g5_graph g5_graph;
typedef struct g5_bfs g5_bfs;
typedef struct g5_msbfs g5_msbfs;
typedef struct g5_validator g5_validator;

/* ---- library ---- */
const char *g5_last_error(void);
int g5_init(int device);
int g5_device_sync(void);
/* Free and total device memory of the current device (hipMemGetInfo), for
 * admission by memory. */
int g5_mem_info(int64_t *free_bytes, int64_t *total_bytes);
int g5_dmalloc(void **d_ptr, size_t bytes);
int g5_dfree(void *d_ptr);
int g5_memcpy_h2d(void *d_dst, const void *src, size_t bytes);
int g5_memcpy_d2h(void *dst, const void *d_src, size_t bytes);
int g5_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
int g5_fill_u64(uint64_t *d, int64_t n, uint64_t value);
/* *ndiff = positions i < n where d_a[i] != d_b[i] (device arrays; e.g. two
 * predecessor maps viewed as 2n words), so results are compared where they
 * lie. */
int g5_diff_u32(const uint32_t *d_a, const uint32_t *d_b, int64_t n, int64_t *ndiff);
/* Device bitmap <-> sorted queue (the frontier representations the 2D
 * exchange moves; cuda/cuda_bfs.cu:380-475 sorts and uniques on the host).
 * d_queue needs room for every set bit (n_words * 64 at most); *n = length.
 * g5_queue_to_bitmap clears d_bits first and rejects ids >= n_words * 64. */
int g5_bitmap_to_queue(const uint64_t *d_bits, int64_t n_words, uint32_t *d_queue, int64_t *n);
int g5_queue_to_bitmap(const uint32_t *d_queue, int64_t n, uint64_t *d_bits, int64_t n_words);
/* *count = set bits of d_bits[n_words] (the frontier size a direction switch
 * needs, without building the queue). */
int g5_bitmap_count(const uint64_t *d_bits, int64_t n_words, int64_t *count);

/* ---- the DistMatrix2d block (cpu_2d/distmatrix2d.hh:30-130) ----
 * Upload a host CSR block: rowptr int64[n_rows+1], cols uint32[nnz] local
 * column ids (ascending, unique per row), gdeg int64[n_rows] generated
 * directed degrees (nullable).  n_rows and n_cols must be multiples of 64.
 * Replaces CsrProblem::FromHostProblem (csr_problem_2d.cuh:314) as called
 * from CUDA_BFS's constructor (cuda/cuda_bfs.cu:79-87). */
int g5_graph_upload(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows, int64_t n_cols,
                    const int64_t *gdeg, int64_t row_start, int64_t col_start, g5_graph **out);
/* Generate the Graph500 Kronecker graph (cpu_2d/generator/make_graph.c:34-69,
 * seeds as cpu_2d/main.cpp:221) and build the 1x1 block on the device with
 * setupMatrix2's rules (cpu_2d/distmatrix2d.hh:745-1166): sorted, deduplicated
 * columns, self-loops dropped, rows padded to 64. */
int g5_graph_kronecker(int scale, int edgefactor, uint64_t seed1, uint64_t seed2, g5_graph **out);
/* The R x C blocks of the same graph as R x C ranks hold them (setupMatrix2
 * at an R x C grid, distmatrix2d.hh:745-1166, with the CUDA path's padding of
 * the vertex count to a multiple of R, cuda/cuda_bfs.h:40): blocks[r*C + c]
 * keeps the symmetrised entries with row in row part r and column in column
 * part c, local ids; gdeg counts the block's generated entries, so a row's
 * generated degree is the sum over its row of blocks.  1 <= R, C <= 64. */
int g5_graph_kronecker_2d(int scale, int edgefactor, uint64_t seed1, uint64_t seed2, int R, int C,
                          g5_graph **blocks);
/* Block (r, c)'s bounds for a global vertex count: parts in 64-vertex units,
 * the first (units mod parts) parts one unit longer (distmatrix2d.hh:807-815
 * with ALG = 64), after padding nglobal to a multiple of R.  Host only. */
int g5_block_bounds(int64_t nglobal, int R, int C, int r, int c, int64_t *row_start, int64_t *n_rows,
                    int64_t *col_start, int64_t *n_cols);
int g5_graph_info(const g5_graph *g, int64_t *n_rows, int64_t *n_cols, int64_t *nnz,
                  int64_t *nglobal, int64_t *n_gen_edges);
/* global ids of the block's local row 0 and column 0 */
int g5_graph_bounds(const g5_graph *g, int64_t *row_start, int64_t *col_start);
/* Degree histogram of the block's rows: hist[0] = rows without neighbours,
 * hist[k] = rows with 2^(k-1) <= degree < 2^k, hist[nbins-1] open-ended
 * (2 <= nbins <= 64).  The batched search's shapes are kept per scale and
 * histogram (its degree skew). */
int g5_graph_degree_hist(const g5_graph *g, int64_t *hist, int nbins);
/* generated directed degree per local row, gdeg[n_rows] */
int g5_graph_gdeg(const g5_graph *g, int64_t *gdeg);
int g5_graph_download(const g5_graph *g, int64_t *rowptr, uint32_t *cols, int64_t *gdeg);
int g5_graph_free(g5_graph *g);
/* Edges [start, end) of the Kronecker list generated on the device and copied
 * to edges[2*(end-start)] (generate_kronecker_range, graph_generator.c:157-187). */
int g5_kronecker_edges(int scale, uint64_t seed1, uint64_t seed2, int64_t start, int64_t end,
                       int64_t *edges);
/* The benchmark's search keys (cpu_2d/main.cpp:351-411): up to nroots
 * distinct vertices with nonzero degree, drawn from a default-seeded
 * std::knuth_b via uniform_int_distribution<int64>(0, vertices-1), at most
 * 32 draws per root; *found = number selected. */
int g5_select_roots(const g5_graph *g, int64_t vertices, int nroots, int64_t *roots, int *found);

/* ---- local expansion plug-in: the CUDA_BFS hooks (cuda/cuda_bfs.h:68-83) ---- */
int g5_bfs_create(g5_graph *g, g5_bfs **out);
int g5_bfs_free(g5_bfs *b);
/* setStartVertex (cuda_bfs.cu:547-640); root = -1 on ranks that do not own it */
int g5_bfs_start(g5_bfs *b, int64_t root);
/* runLocalBFS (cuda_bfs.cu:642-659) + istheresomethingnew (globalbfs.hh:716):
 * expands the in-queue; *any_new = 1 if the out-queue is non-empty.  New
 * vertices get their minimum-id frontier neighbour as parent. */
int g5_bfs_expand(g5_bfs *b, int *any_new);
/* getBackOutqueue (cuda_bfs.cu:380-475): the column out-queue, sorted and
 * unique, into q (capacity *n entries); *n = queue length. */
int g5_bfs_outqueue(g5_bfs *b, uint32_t *q, int64_t *n);
/* setModOutgoingFQ (cuda_bfs.cu:230-270): adopt the column-reduced queue
 * (q = NULL: keep the own out-queue) and mark it visited at depth + 1. */
int g5_bfs_set_outqueue(g5_bfs *b, const uint32_t *q, int64_t n);
/* setIncommingFQ + setBackInqueue (cuda_bfs.cu:275-321,477-545): the row
 * queue for the next level (q = NULL: the adopted out-queue). */
int g5_bfs_set_inqueue(g5_bfs *b, const uint32_t *q, int64_t n);
/* getBackPredecessor (cuda_bfs.cu:328-378): pred int64[n_cols] global ids
 * (-1 unreached, root -> root), level int32[n_cols] (-1 unreached); either
 * pointer may be NULL. */
int g5_bfs_result(g5_bfs *b, int64_t *pred, int32_t *level);
/* The same hooks with the frontiers left in HBM, for an exchange that moves
 * device bitmaps (as RCCL would between GPUs) instead of host queues:
 *   g5_bfs_out_bitmap      the column out-bitmap of the last expand (n_cols/64
 *                          words, borrowed, valid until the next call on b);
 *                          the device side of getOutgoingFQ (cuda_bfs.cu:185-228)
 *   g5_bfs_set_out_bitmap  setModOutgoingFQ from a column-reduced device
 *                          bitmap (NULL = keep the own out-bitmap)
 *   g5_bfs_set_in_bitmap   setIncommingFQ + setBackInqueue from a device bitmap
 *                          over the row slice (n_rows/64 words)
 *   g5_bfs_pred_device     getBackPredecessor into device memory: int64[n_cols]
 *                          global ids, -1 where this block found no parent,
 *                          root -> root; combine blocks with g5_vreduce_min_pred */
int g5_bfs_out_bitmap(g5_bfs *b, const uint64_t **d_bits, int64_t *n_words);
int g5_bfs_set_out_bitmap(g5_bfs *b, const uint64_t *d_bits);
int g5_bfs_set_in_bitmap(g5_bfs *b, const uint64_t *d_bits);
int g5_bfs_pred_device(g5_bfs *b, int64_t *d_pred);
/* the level map int32[n_cols] of the last search (-1 unreached) copied to
 * device memory d_level */
int g5_bfs_level_device(g5_bfs *b, int32_t *d_level);
/* A bottom-up level of the hook path, the alternative to g5_bfs_expand for
 * dense levels (no counterpart in the reference, whose CUDA path is
 * top-down only): every unvisited column vertex of the block scans its
 * neighbours in the block's row range in ascending order and stops at the
 * first vertex of the row frontier, which must have been set with
 * g5_bfs_set_in_bitmap.  gt is the transposed block (for a symmetric graph
 * on an R x R grid, block (c, r) of rank (r, c)): its rows are this block's
 * columns.  The out-bitmap, parents and *any_new are left as g5_bfs_expand
 * leaves them, so the column and row steps are unchanged; each rank's
 * parents are again the minimum-id frontier vertex of its row range. */
int g5_bfs_expand_bottom_up(g5_bfs *b, const g5_graph *gt, int *any_new);
/* The whole 1x1 runBFS (globalbfs.hh:548-1263) resident in HBM: no host
 * round trip per level.  Dense levels run bottom-up (every unvisited vertex
 * stops at its first -- minimum-id -- frontier neighbour), chosen by the
 * direction switch of Beamer et al. (SC'12: bottom-up once the frontier's
 * edges exceed the unexplored edges / 20, top-down again below n / 24
 * frontier vertices); results are identical in both directions.  flags:
 * G5_RUN_CODEC routes every level's queue through s4-bp128-d4 encode ->
 * decode on the device, as the row broadcast delivers it to non-root ranks
 * (globalbfs.hh:864-905); G5_RUN_TOPDOWN keeps every level top-down.
 * *ms = device time. */
enum { G5_RUN_CODEC = 1, G5_RUN_TOPDOWN = 2 };
int g5_bfs_run(g5_bfs *b, int64_t root, int flags, int32_t *nlevels, float *ms);
/* The direction switch of g5_bfs_run: bottom-up once the frontier's edges
 * exceed the edges not yet checked top-down / alpha (20), top-down again
 * below n / beta frontier vertices (24).  Results never depend on them. */
int g5_bfs_tune(g5_bfs *b, int64_t alpha, int64_t beta);
/* Per-kernel accounting of the expansion kernel across g5_bfs_run and
 * g5_bfs_expand calls:
 * enable (resets the totals), then read total device ms, launches,
 * algorithmic bytes (16 B per frontier vertex, 12 B per frontier edge, 12 B
 * per newly reached vertex) and frontier edges. */
int g5_bfs_profile(g5_bfs *b, int enable);
int g5_bfs_stats(g5_bfs *b, double *ms, int64_t *launches, double *bytes, int64_t *edges);
/* The same per kernel: 0 = k_expand (as g5_bfs_stats), 1 = k_bottom_up (16 B
 * per unvisited vertex scanned, 12 B per neighbour checked, 4 B per newly
 * reached vertex, 16 B per 64-vertex word; edges = neighbours checked). */
int g5_bfs_kernel_stats(g5_bfs *b, int kernel, double *ms, int64_t *launches, double *bytes, int64_t *edges);
/* TEPS numerator of the last search: generated undirected edges with a
 * reached endpoint (validate/validate.h:439-485). */
int g5_bfs_nedge(g5_bfs *b, int64_t *nedge);
/* After g5_bfs_run: nedge as above and the checksum of the predecessor map,
 * sum over reached v of mix64(v << 32 | (pred(v) + 1)) mod 2^64 (mix64 =
 * SplitMix64's finaliser), in one pass on the device -- what
 * g5_msbfs_leave reports per lane, so a server can answer a query from
 * either search.  Needs a diagonal (1x1) block. */
int g5_bfs_summary(g5_bfs *b, int64_t *nedge, uint64_t *checksum);

/* ---- batched multi-source BFS (no counterpart in the reference: it stands
 * beside the single-root plug-in; results equal nroots separate searches
 * with the minimum-id parent rule of cpu_2d/cpubfs_bin.cpp:150-173) ----
 * Needs a diagonal (1x1) block.  Up to 64 roots per run, one uint64 lane mask
 * per vertex. */
int g5_msbfs_create(g5_graph *g, g5_msbfs **out);
int g5_msbfs_free(g5_msbfs *m);
/* The device memory of a batched search over g, for admission by memory
 * (host arithmetic, nothing is allocated): *graph_bytes held by the graph
 * itself (offsets, columns, degrees); *fixed_bytes that g5_msbfs_create
 * allocates whatever the lanes in use; *lane_bytes per lane (its parent and
 * level columns, 5 bytes per vertex, and its bit in the visited and frontier
 * masks), allocated for all 64 lanes at creation.  Any pointer may be NULL. */
int g5_msbfs_footprint(const g5_graph *g, int64_t *graph_bytes, int64_t *fixed_bytes, int64_t *lane_bytes);
/* Searches all roots; *nlevels = levels of the deepest search, *ms = device
 * time from the start to the final maps (g5_msbfs_maps), as the reference's
 * runBFS window ends with the predecessors final (cpu_2d/main.cpp:419-462,
 * globalbfs.hh:985-1082).  The search and its last (finalising) pass are
 * split by g5_msbfs_times. */
int g5_msbfs_run(g5_msbfs *m, const int64_t *roots, int nroots, int32_t *nlevels, float *ms);
/* The result of the last batch run as it lies in HBM (borrowed; valid until
 * the next run, open or free on m):
 *   d_pred   int32[n][64]: row v (original vertex id), slot lane_of[i] holds
 *            root i's parent of v as an original id (root -> root, -1
 *            unreached); slots of lanes >= nroots are undefined
 *   d_level  uint8[n][64], the same slots: distance from the root (255
 *            unreached) -- the level map, when *levels_final (a search of
 *            fewer than 255 levels; deeper ones take g5_msbfs_result)
 *   lane_of  int32[nroots]: root i's lane (roots are ranked by reach)
 * Any pointer may be NULL. */
int g5_msbfs_maps(g5_msbfs *m, const int32_t **d_pred, const uint8_t **d_level, int32_t *lane_of,
                  int *levels_final);
/* Device time of the last batch run's search (start to the last level) and
 * of its finalising pass (unreached slots of the used lanes). */
int g5_msbfs_times(g5_msbfs *m, float *search_ms, float *finalize_ms);
/* Root `lane`'s result as the single-root search returns it: pred int64[n]
 * (-1 unreached, root -> root) and level int32[n] (-1 unreached); either may
 * be NULL.  Allocates device scratch.  Levels come from the level bytes, or,
 * for a search of 255 levels or more, from the parent tree by pointer
 * jumping (G5_E_STATE if that tree is broken: a cycle, or a vertex whose
 * parent is unreached). */
int g5_msbfs_result(g5_msbfs *m, int lane, int64_t *pred, int32_t *level);
/* The same pred int64[n] written to device memory d_pred (no allocation), for
 * g5_validate_device without a host round trip. */
int g5_msbfs_pred_device(g5_msbfs *m, int lane, int64_t *d_pred);
/* The same level int32[n] written to device memory d_level (G5_E_STATE for
 * a search of 255 levels or more: use g5_msbfs_result). */
int g5_msbfs_level_device(g5_msbfs *m, int lane, int32_t *d_level);
/* TEPS numerators nedge[nroots] of the last run (validate/validate.h:439-485). */
int g5_msbfs_nedge(g5_msbfs *m, int64_t *nedge);
/* Direction per level: 0 (default) push while the frontier's edges are
 * fewer than nnz/16, pull otherwise; 1 pull only; 2 push only.  Results are
 * identical in every mode. */
int g5_msbfs_set_mode(g5_msbfs *m, int mode);
/* Per-kernel accounting.  Units 0..2 are the pull kernels of the degree bins
 * (0: workgroup per vertex, degree > 2048; 1: wavefront per vertex,
 * 33..2048; 2: thread per vertex, 1..32), unit 3 the push levels and the
 * pushed lanes of split levels.  Enabling runs pull levels as direct launches timed with
 * events (otherwise each pull level replays a captured hipGraph); bytes are
 * the algorithmic bytes of the unit's launches. */
int g5_msbfs_profile(g5_msbfs *m, int enable);
int g5_msbfs_stats(g5_msbfs *m, int unit, double *ms, int64_t *launches, double *bytes);
/* Per level of the last profiled run, 12 doubles each: ms of units 0..3 (as
 * in g5_msbfs_stats), edges read by units 0..3, new (vertex, root) pairs, 1
 * if the level pushed (2 if it pulled some lanes and pushed the others), the
 * frontier's vertices and edges.  *nlev = levels recorded; at most cap are copied. */
int g5_msbfs_levels(g5_msbfs *m, double *out, int cap, int *nlev);
/* Vertex counts of the three degree bins (units 0..2 above; the first two
 * end at multiples of 64, so up to 63 vertices sit one bin early), then the number
 * of split levels since the last g5_msbfs_run / g5_msbfs_open; up to cap. */
int g5_msbfs_info(g5_msbfs *m, int64_t *bins, int cap);
/* Continuous mode (serving): lanes join and leave between levels, so a root
 * that arrives while other searches run starts at the next level instead of
 * waiting for the batch to end.  Each lane's result equals the search of its
 * root alone.
 *   g5_msbfs_open   every lane free (ends any batch run on m)
 *   g5_msbfs_join   roots[i] starts in free lane lanes[i] (n <= 64)
 *   g5_msbfs_step   one level for all lanes in use; *live = lanes whose
 *                   frontier is still non-empty -- a lane in use but not in
 *                   *live has finished; *ms = device time of the level
 *   g5_msbfs_leave  for each lane in the mask: nedge[lane] (TEPS numerator)
 *                   and checksum[lane] = sum over reached vertices v of
 *                   mix64(v << 32 | (pred(v) + 1)) mod 2^64 (original ids;
 *                   mix64 = SplitMix64's finaliser); then the lanes are free.
 *                   g5_msbfs_result may be called for a lane before it leaves. */
/* Shape parameters of the batched search (results never depend on them):
 * G5_TUNE_THREAD_MAX / G5_TUNE_WAVE_MAX bound the degree bins of the
 * thread-per-vertex and wavefront-per-vertex pull kernels (default 32, 2048;
 * higher degrees take a workgroup per vertex); G5_TUNE_PUSH_RATIO pushes a
 * level whose frontier edges are below nnz / ratio (16); G5_TUNE_ACTIVE_RATIO
 * makes pull levels consult the active bitmap below nnz / ratio edges (2);
 * G5_TUNE_LANE_RATIO: a pull level that follows a push level pushes the
 * lanes whose frontier edges are below nnz / ratio (64; 0 = never), which
 * stops one lagging root from making every vertex scan its whole list;
 * G5_TUNE_SPLIT_MAX: such split levels bound the thread-per-vertex bin at
 * min(split_max, thread_max) (16), as every row is scanned to its end there.
 * g5_msbfs_autotune searches these by coordinate descent over candidate
 * values (passes rounds, two timed batch searches of roots per candidate),
 * keeps the fastest and reports its time and the number of candidates run. */
enum {
    G5_TUNE_THREAD_MAX = 0,
    G5_TUNE_WAVE_MAX = 1,
    G5_TUNE_PUSH_RATIO = 2,
    G5_TUNE_ACTIVE_RATIO = 3,
    G5_TUNE_LANE_RATIO = 4,
    G5_TUNE_SPLIT_MAX = 5
};
int g5_msbfs_tune(g5_msbfs *m, int key, int64_t value);
int g5_msbfs_get_tune(g5_msbfs *m, int key, int64_t *value);
int g5_msbfs_autotune(g5_msbfs *m, const int64_t *roots, int nroots, int passes, double *best_ms, int *trials);
/* Placement of the parent records (no counterpart in the reference, whose
 * host predecessor buffers have no such effect): the batched search runs at
 * one of two speeds, about 12 % apart at scale 26, depending on the physical
 * memory behind the parent records.  Tries `trials` (0..16) fresh
 * allocations of the records (each a copy, so results stay valid) and keeps
 * the one whose search of roots is fastest (best of two runs; a move must
 * gain 2 %).  first_ms: the time before, best_ms: after, moves: allocations
 * kept.  Needs the records' size (256 B per vertex) in free memory while it
 * runs; stops trying when an allocation fails.  After it the maps hold the
 * results of roots.  G5_E_STATE while continuous-mode lanes are in use. */
int g5_msbfs_place(g5_msbfs *m, const int64_t *roots, int nroots, int trials, float *first_ms, float *best_ms,
                   int *moves);
int g5_msbfs_open(g5_msbfs *m);
int g5_msbfs_join(g5_msbfs *m, const int64_t *roots, const int *lanes, int n);
/* reach[i] = the edges of roots[i]'s neighbours (the frontier edges of its
 * second level), n <= 64.  Batch runs give roots lanes by descending reach,
 * so that the lanes a vertex gains at one level share parent-record sectors;
 * a server joining several roots can do the same. */
int g5_msbfs_reach(g5_msbfs *m, const int64_t *roots, int n, int64_t *reach);
int g5_msbfs_step(g5_msbfs *m, uint64_t *live, float *ms);
int g5_msbfs_leave(g5_msbfs *m, uint64_t lanes, int64_t *nedge, uint64_t *checksum);

/* ---- s4-bp128-d4 codec (codecfactory.h:103-105) on the device ----
 * Word stream identical to the reference codec's output for 16-byte aligned
 * buffers.  Worst-case size: g5_s4_max_words(n). */
size_t g5_s4_max_words(size_t n);
/* d_in uint32[n] -> d_out words; *nwords = stream length (host). */
int g5_s4_encode(const uint32_t *d_in, size_t n, uint32_t *d_out, size_t out_cap,
                 size_t *nwords);
/* d_in words[nwords] -> d_out uint32 (capacity out_cap); *n = values decoded. */
int g5_s4_decode(const uint32_t *d_in, size_t nwords, uint32_t *d_out, size_t out_cap, size_t *n);
/* The exchange's codec hop in one call (compress at the sender, decompress at
 * the receiver): d_in uint32[n] -> d_words (stream, *nwords long) -> d_out
 * uint32[n] (may alias d_in), with one host synchronisation at the end;
 * G5_E_STREAM unless exactly the n values came back. */
int g5_s4_roundtrip(const uint32_t *d_in, size_t n, uint32_t *d_words, size_t cap_words, uint32_t *d_out,
                    size_t *nwords);
/* The same round trip, allocating nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is a static device array of 16 MiB that every GPU the library runs
 * on holds until the process ends; n is limited to G5_S4_NOALLOC_MAX_VALUES
 * (G5_E_ARG beyond).  Results, streams and errors are those of
 * g5_s4_roundtrip.  All calls on one GPU share the array: calls that overlap
 * on one GPU race. */
enum { G5_S4_NOALLOC_MAX_VALUES = 1 << 22 };
int g5_s4_roundtrip_noalloc(const uint32_t *d_in, size_t n, uint32_t *d_words, size_t cap_words, uint32_t *d_out,
                            size_t *nwords);

/* ---- vreduce combines (vreduce.hpp:41-544, globalbfs.hh:235-524) ----
 * d_maps: k contiguous bitmaps of n_words uint64 each -> d_out = OR. */
int g5_vreduce_or(const uint64_t *d_maps, int k, int64_t n_words, uint64_t *d_out);
/* The same over k <= 64 bitmaps wherever they lie: d_maps is a host array of
 * k device pointers (16-byte aligned), e.g. the ranks' g5_bfs_out_bitmap
 * buffers, so nothing is copied together first.  The reduce_fq_out functor's
 * inputs (globalbfs.hh:617-622) map onto it one to one. */
int g5_vreduce_or_list(const uint64_t *const *d_maps, int k, int64_t n_words, uint64_t *d_out);
/* d_preds: k contiguous int64[n] predecessor arrays (-1 absent) -> d_out =
 * minimum present id (the canonical rule replacing rank precedence). */
int g5_vreduce_min_pred(const int64_t *d_preds, int k, int64_t n, int64_t *d_out);
/* The same over k <= 64 arrays named by a host array of device pointers. */
int g5_vreduce_min_pred_list(const int64_t *const *d_preds, int k, int64_t n, int64_t *d_out);

/* Device-time accounting of the exchange entry points called through this
 * ABI: the vreduce combines (g5_vreduce_*) and the codec's g5_s4_encode and
 * g5_s4_decode.  g5_exchange_profile(1) starts it (and resets the totals), 0
 * stops it; g5_exchange_stats reads per unit the device ms of the unit's
 * kernels, their algorithmic bytes and the calls: OR (k+1)*8*n_words,
 * min-pred (k+1)*8*n, encode 4n + 4*csize, decode 4*csize + 4n.  A
 * combine's time is its kernel's, from start to end; the codec's runs from
 * its first kernel's launch to its last kernel's end (launch gaps included),
 * and a decode's excludes the host round trip between its header search and
 * its unpacking. */
enum { G5_UNIT_OR = 0, G5_UNIT_MIN_PRED = 1, G5_UNIT_ENCODE = 2, G5_UNIT_DECODE = 3, G5_UNIT_COUNT = 4 };
int g5_exchange_profile(int enable);
int g5_exchange_stats(int unit, double *ms, double *bytes, int64_t *calls);

/* ---- result validation (validate_bfs_result, validate/validate.h:249-491) ----
 * For a whole (1x1) graph.  *bad is 0 for a valid predecessor map, else the
 * OR of the failed checks below; a range failure stops the other checks, as
 * in the reference.  *nedge = generated undirected edges whose source has a
 * parent (validate.h:474-477,486), depth int32[n] (-1 unreached; may be
 * NULL), *ms = device time (may be NULL).  The graph must outlive the
 * validator. */
enum {
    G5_BAD_RANGE = 1,         /* a parent or the root out of range (validate.h:69-105,258-264) */
    G5_BAD_ROOT = 2,          /* the root is not its own parent (:285-295) */
    G5_BAD_SELF_PARENT = 4,   /* another vertex is its own parent (:300-342) */
    G5_BAD_DEPTH_MAP = 8,     /* inconsistent depth map (:200-222); a set-once map never is */
    G5_BAD_EDGE_LEVEL = 16,   /* an edge joins depths more than one apart (:425-433) */
    G5_BAD_PARENT_EDGE = 32   /* a vertex has no edge to its parent (:436-471) */
};
int g5_validator_create(const g5_graph *g, g5_validator **out);
int g5_validator_free(g5_validator *v);
/* pred int64[n] in host memory */
int g5_validate(g5_validator *v, int64_t root, int64_t nglobalverts, const int64_t *pred, int *bad,
                int64_t *nedge, int32_t *depth, float *ms);
/* pred int64[n] in device memory (g5_bfs_pred_device, g5_msbfs_pred_device) */
int g5_validate_device(g5_validator *v, int64_t root, int64_t nglobalverts, const int64_t *d_pred, int *bad,
                       int64_t *nedge, int32_t *depth, float *ms);

#ifdef __cplusplus
}
#endif
#endif
