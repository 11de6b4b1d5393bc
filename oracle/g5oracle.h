/*
 * g5oracle.h -- CPU oracle for the gpugraph500 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function restates the reference algorithm it is named after; each
 * definition in g5oracle.c cites the reference file:line it follows
 * (paths relative to the reference checkout, e.g. cpu_2d/cpubfs_bin.cpp).
 *
 * How each function is pinned (i.e. why it is trusted):
 *   orc_make_mrg_seed / orc_generate_kronecker_range
 *       golden edge lists written by the reference generator compiled from
 *       its own sources (oracle/_ref), tests/golden/kron_s*.bin
 *   orc_build_csr      brute force (Python set construction) on small graphs,
 *                      plus sortedness/uniqueness/symmetry/no-self-loop invariants
 *   orc_bfs_bin        brute-force BFS on tiny graphs (all roots), closed forms
 *                      (path, star, complete, disconnected), the validator
 *                      restatement, and level agreement with orc_bfs_simple
 *   orc_bfs_simple     level agreement with orc_bfs_bin and brute force
 *   orc_validate       mutation tests (wrong parent, level skip, missing and
 *                      extra vertex must each fail)
 *   orc_s4bp128d4_*    golden word streams written by the reference codec
 *                      compiled from its own sources (oracle/_ref)
 *   orc_cpusimd_*      gate behaviour pinned by the same golden streams
 *   orc_select_roots   golden root lists written by libstdc++'s std::knuth_b
 *                      and std::uniform_int_distribution (tests/golden/roots_*)
 *   orc_vreduce_*      numpy closed forms (bitwise_or.reduce, minimum.reduce)
 */
#ifndef G5ORACLE_H
#define G5ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Graph500 Kronecker generator ---- */
void orc_make_mrg_seed(uint64_t userseed1, uint64_t userseed2, uint32_t seed[5]);
void orc_generate_kronecker_range(const uint32_t seed[5], int logN, int64_t start_edge,
                                  int64_t end_edge, int64_t *edges /* 2 per edge */);

/* ---- DistMatrix2d at a 1x1 grid (WOLO, ALG = 64) ---- */
int64_t orc_num_vertices(const int64_t *edges, int64_t m);
int64_t orc_padded_rows(int64_t nglobal);
/* cols must hold 2*m entries; returns nnz.  gdeg (nullable, n_rows entries)
 * receives the count of generated directed entries per row, self-loops and
 * duplicates included. */
int64_t orc_build_csr(const int64_t *edges, int64_t m, int64_t n_rows, int64_t *rowptr,
                      uint32_t *cols, int64_t *gdeg);

/* ---- BFS ---- */
int orc_bfs_bin(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows, int64_t root,
                int64_t *pred, int32_t *level, int32_t *nlevels);
int orc_bfs_simple(const int64_t *rowptr, const uint32_t *cols, int64_t n_rows, int64_t root,
                   int64_t *pred, int32_t *level, int32_t *nlevels);

/* ---- validator (bit flags of the failed checks; 0 = valid) ---- */
enum {
    ORC_BAD_RANGE = 1,
    ORC_BAD_ROOT = 2,
    ORC_BAD_SELF_PARENT = 4,
    ORC_BAD_DEPTH_MAP = 8,
    ORC_BAD_EDGE_LEVEL = 16,
    ORC_BAD_PARENT_EDGE = 32
};
int orc_validate(const int64_t *edges, int64_t m, int64_t n_rows, int64_t nglobalverts,
                 int64_t root, const int64_t *pred, int32_t *depth_out, int64_t *nedge);

/* ---- s4-bp128-d4 codec and the CpuSimd gate ---- */
size_t orc_s4bp128d4_max_words(size_t n);
int orc_s4bp128d4_encode(const uint32_t *in, size_t n, uint32_t *out, size_t *nwords);
int orc_s4bp128d4_decode(const uint32_t *in, size_t nwords, uint32_t *out, size_t *n);
int orc_cpusimd_compress(const int64_t *in, size_t n, size_t threshold, uint32_t *out,
                         size_t *csize);
int orc_cpusimd_decompress(const uint32_t *in, size_t csize, size_t threshold, int64_t *out,
                           size_t *n);

/* ---- root selection (std::knuth_b + uniform_int_distribution<int64>) ---- */
int orc_select_roots(const int64_t *rowptr, int64_t n_rows, int64_t vertices, int nroots,
                     int64_t *roots);
void orc_knuth_b_draws(int64_t vertices, int count, int64_t *out);

/* ---- vreduce combines ([k][n] contiguous inputs) ---- */
void orc_vreduce_or(int k, const uint64_t *maps, int64_t nwords, uint64_t *out);
void orc_vreduce_min_pred(int k, const int64_t *preds, int64_t n, int64_t *out);

/* TEPS numerator: generated undirected edges with a reached endpoint */
int64_t orc_nedge(const int64_t *gdeg, const int32_t *level, int64_t n_rows);

#ifdef __cplusplus
}
#endif
#endif
