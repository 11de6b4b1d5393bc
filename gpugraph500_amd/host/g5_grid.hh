// g5_grid.hh -- the GlobalBFS level loop (cpu_2d/globalbfs.hh:548-1263) for
// an R x C grid of HipBFS ranks emulated in one process on one GPU.
//
// Each rank is a HipBFS over its own DistMatrix2d block.  The reference runs
// one MPI process per rank; here the ranks advance in lock-step and the two
// exchange steps of a level are done in-process:
//   column step  the union of the column group's out-queues (vreduce,
//                vreduce.hpp:41-544), adopted by every rank of the column
//                (setModOutgoingFQ)
//   row step     the diagonal rank's reduced queue, handed to every rank of
//                its row (the fold, globalbfs.hh:816-967 with symmetric
//                partitioning: setIncommingFQ + setBackInqueue)
// and the search ends with the predecessor combine per column (the minimum
// over the column's contributions, the canonical rule that replaces the
// reference's rank-order precedence of allReduceBitCompressed,
// globalbfs.hh:235-524).
//
// Two exchange modes:
//   Exchange::HostQueues  the reference's hook protocol: host int64 queues,
//                         every message through a Compression schema
//                         (compress -> decompress) as it would cross the wire
//   Exchange::Device      bitmaps stay in HBM: OR of the column's
//                         out-bitmaps with g5_vreduce_or_list, the fold by device
//                         bitmap, optionally through the GPU codec (bitmap ->
//                         queue -> s4-bp128-d4 encode -> decode -> bitmap),
//                         and g5_vreduce_min_pred for the predecessors
// Both give the same predecessors as the 1 x 1 search (minimum-id parents).
// With Exchange::Device the dense levels can run bottom-up (bottom_up): rank
// (r, c) scans its unvisited columns' neighbours in block (c, r), the
// transposed block, and stops at the first vertex of its row frontier; the
// switch follows the global count of newly reached vertices.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "g5_compression.hh"
#include "g5_hipbfs.hh"

namespace g5host {

enum class Exchange { HostQueues, Device };

struct GridStats {
    int levels = 0;
    int bottom_up_levels = 0;
    double total_s = 0;      // whole emulated search (all ranks in sequence)
    double expand_s = 0;     // runLocalBFS over all ranks
    double column_s = 0;     // column step (reduce + adopt) over all columns
    double fold_s = 0;       // row step over all rows
    double pred_s = 0;       // predecessor combine
    double codec_s = 0;      // encode + decode inside the exchange steps
    double share_expand_s = 0;  // the watched rank's runLocalBFS
    double share_column_s = 0;  // its column's reduce + adopt
    double share_fold_s = 0;    // its row's fold
    int64_t raw_words = 0, coded_words = 0;  // exchanged queue sizes before / after the codec
};

class Grid2D
{
public:
    // blocks[r * C + c] as g5_graph_kronecker_2d returns them (owned by the
    // caller); R == C (the fold assumes symmetric partitioning, as the CUDA
    // path does).  schema may be null (raw transfer).
    Grid2D(int R, int C, g5_graph **blocks, Exchange mode, Compression<int64_t, uint32_t> *schema, bool device_codec,
           bool bottom_up = false);
    ~Grid2D();

    // One search; pred receives nglobal() global predecessors (-1 unreached,
    // root -> root).  (r0, c0) is the rank whose share is timed separately.
    void runBFS(int64_t root, std::vector<int64_t> &pred, GridStats &st, int r0 = 0, int c0 = 0);

    // generated undirected edges with a reached endpoint (validate.h:439-485)
    int64_t nedge(const std::vector<int64_t> &pred) const;

    int64_t nglobal() const { return n_total_; }
    HipBFS &rank(int r, int c) { return *ranks_[r * C_ + c]; }

private:
    int R_, C_;
    Exchange mode_;
    Compression<int64_t, uint32_t> *schema_;
    bool device_codec_;
    bool bottom_up_;
    std::vector<g5_graph *> blocks_;  // [r * C + c], not owned
    int64_t last_new_ = 0;            // vertices reached by the last level (bottom_up_)
    std::vector<std::unique_ptr<HipBFS>> ranks_;
    std::vector<int64_t> col_start_, col_len_;
    int64_t n_total_ = 0, max_cols_ = 0;
    std::vector<int64_t> gdeg_;  // generated degree per global vertex (sum over a row of blocks)
    // device scratch for Exchange::Device
    uint64_t *d_reduced_ = nullptr;   // C x max_cols/64 words
    uint64_t *d_fold_ = nullptr;      // max_cols/64 words
    int64_t *d_preds_ = nullptr;      // R x max_cols
    int64_t *d_pred_out_ = nullptr;   // max_cols
    uint32_t *d_q_ = nullptr, *d_words_ = nullptr, *d_q2_ = nullptr;

    void exchange_host(GridStats &st, int r0, int c0);
    void exchange_device(GridStats &st, int r0, int c0);
    void send(int64_t *q, int n, std::vector<int64_t> &out, GridStats &st);
};

}  // namespace g5host
