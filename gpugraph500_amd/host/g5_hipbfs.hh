// g5_hipbfs.hh -- HipBFS: the GPU "Derived" of the reference's GlobalBFS
// (the hook set of CUDA_BFS, cpu_2d/cuda/cuda_bfs.h:68-83) implemented over
// the gfx950 C ABI (include/g5hip.h).
//
// One HipBFS is one rank's share: a DistMatrix2d block resident in HBM
// (g5_graph).  Queues crossing the hooks are host int64 arrays of local
// column ids, sorted and unique, as the reference's CUDA_BFS hands them to
// vreduce and the row fold (cuda_bfs.cu:155-321).  Buffers returned by
// getOutgoingFQ are owned by the HipBFS and stay valid until its next hook
// call.  Failures of the C ABI are thrown as std::runtime_error (the
// reference aborts through B40CPerror -> MPI_Abort).
#pragma once

#include <cstdint>
#include <vector>

#include "g5hip.h"

namespace g5host {

class HipBFS
{
public:
    typedef int64_t vertexType;

    // the block stays owned by the caller and must outlive the HipBFS
    explicit HipBFS(g5_graph *block);
    ~HipBFS();
    HipBFS(const HipBFS &) = delete;
    HipBFS &operator=(const HipBFS &) = delete;

    // ---- GlobalBFS hooks (cuda/cuda_bfs.h:68-83) ----
    void setStartVertex(vertexType start);  // start = -1 on ranks that own neither its row nor its column
    void runLocalBFS();
    bool istheresomethingnew();
    void getBackOutqueue();
    void reduce_fq_out(vertexType globalstart, long size, vertexType *startaddr, int insize);
    void getOutgoingFQ(vertexType *&startaddr, int &outsize);
    void getOutgoingFQ(vertexType globalstart, long size, vertexType *&startaddr, int &outsize);
    void setModOutgoingFQ(vertexType *startaddr, int insize);  // startaddr = 0: keep the own out-queue
    void setIncommingFQ(vertexType globalstart, long size, vertexType *startaddr, int &insize_max);
    void setBackInqueue();
    void getBackPredecessor();  // predecessor[n_cols]: global ids, -1 unreached here, root -> root
    void generatOwenMask() {}
    void bfsMemCpy(vertexType *&dst, vertexType *src, size_t n);

    std::vector<vertexType> predecessor;

    // ---- the same steps with the frontiers left in HBM ----
    const uint64_t *outBitmap(int64_t &n_words);
    void setModOutgoingBitmap(const uint64_t *d_bits);
    void setIncomingBitmap(const uint64_t *d_bits);
    // runLocalBFS as a bottom-up level over the transposed block (row frontier
    // from setIncomingBitmap)
    void runLocalBFSBottomUp(const g5_graph *transposed);
    void predecessorDevice(int64_t *d_pred);

    g5_bfs *handle() { return b_; }
    int64_t rows() const { return n_rows_; }
    int64_t cols() const { return n_cols_; }
    int64_t rowStart() const { return row_start_; }
    int64_t colStart() const { return col_start_; }

private:
    g5_graph *g_ = nullptr;
    g5_bfs *b_ = nullptr;
    int64_t n_rows_ = 0, n_cols_ = 0, row_start_ = 0, col_start_ = 0;
    bool something_new_ = false;
    std::vector<vertexType> queuebuff_, redbuff_, fq_64_;  // out-queue, reduction scratch, exchange buffer
    int64_t qb_length_ = 0;
    std::vector<vertexType> inqueue_;
    int64_t in_length_ = 0;
    std::vector<uint32_t> q32_;
};

}  // namespace g5host
