// g5_hipbfs.cpp -- HipBFS hooks over the C ABI (see g5_hipbfs.hh).
#include "g5_hipbfs.hh"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace g5host {

namespace {

void check(int rc, const char *what)
{
    if (rc != G5_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

}  // namespace

HipBFS::HipBFS(g5_graph *block) : g_(block)
{
    if (!block) throw std::invalid_argument("HipBFS: null block");
    int64_t nnz = 0, nglobal = 0, gen = 0;
    check(g5_graph_info(block, &n_rows_, &n_cols_, &nnz, &nglobal, &gen), "g5_graph_info");
    check(g5_graph_bounds(block, &row_start_, &col_start_), "g5_graph_bounds");
    check(g5_bfs_create(block, &b_), "g5_bfs_create");
    queuebuff_.resize(n_cols_);
    redbuff_.resize(n_cols_);
    fq_64_.resize(std::max(n_rows_, n_cols_));
    inqueue_.resize(n_rows_);
    q32_.resize(std::max(n_rows_, n_cols_));
    predecessor.assign(n_cols_, -1);
}

HipBFS::~HipBFS() { g5_bfs_free(b_); }

void HipBFS::setStartVertex(vertexType start)
{
    check(g5_bfs_start(b_, start), "g5_bfs_start");
    something_new_ = false;
    qb_length_ = 0;
    in_length_ = 0;
}

void HipBFS::runLocalBFS()
{
    int any = 0;
    check(g5_bfs_expand(b_, &any), "g5_bfs_expand");
    something_new_ = any != 0;
}

void HipBFS::runLocalBFSBottomUp(const g5_graph *transposed)
{
    int any = 0;
    check(g5_bfs_expand_bottom_up(b_, transposed, &any), "g5_bfs_expand_bottom_up");
    something_new_ = any != 0;
}

bool HipBFS::istheresomethingnew() { return something_new_; }

// cuda_bfs.cu:380-475: the column out-queue, sorted and unique, host-visible
void HipBFS::getBackOutqueue()
{
    int64_t n = n_cols_;
    check(g5_bfs_outqueue(b_, q32_.data(), &n), "g5_bfs_outqueue");
    for (int64_t i = 0; i < n; ++i) queuebuff_[i] = q32_[i];
    qb_length_ = n;
}

// cuda_bfs.cu:155-218: the own queue's part in [globalstart, globalstart +
// size) united with the incoming sorted queue becomes the own queue
void HipBFS::reduce_fq_out(vertexType globalstart, long size, vertexType *startaddr, int insize)
{
    vertexType *b = queuebuff_.data(), *e = b + qb_length_;
    vertexType *lo = std::lower_bound(b, e, globalstart);
    vertexType *hi = std::lower_bound(lo, e, globalstart + size);
    vertexType *end = std::set_union(lo, hi, startaddr, startaddr + insize, redbuff_.data());
    qb_length_ = end - redbuff_.data();
    std::swap(queuebuff_, redbuff_);
}

void HipBFS::getOutgoingFQ(vertexType *&startaddr, int &outsize)
{
    startaddr = queuebuff_.data();
    outsize = (int)qb_length_;
}

void HipBFS::getOutgoingFQ(vertexType globalstart, long size, vertexType *&startaddr, int &outsize)
{
    vertexType *b = queuebuff_.data(), *e = b + qb_length_;
    vertexType *lo = std::lower_bound(b, e, globalstart);
    vertexType *hi = std::lower_bound(lo, e, globalstart + size);
    startaddr = lo;
    outsize = (int)(hi - lo);
}

// cuda_bfs.cu:230-270: adopt the column-reduced queue, update visited
void HipBFS::setModOutgoingFQ(vertexType *startaddr, int insize)
{
    if (startaddr) {
        if (insize > n_cols_) throw std::length_error("HipBFS::setModOutgoingFQ: queue longer than the column slice");
        if (startaddr != queuebuff_.data()) std::copy(startaddr, startaddr + insize, queuebuff_.data());
        qb_length_ = insize;
    }
    for (int64_t i = 0; i < qb_length_; ++i) q32_[i] = (uint32_t)queuebuff_[i];
    check(g5_bfs_set_outqueue(b_, q32_.data(), qb_length_), "g5_bfs_set_outqueue");
}

// cuda_bfs.cu:275-321 (symmetric partitioning: the range is the whole row slice)
void HipBFS::setIncommingFQ(vertexType, long, vertexType *startaddr, int &insize_max)
{
    if (insize_max > n_rows_) throw std::length_error("HipBFS::setIncommingFQ: queue longer than the row slice");
    if (startaddr != inqueue_.data()) std::copy(startaddr, startaddr + insize_max, inqueue_.data());
    in_length_ = insize_max;
}

// cuda_bfs.cu:477-545
void HipBFS::setBackInqueue()
{
    for (int64_t i = 0; i < in_length_; ++i) q32_[i] = (uint32_t)inqueue_[i];
    check(g5_bfs_set_inqueue(b_, q32_.data(), in_length_), "g5_bfs_set_inqueue");
}

// cuda_bfs.cu:328-378
void HipBFS::getBackPredecessor() { check(g5_bfs_result(b_, predecessor.data(), nullptr), "g5_bfs_result"); }

void HipBFS::bfsMemCpy(vertexType *&dst, vertexType *src, size_t n)
{
    if (n > fq_64_.size()) throw std::length_error("HipBFS::bfsMemCpy: exceeds the exchange buffer");
    std::copy(src, src + n, fq_64_.data());
    dst = fq_64_.data();
}

const uint64_t *HipBFS::outBitmap(int64_t &n_words)
{
    const uint64_t *p = nullptr;
    check(g5_bfs_out_bitmap(b_, &p, &n_words), "g5_bfs_out_bitmap");
    return p;
}

void HipBFS::setModOutgoingBitmap(const uint64_t *d_bits)
{
    check(g5_bfs_set_out_bitmap(b_, d_bits), "g5_bfs_set_out_bitmap");
}

void HipBFS::setIncomingBitmap(const uint64_t *d_bits)
{
    check(g5_bfs_set_in_bitmap(b_, d_bits), "g5_bfs_set_in_bitmap");
}

void HipBFS::predecessorDevice(int64_t *d_pred) { check(g5_bfs_pred_device(b_, d_pred), "g5_bfs_pred_device"); }

}  // namespace g5host
