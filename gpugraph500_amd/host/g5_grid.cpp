// g5_grid.cpp -- the emulated R x C GlobalBFS level loop (see g5_grid.hh).
#include "g5_grid.hh"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace g5host {

namespace {

void check(int rc, const char *what)
{
    if (rc != G5_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename X>
X *dalloc(size_t n)
{
    void *p = nullptr;
    check(g5_dmalloc(&p, (n ? n : 1) * sizeof(X)), "g5_dmalloc");
    return static_cast<X *>(p);
}

}  // namespace

Grid2D::Grid2D(int R, int C, g5_graph **blocks, Exchange mode, Compression<int64_t, uint32_t> *schema,
               bool device_codec, bool bottom_up)
    : R_(R), C_(C), mode_(mode), schema_(schema), device_codec_(device_codec),
      bottom_up_(bottom_up && mode == Exchange::Device), blocks_(blocks, blocks + (size_t)R * C)
{
    if (R < 1 || C < 1 || R != C) throw std::invalid_argument("Grid2D: needs R == C >= 1 (symmetric partitioning)");
    for (int i = 0; i < R * C; ++i) ranks_.emplace_back(new HipBFS(blocks[i]));
    col_start_.resize(C);
    col_len_.resize(C);
    for (int c = 0; c < C; ++c) {
        col_start_[c] = rank(0, c).colStart();
        col_len_[c] = rank(0, c).cols();
        max_cols_ = std::max(max_cols_, col_len_[c]);
        for (int r = 0; r < R; ++r)
            if (rank(r, c).colStart() != col_start_[c] || rank(r, c).cols() != col_len_[c])
                throw std::invalid_argument("Grid2D: blocks of a column disagree on its bounds");
    }
    for (int r = 0; r < R; ++r)
        if (rank(r, r).rowStart() != col_start_[r] || rank(r, r).rows() != col_len_[r])
            throw std::invalid_argument("Grid2D: row part r must equal column part r");
    n_total_ = col_start_[C - 1] + col_len_[C - 1];
    gdeg_.assign(n_total_, 0);
    std::vector<int64_t> t;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            HipBFS &h = rank(r, c);
            t.resize(h.rows());
            check(g5_graph_gdeg(blocks[r * C + c], t.data()), "g5_graph_gdeg");
            for (int64_t i = 0; i < h.rows(); ++i) gdeg_[h.rowStart() + i] += t[i];
        }
    if (mode_ == Exchange::Device) {
        const int64_t w = max_cols_ / 64;
        d_reduced_ = dalloc<uint64_t>((size_t)C * w);
        d_fold_ = dalloc<uint64_t>(w);
        d_preds_ = dalloc<int64_t>((size_t)R * max_cols_);
        d_pred_out_ = dalloc<int64_t>(max_cols_);
        if (device_codec_) {
            d_q_ = dalloc<uint32_t>(max_cols_);
            d_q2_ = dalloc<uint32_t>(max_cols_);
            d_words_ = dalloc<uint32_t>(g5_s4_max_words(max_cols_));
        }
    }
}

Grid2D::~Grid2D()
{
    void *bufs[] = {d_reduced_, d_fold_, d_preds_, d_pred_out_, d_q_, d_words_, d_q2_};
    for (void *p : bufs)
        if (p) g5_dfree(p);
}

// one message: through the schema (compress, then decompress at the
// receiver) or a plain copy
void Grid2D::send(int64_t *q, int n, std::vector<int64_t> &out, GridStats &st)
{
    st.raw_words += n;
    if (!schema_) {
        out.assign(q, q + n);
        st.coded_words += n;
        return;
    }
    const double t0 = now();
    uint32_t *coded = nullptr;
    size_t csize = 0;
    schema_->compress(q, (size_t)n, &coded, csize);
    int64_t *back = nullptr;
    size_t m = (size_t)n;
    schema_->decompress(coded, (int)csize, &back, m);
    out.assign(back, back + m);
    free(coded);
    free(back);
    st.coded_words += (int64_t)csize;
    st.codec_s += now() - t0;
}

void Grid2D::exchange_host(GridStats &st, int r0, int c0)
{
    std::vector<int64_t> msg;
    // column step: union into the column's first rank, then every rank adopts it
    for (int c = 0; c < C_; ++c) {
        const double t0 = now();
        HipBFS &acc = rank(0, c);
        for (int r = 0; r < R_; ++r) rank(r, c).getBackOutqueue();
        for (int r = 1; r < R_; ++r) {
            int64_t *p = nullptr;
            int n = 0;
            rank(r, c).getOutgoingFQ(p, n);
            send(p, n, msg, st);
            acc.reduce_fq_out(0, (long)col_len_[c], msg.data(), (int)msg.size());
        }
        int64_t *p = nullptr;
        int n = 0;
        acc.getOutgoingFQ(p, n);
        send(p, n, msg, st);  // the allgather of the reduced queue
        for (int r = 0; r < R_; ++r) {
            int64_t *dst = nullptr;
            rank(r, c).bfsMemCpy(dst, msg.data(), msg.size());
            rank(r, c).setModOutgoingFQ(dst, (int)msg.size());
        }
        const double dt = now() - t0;
        st.column_s += dt;
        if (c == c0) st.share_column_s += dt;
    }
    // row step: the diagonal rank's reduced queue to every rank of its row
    for (int r = 0; r < R_; ++r) {
        const double t0 = now();
        int64_t *p = nullptr;
        int n = 0;
        rank(r, r).getOutgoingFQ(0, (long)col_len_[r], p, n);
        send(p, n, msg, st);
        for (int c = 0; c < C_; ++c) {
            int len = (int)msg.size();
            rank(r, c).setIncommingFQ(0, (long)col_len_[r], msg.data(), len);
            rank(r, c).setBackInqueue();
        }
        const double dt = now() - t0;
        st.fold_s += dt;
        if (r == r0) st.share_fold_s += dt;
    }
}

void Grid2D::exchange_device(GridStats &st, int r0, int c0)
{
    const int64_t wmax = max_cols_ / 64;
    for (int c = 0; c < C_; ++c) {
        const double t0 = now();
        const int64_t w = col_len_[c] / 64;
        // the column's out-bitmaps where the ranks keep them (an all-gather
        // into one buffer with RCCL)
        std::vector<const uint64_t *> maps(R_);
        for (int r = 0; r < R_; ++r) {
            int64_t nw = 0;
            maps[r] = rank(r, c).outBitmap(nw);
        }
        uint64_t *red = d_reduced_ + (size_t)c * wmax;
        check(g5_vreduce_or_list(maps.data(), R_, w, red), "g5_vreduce_or_list");
        if (bottom_up_) {
            int64_t k = 0;
            check(g5_bitmap_count(red, w, &k), "g5_bitmap_count");
            last_new_ += k;
        }
        for (int r = 0; r < R_; ++r) rank(r, c).setModOutgoingBitmap(red);
        st.raw_words += 2 * w * (R_ - 1);  // 32-bit words of the bitmaps gathered
        st.coded_words += 2 * w * (R_ - 1);
        const double dt = now() - t0;
        st.column_s += dt;
        if (c == c0) st.share_column_s += dt;
    }
    for (int r = 0; r < R_; ++r) {
        const double t0 = now();
        const int64_t w = col_len_[r] / 64;
        const uint64_t *src = d_reduced_ + (size_t)r * wmax;
        if (device_codec_) {
            const double tc = now();
            int64_t n = 0;
            check(g5_bitmap_to_queue(src, w, d_q_, &n), "g5_bitmap_to_queue");
            size_t words = 0;
            check(g5_s4_roundtrip(d_q_, (size_t)n, d_words_, g5_s4_max_words((size_t)n), d_q2_, &words),
                  "g5_s4_roundtrip");
            check(g5_queue_to_bitmap(d_q2_, n, d_fold_, w), "g5_queue_to_bitmap");
            src = d_fold_;
            st.raw_words += n;
            st.coded_words += (int64_t)words;
            st.codec_s += now() - tc;
        }
        for (int c = 0; c < C_; ++c) rank(r, c).setIncomingBitmap(src);
        const double dt = now() - t0;
        st.fold_s += dt;
        if (r == r0) st.share_fold_s += dt;
    }
}

void Grid2D::runBFS(int64_t root, std::vector<int64_t> &pred, GridStats &st, int r0, int c0)
{
    if (root < 0 || root >= n_total_) throw std::out_of_range("Grid2D::runBFS: root out of range");
    st = GridStats();
    const double t_begin = now();
    for (auto &h : ranks_) h->setStartVertex(root);
    bool bu = false;  // the first level is top-down
    for (;;) {
        bool any = false;
        for (int r = 0; r < R_; ++r)
            for (int c = 0; c < C_; ++c) {
                const double t0 = now();
                if (bu) rank(r, c).runLocalBFSBottomUp(blocks_[(size_t)c * C_ + r]);
                else rank(r, c).runLocalBFS();
                const double dt = now() - t0;
                st.expand_s += dt;
                if (r == r0 && c == c0) st.share_expand_s += dt;
                any = any || rank(r, c).istheresomethingnew();
            }
        if (!any) break;  // the allreduce of istheresomethingnew (globalbfs.hh:716-739)
        st.bottom_up_levels += bu ? 1 : 0;
        last_new_ = 0;
        if (mode_ == Exchange::HostQueues) exchange_host(st, r0, c0);
        else exchange_device(st, r0, c0);
        ++st.levels;
        // next direction: bottom-up once a level reaches more than n / 64
        // vertices, top-down again below n / 24 (the 1 x 1 search's return rule)
        if (bottom_up_) {
            if (!bu && last_new_ > n_total_ / 64) bu = true;
            else if (bu && last_new_ < n_total_ / 24) bu = false;
        }
    }
    // predecessors: per column, the minimum over the column's contributions
    const double t0 = now();
    pred.assign(n_total_, -1);
    for (int c = 0; c < C_; ++c) {
        const int64_t n = col_len_[c];
        if (mode_ == Exchange::HostQueues) {
            for (int r = 0; r < R_; ++r) {
                HipBFS &h = rank(r, c);
                h.getBackPredecessor();
                for (int64_t i = 0; i < n; ++i) {
                    const int64_t p = h.predecessor[i];
                    int64_t &d = pred[col_start_[c] + i];
                    if (p >= 0 && (d < 0 || p < d)) d = p;
                }
            }
        } else {
            for (int r = 0; r < R_; ++r) rank(r, c).predecessorDevice(d_preds_ + (size_t)r * n);
            check(g5_vreduce_min_pred(d_preds_, R_, n, d_pred_out_), "g5_vreduce_min_pred");
            check(g5_memcpy_d2h(pred.data() + col_start_[c], d_pred_out_, (size_t)n * 8), "g5_memcpy_d2h");
        }
    }
    st.pred_s = now() - t0;
    st.total_s = now() - t_begin;
}

int64_t Grid2D::nedge(const std::vector<int64_t> &pred) const
{
    int64_t s = 0;
    for (int64_t v = 0; v < n_total_ && v < (int64_t)pred.size(); ++v)
        if (pred[v] >= 0) s += gdeg_[v];
    return s / 2;
}

}  // namespace g5host
