// g5_serve.cpp -- the per-graph serving engine (see g5_serve.hh).
#include "g5_serve.hh"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace g5host {

namespace {

void check(int rc, const char *what)
{
    if (rc != G5_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

int popc(uint64_t x) { return __builtin_popcountll(x); }

}  // namespace

double now_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool MemoryBudget::try_take(int64_t bytes)
{
    std::lock_guard<std::mutex> lk(mu_);
    if (used_ + bytes > limit_) return false;
    used_ += bytes;
    peak_ = std::max(peak_, used_);
    return true;
}

void MemoryBudget::give(int64_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= bytes;
        ++gen_;
    }
    cv_.notify_all();
}

uint64_t MemoryBudget::gen() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return gen_;
}

int64_t MemoryBudget::peak() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return peak_;
}

void MemoryBudget::wait_for_release(uint64_t seen, double timeout_s)
{
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return gen_ != seen; });
}

Engine::Engine(int scale, int edgefactor, int lanes, int window, int small, MemoryBudget *budget, int place)
    : scale_(scale), max_lanes_(lanes), window_(window), small_(small), budget_(budget)
{
    if (lanes < 1 || lanes > 64) throw std::invalid_argument("Engine: lanes must be in [1, 64]");
    // graph admission: the graph (its generation's peak, the edge list and
    // its sorted copy included, is bounded by the same estimate) and its
    // search state must fit the device's free memory
    {
        const int64_t nv = (int64_t)1 << scale, m = (int64_t)edgefactor << scale;
        const int64_t need = 8 * (nv + 1) + 4 * 2 * m + 4 * nv     // the graph
                             + 8 * (nv + 1) + 4 * 4 * m            // its degree-ordered copy (two id arrays)
                             + 64 * 5 * nv + 3 * 8 * nv + 40 * nv;  // records, masks, queues, ids
        int64_t free_b = 0, total_b = 0;
        check(g5_mem_info(&free_b, &total_b), "g5_mem_info");
        if (need > free_b)
            throw std::runtime_error("graph of scale " + std::to_string(scale) + " not admitted: it needs about " +
                                     std::to_string(need >> 20) + " MiB of HBM, " + std::to_string(free_b >> 20) +
                                     " MiB are free");
    }
    check(g5_graph_kronecker(scale, edgefactor, 1, 2, &g_), "g5_graph_kronecker");
    int64_t nc = 0, nnz = 0, ng = 0, gen = 0;
    check(g5_graph_info(g_, &n_, &nc, &nnz, &ng, &gen), "g5_graph_info");
    int64_t graph_b = 0, fixed_b = 0;
    check(g5_msbfs_footprint(g_, &graph_b, &fixed_b, &lane_bytes_), "g5_msbfs_footprint");
    device_bytes_ = graph_b + fixed_b + 64 * lane_bytes_;
    keys_.resize(64);
    int found = 0;
    check(g5_select_roots(g_, (int64_t)1 << scale, 64, keys_.data(), &found), "g5_select_roots");
    keys_.resize(found);
    check(g5_msbfs_create(g_, &m_), "g5_msbfs_create");
    key_reach_.assign(keys_.size(), 0);
    check(g5_msbfs_reach(m_, keys_.data(), (int)keys_.size(), key_reach_.data()), "g5_msbfs_reach");
    if (place > 0 && !keys_.empty()) {
        float first = 0, best = 0;
        int moves = 0;
        check(g5_msbfs_place(m_, keys_.data(), (int)std::min<size_t>(keys_.size(), 64), place, &first, &best, &moves),
              "g5_msbfs_place");
    }
    check(g5_msbfs_open(m_), "g5_msbfs_open");
    // allocated up front: a device allocation in the middle of serving can
    // stall the other engines' streams
    if (small_ > 0) check(g5_bfs_create(g_, &b_), "g5_bfs_create");
}

Engine::~Engine()
{
    if (th_.joinable()) {
        try {
            finish();
        } catch (const std::exception &) {
        }
    }
    g5_msbfs_free(m_);
    g5_bfs_free(b_);
    g5_graph_free(g_);
}

double Engine::clock() const { return now_seconds() - epoch_; }

void Engine::start() { th_ = std::thread(&Engine::loop, this); }

void Engine::submit(Query *q)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        pending_.push_back(q);
    }
    cv_.notify_one();
}

void Engine::finish()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_one();
    if (th_.joinable()) th_.join();
    if (!error_.empty()) throw std::runtime_error("engine (scale " + std::to_string(scale_) + "): " + error_);
}

bool Engine::in_window() const
{
    if (window_ < 0) return true;
    for (int lane = 0; lane < 64; ++lane)
        if ((busy_ >> lane & 1) && level_ - slots_[lane].joined > window_) return false;
    return true;
}

int64_t Engine::reach_of(int64_t root) const
{
    for (size_t i = 0; i < keys_.size(); ++i)
        if (keys_[i] == root) return key_reach_[i];
    return 0;
}

int Engine::admit(std::deque<Query *> &from)
{
    if (!in_window()) return 0;  // wait for the running cohort
    const uint64_t allowed = max_lanes_ == 64 ? ~0ull : ((1ull << max_lanes_) - 1);
    // the queries that fit, FIFO, and their roots
    struct Pending {
        Query *q;
        int idx;
        int64_t reach;
    };
    std::vector<Pending> adm;
    int free_lanes = popc(allowed & ~busy_);
    while (!from.empty()) {
        Query *q = from.front();
        const int need = (int)q->roots.size();
        if (need > max_lanes_) throw std::invalid_argument("query with more roots than lanes");
        if (free_lanes < need) break;  // FIFO: later queries wait behind it
        if (budget_ && need * lane_bytes_ > budget_->limit())
            throw std::invalid_argument("query whose roots need more memory than the whole budget");
        if (budget_ && !budget_->try_take(need * lane_bytes_)) {
            if (blocked_ != q) ++stats_.budget_waits;  // its roots' memory is not free yet
            blocked_ = q;
            break;
        }
        from.pop_front();
        free_lanes -= need;
        q->admitted = clock();
        q->remaining = need;
        q->nedge.assign(need, 0);
        q->checksum.assign(need, 0);
        for (int i = 0; i < need; ++i) adm.push_back({q, i, reach_of(q->roots[i])});
    }
    // lanes in ascending order by descending reach, as a batch run assigns
    // them: roots of similar reach share parent-record sectors
    std::stable_sort(adm.begin(), adm.end(), [](const Pending &a, const Pending &b) { return a.reach > b.reach; });
    std::vector<int64_t> roots;
    std::vector<int> lanes;
    for (const Pending &a : adm) {
        const int lane = __builtin_ctzll(allowed & ~busy_);
        busy_ |= 1ull << lane;
        slots_[lane].q = a.q;
        slots_[lane].idx = a.idx;
        slots_[lane].joined = level_;
        roots.push_back(a.q->roots[a.idx]);
        lanes.push_back(lane);
    }
    if (!roots.empty()) {
        check(g5_msbfs_join(m_, roots.data(), lanes.data(), (int)roots.size()), "g5_msbfs_join");
        stats_.joins += (int64_t)roots.size();
    }
    return (int)roots.size();
}

bool Engine::run_small(std::deque<Query *> &from)
{
    if (busy_ || small_ <= 0 || from.size() != 1 || (int)from.front()->roots.size() > small_) return false;
    Query *q = from.front();
    from.pop_front();
    q->admitted = clock();
    const size_t k = q->roots.size();
    q->nedge.assign(k, 0);
    q->checksum.assign(k, 0);
    for (size_t i = 0; i < k; ++i) {
        int32_t nl = 0;
        float ms = 0;
        check(g5_bfs_run(b_, q->roots[i], 0, &nl, &ms), "g5_bfs_run");
        check(g5_bfs_summary(b_, &q->nedge[i], &q->checksum[i]), "g5_bfs_summary");
        stats_.device_ms += ms;
    }
    stats_.single_roots += (int64_t)k;
    q->remaining = 0;
    q->done = clock();
    return true;
}

bool Engine::advance()
{
    if (!busy_) return false;
    uint64_t live = 0;
    float ms = 0;
    check(g5_msbfs_step(m_, &live, &ms), "g5_msbfs_step");
    int64_t info[4] = {0, 0, 0, 0};
    check(g5_msbfs_info(m_, info, 4), "g5_msbfs_info");
    stats_.split_levels = info[3];
    stats_.levels += 1;
    level_ += 1;
    stats_.device_ms += ms;
    const uint64_t finished = busy_ & ~live;
    if (finished) {
        int64_t ne[64];
        uint64_t ck[64];
        check(g5_msbfs_leave(m_, finished, ne, ck), "g5_msbfs_leave");
        const double t = clock();
        for (int lane = 0; lane < 64; ++lane) {
            if (!(finished >> lane & 1)) continue;
            Slot &sl = slots_[lane];
            sl.q->nedge[sl.idx] = ne[lane];
            sl.q->checksum[sl.idx] = ck[lane];
            if (--sl.q->remaining == 0) sl.q->done = t;
            sl.q = nullptr;
        }
        busy_ &= ~finished;
        stats_.leaves += popc(finished);
        if (budget_) budget_->give(popc(finished) * lane_bytes_);
    }
    return busy_ != 0;
}

void Engine::loop()
{
    try {
        loop_body();
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu_);
        error_ = e.what();
    }
}

void Engine::loop_body()
{
    std::deque<Query *> local;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (busy_ == 0 && local.empty())
                cv_.wait(lk, [&] { return stop_ || !pending_.empty(); });
            while (!pending_.empty()) {
                local.push_back(pending_.front());
                pending_.pop_front();
            }
            if (stop_ && busy_ == 0 && local.empty()) return;
        }
        if (run_small(local)) continue;
        const uint64_t seen = budget_ ? budget_->gen() : 0;
        const int joined = admit(local);  // between two levels
        if (!advance() && !joined && !local.empty() && budget_) {
            // idle, with queries held back by memory that other engines hold:
            // wait until some of it is given back (or new queries arrive)
            budget_->wait_for_release(seen, 0.002);
        }
    }
}

void Engine::run_alone(Query *q)
{
    if (busy_) throw std::logic_error("run_alone on a busy engine");
    std::deque<Query *> one{q};
    MemoryBudget *b = budget_;
    budget_ = nullptr;  // a reference run by itself: no other query holds memory
    admit(one);
    budget_ = b;
    if (!one.empty()) throw std::invalid_argument("query does not fit the engine's lanes");
    while (advance()) {
    }
}

}  // namespace g5host
