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

Engine::Engine(int scale, int edgefactor, int lanes, int window, int small)
    : scale_(scale), max_lanes_(lanes), window_(window), small_(small)
{
    if (lanes < 1 || lanes > 64) throw std::invalid_argument("Engine: lanes must be in [1, 64]");
    check(g5_graph_kronecker(scale, edgefactor, 1, 2, &g_), "g5_graph_kronecker");
    int64_t nc = 0, nnz = 0, ng = 0, gen = 0;
    check(g5_graph_info(g_, &n_, &nc, &nnz, &ng, &gen), "g5_graph_info");
    keys_.resize(64);
    int found = 0;
    check(g5_select_roots(g_, (int64_t)1 << scale, 64, keys_.data(), &found), "g5_select_roots");
    keys_.resize(found);
    check(g5_msbfs_create(g_, &m_), "g5_msbfs_create");
    key_reach_.assign(keys_.size(), 0);
    check(g5_msbfs_reach(m_, keys_.data(), (int)keys_.size(), key_reach_.data()), "g5_msbfs_reach");
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
        admit(local);  // between two levels
        advance();
    }
}

void Engine::run_alone(Query *q)
{
    if (busy_) throw std::logic_error("run_alone on a busy engine");
    std::deque<Query *> one{q};
    admit(one);
    if (!one.empty()) throw std::invalid_argument("query does not fit the engine's lanes");
    while (advance()) {
    }
}

}  // namespace g5host
