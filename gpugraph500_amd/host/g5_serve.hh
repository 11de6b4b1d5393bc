// g5_serve.hh -- BFS serving over several graphs resident in HBM, with
// continuous batching on the 64-lane multi-source search.
//
// One Engine per resident graph: its own g5_msbfs in continuous mode (own
// stream) driven by its own host thread, so searches on different graphs
// overlap on the device.  A query names a graph and 1..8 roots.  Admission is
// by lane: each lane is a preallocated memory slot of the engine (the
// int32[n] parent column plus one bit per vertex in the visited and frontier
// masks), and a query is admitted, FIFO, as soon as enough lanes are free for
// all of its roots -- between two levels of the searches already running,
// not when the batch ends -- provided every running lane is still within the
// join window (its first `window` levels).  Lanes that join later than that
// run out of step with the others: every level then holds some lane in its
// dense middle levels, which defeats the pull kernels' early exit for all of
// them, so outside the window new roots wait for the running cohort to
// drain (window < 0: join whenever lanes are free).  A root's lane leaves as
// soon as its own search is complete; its result is (nedge, checksum of the
// predecessor map), and is identical to the root searched alone.
// Small queries on an idle engine: a query of at most `small` roots that
// finds the engine idle with nothing queued behind it is answered by the
// direction-optimised single-root search, root after root (a few ms at
// scale 26), instead of a multi-lane search whose pull levels sweep every
// vertex for a handful of lanes; the results are the same.
// Admission by memory: with a MemoryBudget shared by the engines, a root in
// a lane holds that lane's bytes of HBM (g5_msbfs_footprint: its parent and
// level columns and its bits in the lane masks) against the budget from the
// moment it joins until its lane leaves, and a query is admitted only when
// the budget holds all of its roots' bytes (FIFO: later queries wait behind
// it); a graph is admitted at start-up only if its graph and search state fit
// the device's free memory (g5_mem_info).  The lanes' memory is allocated
// when the engine is created, so nothing is allocated while serving; the
// budget bounds how much of it is in use at once, across all graphs.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "g5hip.h"

namespace g5host {

struct Query {
    int id = 0;
    int graph = 0;
    std::vector<int64_t> roots;
    double arrival = 0;   // seconds since the serving start (the replay clock)
    // filled by the engine
    double admitted = -1, done = -1;
    std::vector<int64_t> nedge;
    std::vector<uint64_t> checksum;
    int remaining = 0;
};

// Bytes of HBM the running roots of all engines may hold at once.
class MemoryBudget
{
public:
    explicit MemoryBudget(int64_t limit) : limit_(limit) {}
    int64_t limit() const { return limit_; }
    // takes `bytes` if they fit, else nothing (returns false)
    bool try_take(int64_t bytes);
    void give(int64_t bytes);
    // waits until bytes were given back since `seen` (a count from gen()) or
    // the timeout passes
    void wait_for_release(uint64_t seen, double timeout_s);
    uint64_t gen() const;
    int64_t peak() const;

private:
    int64_t limit_;
    int64_t used_ = 0, peak_ = 0;
    uint64_t gen_ = 0;
    mutable std::mutex mu_;
    std::condition_variable cv_;
};

struct EngineStats {
    int64_t levels = 0, joins = 0, leaves = 0;
    int64_t split_levels = 0;  // levels that pulled some lanes and pushed the others
    int64_t single_roots = 0;  // roots answered by the single-root search
    int64_t budget_waits = 0;  // queries that had to wait because the memory budget was held
    double device_ms = 0;  // sum of the levels' device time
};

class Engine
{
public:
    // owns the graph (built on the device from the Kronecker generator);
    // budget (may be null: lanes only) is shared by the engines; place > 0:
    // try that many placements of the parent records at start
    // (g5_msbfs_place on the graph's search keys) and keep the fastest
    Engine(int scale, int edgefactor, int lanes, int window = 1, int small = 0, MemoryBudget *budget = nullptr,
           int place = 0);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    void start();                 // launch the engine thread
    void submit(Query *q);        // hand a query to the engine (thread-safe)
    void finish();                // no more queries: drain, join the thread, rethrow its error
    // the whole query alone on an idle engine (no thread), for reference results
    void run_alone(Query *q);

    int scale() const { return scale_; }
    int64_t vertices() const { return n_; }
    int64_t lane_bytes() const { return lane_bytes_; }
    int64_t device_bytes() const { return device_bytes_; }  // graph + search state held by the engine
    const std::vector<int64_t> &keys() const { return keys_; }  // search keys (degree > 0)
    EngineStats stats() const { return stats_; }
    double clock() const;         // seconds since the shared epoch
    void set_epoch(double t0) { epoch_ = t0; }

private:
    int scale_;
    int max_lanes_;
    int window_;
    int small_;
    int64_t level_ = 0;  // levels run by this engine
    int64_t n_ = 0;
    MemoryBudget *budget_ = nullptr;
    int64_t lane_bytes_ = 0, device_bytes_ = 0;
    const Query *blocked_ = nullptr;  // the query the budget last held back
    g5_graph *g_ = nullptr;
    g5_msbfs *m_ = nullptr;
    g5_bfs *b_ = nullptr;  // single-root search for small queries (small > 0)
    std::vector<int64_t> keys_;
    std::vector<int64_t> key_reach_;  // g5_msbfs_reach of each key
    int64_t reach_of(int64_t root) const;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Query *> pending_;
    bool stop_ = false;
    struct Slot {
        Query *q = nullptr;
        int idx = 0;
        int64_t joined = 0;  // level_ when the lane joined
    };
    Slot slots_[64];
    uint64_t busy_ = 0;
    EngineStats stats_;
    double epoch_ = 0;
    std::string error_;  // set when the engine thread stopped on an error

    void loop();
    void loop_body();
    // admit from `from` while lanes are free and the running lanes are within
    // the join window; returns the number of roots joined
    int admit(std::deque<Query *> &from);
    bool in_window() const;
    // answers the only queued query by single-root searches if it is small
    // and the engine idle; returns true if it did
    bool run_small(std::deque<Query *> &from);
    // one level plus the departures it causes; returns true while lanes are in use
    bool advance();
};

double now_seconds();

}  // namespace g5host
