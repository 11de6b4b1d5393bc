// g500_serve -- BFS serving over several resident Kronecker graphs with
// continuous batching (g5_serve.hh), replaying a seeded Poisson stream of
// queries in real time.
//
//   g500_serve [-g 18,20,22,24,25,26] [-e 16] [-rate q/s] [-nq queries] [-seed s]
//              [-maxroots k] [-lanes l] [-window w] [-small k] [-budget MiB] [-place t] [-alone]
//              [-compare] [-dump file]
//
//   -rate      mean arrival rate; inter-arrival times are exponential
//   -maxroots  roots per query drawn uniformly from 1..k (default 8)
//   -lanes     lanes (memory slots) per graph engine, at most 64
//   -window    new roots join a running cohort only while all its lanes are
//              within their first w levels (default 1; -1: whenever lanes
//              are free)
//   -small     a query of at most k roots that finds its engine idle and
//              nothing queued behind it runs as single-root searches
//              (default 4; 0: always the multi-lane search)
//   -budget    HBM (MiB, fractions allowed) that the running roots of all
//              graphs may hold at once: a root holds its lane's bytes
//              (g5_msbfs_footprint) from joining to leaving, and a query
//              waits until its roots' bytes are free (default: no budget,
//              lanes only).  Graphs are admitted at start only if they fit
//              the device's free memory.
//   -place     at start, each engine tries t fresh placements of its parent
//              records and keeps the one whose search of its 64 keys is
//              fastest (g5_msbfs_place; default 0: none)
//   -alone     every query by itself on an idle engine, one after another
//   -compare   after the served run, run every query alone and compare each
//              root's (nedge, checksum): results must be identical
//   -dump f    per query: "id graph arrival done nroots" then per root
//              "root nedge checksum" (one line per query)
// Output: one JSON line (latency p50/p90/p99/max in ms from arrival to the
// last root's result, queries/s, GTEPS = sum of nedge over the wall time).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "g5_serve.hh"

using namespace g5host;

namespace {

struct Opts {
    std::vector<int> scales{18, 20, 22, 24, 25, 26};
    int edgefactor = 16, nq = 400, maxroots = 8, lanes = 64, window = 1, small = 4, place = 0;
    double rate = 1000;
    uint64_t seed = 1;
    bool alone = false, compare = false;
    std::string dump;
    double budget_mib = -1;
};

Opts parse(int argc, char **argv)
{
    Opts o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw std::invalid_argument("missing value after " + a);
            return argv[++i];
        };
        if (a == "-g") {
            o.scales.clear();
            std::stringstream ss(next());
            std::string t;
            while (std::getline(ss, t, ',')) o.scales.push_back(std::stoi(t));
        } else if (a == "-e") o.edgefactor = std::stoi(next());
        else if (a == "-rate") o.rate = std::stod(next());
        else if (a == "-nq") o.nq = std::stoi(next());
        else if (a == "-seed") o.seed = std::stoull(next());
        else if (a == "-maxroots") o.maxroots = std::stoi(next());
        else if (a == "-lanes") o.lanes = std::stoi(next());
        else if (a == "-window") o.window = std::stoi(next());
        else if (a == "-small") o.small = std::stoi(next());
        else if (a == "-place") o.place = std::stoi(next());
        else if (a == "-budget") o.budget_mib = std::stod(next());
        else if (a == "-alone") o.alone = true;
        else if (a == "-compare") o.compare = true;
        else if (a == "-dump") o.dump = next();
        else throw std::invalid_argument("unknown option " + a);
    }
    if (o.scales.empty() || o.nq < 1 || o.rate <= 0 || o.maxroots < 1 || o.maxroots > 64 || o.lanes < 1 ||
        o.lanes > 64 || o.maxroots > o.lanes)
        throw std::invalid_argument("bad option values");
    return o;
}

double pct(std::vector<double> v, double p)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const size_t k = std::min(v.size() - 1, (size_t)std::ceil(p / 100.0 * v.size()) - (p > 0 ? 1 : 0));
    return v[k];
}

void check(int rc, const char *what)
{
    if (rc != G5_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

int run(const Opts &o)
{
    check(g5_init(0), "g5_init");
    std::unique_ptr<MemoryBudget> budget;
    if (o.budget_mib >= 0) budget.reset(new MemoryBudget((int64_t)(o.budget_mib * 1048576.0)));
    std::vector<std::unique_ptr<Engine>> eng;
    for (int s : o.scales) eng.emplace_back(new Engine(s, o.edgefactor, o.lanes, o.window, o.small, budget.get(), o.place));

    // the query stream: seeded, independent of the serving run
    std::mt19937_64 rng(o.seed);
    std::exponential_distribution<double> gap(o.rate);
    std::uniform_int_distribution<int> pick_graph(0, (int)eng.size() - 1), pick_n(1, o.maxroots);
    std::vector<Query> qs(o.nq);
    double t = 0;
    for (int i = 0; i < o.nq; ++i) {
        Query &q = qs[i];
        q.id = i;
        t += gap(rng);
        q.arrival = t;
        q.graph = pick_graph(rng);
        const auto &keys = eng[q.graph]->keys();
        std::uniform_int_distribution<size_t> pick_key(0, keys.size() - 1);
        const int n = pick_n(rng);
        for (int k = 0; k < n; ++k) q.roots.push_back(keys[pick_key(rng)]);
    }

    double wall = 0;
    if (o.alone) {
        const double t0 = now_seconds();
        for (auto &e : eng) e->set_epoch(t0);
        for (Query &q : qs) {
            q.arrival = eng[q.graph]->clock();
            eng[q.graph]->run_alone(&q);
        }
        wall = now_seconds() - t0;
    } else {
        const double t0 = now_seconds();
        for (auto &e : eng) {
            e->set_epoch(t0);
            e->start();
        }
        for (Query &q : qs) {
            const double wait = q.arrival - (now_seconds() - t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            eng[q.graph]->submit(&q);
        }
        for (auto &e : eng) e->finish();
        wall = now_seconds() - t0;
    }

    std::vector<double> lat;
    double nedge = 0;
    for (const Query &q : qs) {
        lat.push_back((q.done - q.arrival) * 1e3);
        for (int64_t x : q.nedge) nedge += (double)x;
    }
    int mismatches = -1;
    if (o.compare && !o.alone) {
        mismatches = 0;
        for (const Query &q : qs) {
            Query a;
            a.roots = q.roots;
            eng[q.graph]->run_alone(&a);
            if (a.nedge != q.nedge || a.checksum != q.checksum) ++mismatches;
        }
    }
    if (!o.dump.empty()) {
        FILE *f = fopen(o.dump.c_str(), "w");
        if (!f) throw std::runtime_error("cannot open " + o.dump);
        for (const Query &q : qs) {
            fprintf(f, "%d %d %.6f %.6f %zu", q.id, eng[q.graph]->scale(), q.arrival, q.done, q.roots.size());
            for (size_t k = 0; k < q.roots.size(); ++k)
                fprintf(f, " %lld %lld %llu", (long long)q.roots[k], (long long)q.nedge[k],
                        (unsigned long long)q.checksum[k]);
            fprintf(f, "\n");
        }
        fclose(f);
    }
    int64_t levels = 0, split = 0, single = 0, waits = 0;
    double dev_ms = 0;
    std::string per;
    for (auto &e : eng) {
        const EngineStats st = e->stats();
        levels += st.levels;
        split += st.split_levels;
        single += st.single_roots;
        waits += st.budget_waits;
        dev_ms += st.device_ms;
        per += (per.empty() ? "" : ", ") + std::string("\"") + std::to_string(e->scale()) + "\": {\"levels\": " +
               std::to_string(st.levels) + ", \"split_levels\": " + std::to_string(st.split_levels) +
               ", \"roots\": " + std::to_string(st.joins) + ", \"single_roots\": " +
               std::to_string(st.single_roots) + ", \"budget_waits\": " + std::to_string(st.budget_waits) +
               ", \"lane_bytes\": " + std::to_string(e->lane_bytes()) + ", \"device_bytes\": " +
               std::to_string(e->device_bytes()) + ", \"device_ms\": " + std::to_string(st.device_ms) + "}";
    }
    std::string scales;
    for (size_t i = 0; i < o.scales.size(); ++i) scales += (i ? "," : "") + std::to_string(o.scales[i]);
    printf("{\"mode\": \"%s\", \"graphs\": [%s], \"queries\": %d, \"rate_qps\": %.1f, \"lanes\": %d, \"window\": %d, "
           "\"small\": %d, \"place\": %d, "
           "\"wall_s\": %.4f, \"achieved_qps\": %.1f, \"gteps\": %.4f, \"latency_ms\": {\"p50\": %.3f, "
           "\"p90\": %.3f, \"p99\": %.3f, \"max\": %.3f}, \"levels\": %lld, \"split_levels\": %lld, "
           "\"single_roots\": %lld, \"device_ms\": %.3f, \"budget_bytes\": %lld, \"budget_peak_bytes\": %lld, "
           "\"budget_waits\": %lld, "
           "\"mismatches\": %d, \"engines\": {%s}}\n",
           o.alone ? "alone" : "served", scales.c_str(), o.nq, o.rate, o.lanes, o.window, o.small, o.place, wall,
           o.nq / wall, nedge / wall / 1e9, pct(lat, 50), pct(lat, 90), pct(lat, 99), pct(lat, 100),
           (long long)levels, (long long)split, (long long)single, dev_ms,
           (long long)(budget ? budget->limit() : -1), (long long)(budget ? budget->peak() : -1), (long long)waits,
           mismatches, per.c_str());
    return mismatches > 0 ? 3 : 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(parse(argc, argv));
    } catch (const std::exception &e) {
        fprintf(stderr, "g500_serve: %s\n", e.what());
        return 1;
    }
}
