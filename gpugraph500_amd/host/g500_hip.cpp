// g500_hip -- the benchmark driver of the 2D BFS (cpu_2d/main.cpp) for one
// MI355X: the Kronecker graph built on the device and split into R x C
// blocks, search keys chosen as main.cpp:351-411 does, and every search run
// through the emulated R x C GlobalBFS level loop (g5_grid.hh).
//
//   g500_hip [-s scale] [-e edgefactor] [-R rows] [-C cols] [-n roots]
//            [-x host|device] [-be codec] [-btr threshold] [-share r c]
//            [-td] [-verify] [-validate] [-dump file] [-dumpk k]
//
//   -x host     the reference's hook protocol (host queues, every message
//               through the Compression schema); device: bitmaps in HBM
//   -be codec   "" (raw) or s4-bp128-d4 (GPU codec); in device mode the codec
//               runs on the fold's queues
//   -share r c  the rank whose share (expansion, its column's reduce, its
//               row's fold) is timed separately (default 0 0)
//   -td         top-down levels only (device exchange otherwise runs dense
//               levels bottom-up over the transposed blocks)
//   -verify     compare every search with the 1 x 1 search (g5_bfs_run on the
//               undivided graph): predecessors must be identical
//   -validate   check every search's predecessors on the device as the
//               reference's validate_bfs_result does (main.cpp:511-533,
//               validate/validate.h:249-491); its nedge must equal the TEPS
//               numerator used
//   -dump f     write the first k searches' predecessors: int64 count, then
//               per search int64 root, int64 n, int64 pred[n]
// Output: per-search lines, then the reference's statistics names
// (main.cpp:915-963) and one JSON line.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "g5_compression.hh"
#include "g5_grid.hh"

using namespace g5host;

namespace {

void check(int rc, const char *what)
{
    if (rc != G5_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

struct Opts {
    int scale = 16, edgefactor = 16, R = 1, C = 1, roots = 64, r0 = 0, c0 = 0, dumpk = 4, threshold = 64;
    std::string exchange = "device", codec, dump;
    bool verify = false, validate = false, td = false;
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
        if (a == "-s") o.scale = std::stoi(next());
        else if (a == "-e") o.edgefactor = std::stoi(next());
        else if (a == "-R") o.R = std::stoi(next());
        else if (a == "-C") o.C = std::stoi(next());
        else if (a == "-n") o.roots = std::stoi(next());
        else if (a == "-x") o.exchange = next();
        else if (a == "-be") o.codec = next();
        else if (a == "-btr") o.threshold = std::stoi(next());
        else if (a == "-share") {
            o.r0 = std::stoi(next());
            o.c0 = std::stoi(next());
        } else if (a == "-verify") o.verify = true;
        else if (a == "-validate") o.validate = true;
        else if (a == "-td") o.td = true;
        else if (a == "-dump") o.dump = next();
        else if (a == "-dumpk") o.dumpk = std::stoi(next());
        else throw std::invalid_argument("unknown option " + a);
    }
    if (o.exchange != "host" && o.exchange != "device") throw std::invalid_argument("-x must be host or device");
    if (o.roots < 1 || o.roots > 64) throw std::invalid_argument("-n must be in [1, 64]");
    if (o.r0 < 0 || o.r0 >= o.R || o.c0 < 0 || o.c0 >= o.C) throw std::invalid_argument("-share outside the grid");
    return o;
}

double hmean(const std::vector<double> &x)
{
    double s = 0;
    for (double v : x) s += 1.0 / v;
    return x.size() / s;
}

int run(const Opts &o)
{
    check(g5_init(0), "g5_init");
    const uint64_t s1 = 1, s2 = 2;  // main.cpp:221
    std::vector<g5_graph *> blocks((size_t)o.R * o.C, nullptr);
    check(g5_graph_kronecker_2d(o.scale, o.edgefactor, s1, s2, o.R, o.C, blocks.data()), "g5_graph_kronecker_2d");
    // the undivided graph: degrees for the search keys, and the 1 x 1 search for -verify
    g5_graph *whole = nullptr;
    check(g5_graph_kronecker(o.scale, o.edgefactor, s1, s2, &whole), "g5_graph_kronecker");
    std::vector<int64_t> roots(o.roots);
    int found = 0;
    check(g5_select_roots(whole, (int64_t)1 << o.scale, o.roots, roots.data(), &found), "g5_select_roots");
    roots.resize(found);
    g5_bfs *single = nullptr;
    if (o.verify) check(g5_bfs_create(whole, &single), "g5_bfs_create");
    g5_validator *validator = nullptr;
    int64_t whole_rows = 0, whole_nglobal = 0;
    {
        int64_t nc = 0, nnz = 0, gen = 0;
        check(g5_graph_info(whole, &whole_rows, &nc, &nnz, &whole_nglobal, &gen), "g5_graph_info");
    }
    if (o.validate) check(g5_validator_create(whole, &validator), "g5_validator_create");

    Compression<int64_t, uint32_t> *schema = nullptr;
    bool device_codec = false;
    if (!o.codec.empty()) {
        if (o.exchange == "host") {
            schema = CompressionFactory<int64_t, uint32_t>::getFromName("hipsimd");
            schema->reconfigure(o.threshold, o.codec);
        } else {
            if (o.codec != "s4-bp128-d4") throw std::invalid_argument("only s4-bp128-d4 runs on the GPU");
            device_codec = true;
        }
    }
    std::unique_ptr<Grid2D> gridp(new Grid2D(o.R, o.C, blocks.data(),
                                             o.exchange == "host" ? Exchange::HostQueues : Exchange::Device, schema,
                                             device_codec, !o.td));
    Grid2D &grid = *gridp;

    FILE *dump = nullptr;
    if (!o.dump.empty()) {
        dump = fopen(o.dump.c_str(), "wb");
        if (!dump) throw std::runtime_error("cannot open " + o.dump);
        const int64_t k = std::min<int64_t>(o.dumpk, (int64_t)roots.size());
        fwrite(&k, 8, 1, dump);
    }
    std::vector<double> teps, share_expand, share_column, share_fold, totals;
    std::vector<int64_t> pred, ref, vpred;
    std::vector<double> valid_time;
    int mismatches = 0, invalid = 0, bu_levels = 0;
    double codec_s = 0, pred_s = 0;
    int64_t raw_words = 0, coded_words = 0;
    // one untimed search first (device buffers, code objects), then the
    // watched rank's expansion kernel is accounted (g5_bfs_profile)
    GridStats st;
    grid.runBFS(roots[0], pred, st, o.r0, o.c0);
    check(g5_bfs_profile(grid.rank(o.r0, o.c0).handle(), 1), "g5_bfs_profile");
    for (size_t i = 0; i < roots.size(); ++i) {
        grid.runBFS(roots[i], pred, st, o.r0, o.c0);
        const int64_t ne = grid.nedge(pred);
        teps.push_back(ne / st.total_s);
        totals.push_back(st.total_s);
        share_expand.push_back(st.share_expand_s);
        share_column.push_back(st.share_column_s);
        share_fold.push_back(st.share_fold_s);
        codec_s += st.codec_s;
        pred_s += st.pred_s;
        raw_words += st.raw_words;
        coded_words += st.coded_words;
        if (o.verify) {
            int32_t nl = 0;
            float ms = 0;
            check(g5_bfs_run(single, roots[i], 0, &nl, &ms), "g5_bfs_run");
            int64_t n1 = 0, nc = 0, nnz = 0, ng = 0, gen = 0;
            check(g5_graph_info(whole, &n1, &nc, &nnz, &ng, &gen), "g5_graph_info");
            ref.assign(n1, -1);
            check(g5_bfs_result(single, ref.data(), nullptr), "g5_bfs_result");
            const int64_t n = std::min<int64_t>(n1, (int64_t)pred.size());
            bool same = true;
            for (int64_t v = 0; v < n && same; ++v) same = pred[v] == ref[v];
            for (int64_t v = n; v < (int64_t)pred.size() && same; ++v) same = pred[v] == -1;
            if (!same) ++mismatches;
        }
        if (o.validate) {
            // the grid's map may be padded past the graph's rows: those entries must be empty
            vpred.assign(pred.begin(), pred.begin() + std::min<int64_t>(whole_rows, (int64_t)pred.size()));
            vpred.resize(whole_rows, -1);
            bool pad_empty = true;
            for (size_t v = vpred.size(); v < pred.size(); ++v) pad_empty = pad_empty && pred[v] == -1;
            int bad = 0;
            int64_t vne = 0;
            float vms = 0;
            check(g5_validate(validator, roots[i], whole_nglobal, vpred.data(), &bad, &vne, nullptr, &vms),
                  "g5_validate");
            const bool ok = bad == 0 && pad_empty && vne == ne;
            if (!ok) ++invalid;
            valid_time.push_back(vms * 1e-3);
            printf("Validation of iteration %zu finished in %fs\n", i, vms * 1e-3);
            printf("Result: %s %lld Edge(s) processed (flags %d)\n", ok ? "Valid" : "Invalid", (long long)vne, bad);
        }
        if (dump && (int64_t)i < o.dumpk) {
            const int64_t root = roots[i], n = (int64_t)pred.size();
            fwrite(&root, 8, 1, dump);
            fwrite(&n, 8, 1, dump);
            fwrite(pred.data(), 8, (size_t)n, dump);
        }
        printf("search %zu root %lld levels %d (bottom-up %d) nedge %lld time %.6f s TEPS %.4e share expand %.6f "
               "column %.6f fold %.6f s\n",
               i, (long long)roots[i], st.levels, st.bottom_up_levels, (long long)ne, st.total_s, ne / st.total_s,
               st.share_expand_s, st.share_column_s, st.share_fold_s);
        bu_levels += st.bottom_up_levels;
    }
    if (dump) fclose(dump);
    const double hm = hmean(teps);
    double mean_t = 0, mean_se = 0, mean_sc = 0, mean_sf = 0;
    for (size_t i = 0; i < teps.size(); ++i) {
        mean_t += totals[i];
        mean_se += share_expand[i];
        mean_sc += share_column[i];
        mean_sf += share_fold[i];
    }
    const double k = (double)teps.size();
    printf("harmonic_mean_TEPS: %.3e\n", hm);
    printf("mean_time: %.6f\n", mean_t / k);
    printf("mean_share_expand_time: %.6f\n", mean_se / k);
    printf("mean_share_column_time: %.6f\n", mean_sc / k);
    printf("mean_share_fold_time: %.6f\n", mean_sf / k);
    if (o.verify) printf("verify_mismatches: %d\n", mismatches);
    if (o.validate) {
        double vt = 0;
        for (double t : valid_time) vt += t;
        printf("Validation: %s\n", invalid ? "failed!" : "passed");
        printf("mean_validation_time: %.6f\n", valid_time.empty() ? 0.0 : vt / valid_time.size());
    }
    double kx_ms = 0, kx_bytes = 0;
    int64_t kx_launches = 0, kx_edges = 0;
    check(g5_bfs_stats(grid.rank(o.r0, o.c0).handle(), &kx_ms, &kx_launches, &kx_bytes, &kx_edges), "g5_bfs_stats");
    printf("{\"scale\": %d, \"R\": %d, \"C\": %d, \"roots\": %zu, \"exchange\": \"%s\", \"codec\": \"%s\", "
           "\"hmean_teps\": %.6e, \"mean_time_s\": %.6f, \"share\": [%d, %d], \"share_expand_s\": %.6f, "
           "\"share_column_s\": %.6f, \"share_fold_s\": %.6f, \"codec_s\": %.6f, \"pred_s\": %.6f, "
           "\"raw_words\": %lld, \"coded_words\": %lld, \"verify_mismatches\": %d, \"share_kexpand_ms\": %.6f, "
           "\"share_kexpand_launches\": %lld, \"share_kexpand_bytes\": %.6e, \"share_kexpand_edges\": %lld, "
           "\"validate_failures\": %d, \"bottom_up_levels\": %d}\n",
           o.scale, o.R, o.C, teps.size(), o.exchange.c_str(), o.codec.c_str(), hm, mean_t / k, o.r0, o.c0,
           mean_se / k, mean_sc / k, mean_sf / k, codec_s / k, pred_s / k, (long long)raw_words,
           (long long)coded_words, o.verify ? mismatches : -1, kx_ms, (long long)kx_launches, kx_bytes,
           (long long)kx_edges, o.validate ? invalid : -1, bu_levels);
    gridp.reset();
    if (validator) g5_validator_free(validator);
    if (single) g5_bfs_free(single);
    g5_graph_free(whole);
    for (g5_graph *g : blocks) g5_graph_free(g);
    return ((o.verify && mismatches) || (o.validate && invalid)) ? 3 : 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(parse(argc, argv));
    } catch (const std::exception &e) {
        fprintf(stderr, "g500_hip: %s\n", e.what());
        return 1;
    }
}
