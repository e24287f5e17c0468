// flip_graph: the reference's flip-graph search CLI (src/main_flip_graph.cu)
// on the MI355X engine.  Same flags, defaults, input format, report table
// and JSON output naming; the walk runs through the fg.h C ABI.
//
// Extensions (not in the reference): --ring picks Z2 or ZT at run time
// (the reference fixes it at compile time, src/config.cuh:6);
// --max-launches and --target-rank give the otherwise endless loop
// (src/entities/flip_graph.cu:286) a stop condition; --max-rank sets the
// term capacity that resized schemes may grow to.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "cli_common.h"
#include "fg.h"
#include "scheme_host.h"

using fgcli::ArgType;

namespace {

struct Run {
    fg_ctx *ctx = nullptr;
    int ring = 0;
    int walkers = 0;
    std::string path;
    std::map<std::string, int> best_by_key;  // n2bestRank
    std::string main_key;                    // format of -n1 -n2 -n3
};

// FlipGraph::updateRanks(iteration, true) (src/entities/flip_graph.cu:249-277)
// after a resize pass: a walker whose rank is below its format's best so far
// makes its current scheme its best; every walker's target rank becomes its
// format's best; each format that improved saves the scheme that improved it.
int update_ranks(Run &r, int iteration) {
    const int W = r.walkers;
    std::vector<int32_t> m(W), n((size_t)3 * W), ranks(W);
    if (fg_summary(r.ctx, m.data(), nullptr, nullptr, nullptr, n.data()) < 0) return -1;
    std::vector<std::string> keys(W);
    std::map<std::string, int> best_index;  // n2bestIndex
    std::vector<int32_t> snap;
    for (int w = 0; w < W; ++w) {
        keys[w] = fg::format_key(n[3 * w], n[3 * w + 1], n[3 * w + 2]);
        auto it = r.best_by_key.find(keys[w]);
        if (it == r.best_by_key.end() || m[w] < it->second) {
            r.best_by_key[keys[w]] = m[w];
            best_index[keys[w]] = w;
            snap.push_back(w);
        }
    }
    if (fg_snapshot_best(r.ctx, snap.data(), (int)snap.size()) < 0) return -1;
    for (int w = 0; w < W; ++w) ranks[w] = r.best_by_key[keys[w]];
    if (fg_set_best_ranks(r.ctx, ranks.data()) < 0) return -1;
    for (const auto &kv : best_index) {
        fg::Scheme s;
        if (!fgcli::fetch_scheme(r.ctx, r.ring, kv.second, 0, &s) || !fgcli::verified(r.ring, s)) {
            std::cout << "Scheme of walker " << kv.second << " failed verification, not saved" << std::endl;
            continue;
        }
        const std::string p = fgcli::save_path(r.path, r.ring, s, iteration, kv.second);
        fgcli::make_dirs(r.path);
        fg::save_json(s, p);
        std::cout << "Best rank of " << kv.first << " was improved to " << r.best_by_key[kv.first]
                  << " (known: " << fgcli::known_rank(kv.first, r.ring) << ")! Scheme saved to \"" << p << "\""
                  << std::endl;
    }
    if (!best_index.empty()) std::cout << std::endl;
    return 0;
}

// FlipGraph::report (src/entities/flip_graph.cu:303-380).  Returns the best
// rank seen for the format of -n1 -n2 -n3 (-1 when no walker has it).
int report(Run &r, std::chrono::steady_clock::time_point start, int iteration, const std::vector<double> &times,
           int log_period, int count = 3) {
    const int W = r.walkers;
    std::vector<int32_t> m(W), bm(W), br(W), fc(W), n((size_t)3 * W);
    if (fg_summary(r.ctx, m.data(), bm.data(), br.data(), fc.data(), n.data()) < 0) {
        std::cerr << "fg_summary: " << fg_last_error() << std::endl;
        return -1;
    }
    std::map<std::string, std::vector<int>> by_key;
    for (int w = 0; w < W; ++w) by_key[fg::format_key(n[3 * w], n[3 * w + 1], n[3 * w + 2])].push_back(w);
    for (auto &kv : by_key) {
        std::vector<int> &idx = kv.second;
        const int k = std::min<int>(count, (int)idx.size());
        std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](int a, int b) { return bm[a] < bm[b]; });
        const int w = idx[0];
        auto it = r.best_by_key.find(kv.first);
        const int prev = it == r.best_by_key.end() ? 1 << 30 : it->second;
        if (br[w] < prev) {
            fg::Scheme s;
            if (fgcli::fetch_scheme(r.ctx, r.ring, w, 1, &s) && fgcli::verified(r.ring, s)) {
                const std::string p = fgcli::save_path(r.path, r.ring, s, iteration, w);
                fgcli::make_dirs(r.path);
                fg::save_json(s, p);
                std::cout << "Best rank of " << kv.first << " was improved from " << prev << " to " << br[w]
                          << " (known: " << fgcli::known_rank(kv.first, r.ring) << ")! Scheme saved to \"" << p
                          << "\"" << std::endl;
                r.best_by_key[kv.first] = br[w];
            } else {
                std::cout << "Best scheme of walker " << w << " failed verification, not saved" << std::endl;
            }
        }
    }
    if (log_period == 0 || iteration % log_period == 0) {
        const double last = times.back();
        const double mn = *std::min_element(times.begin(), times.end());
        const double mx = *std::max_element(times.begin(), times.end());
        const double mean = std::accumulate(times.begin(), times.end(), 0.0) / times.size();
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::cout << std::endl << std::left;
        std::cout << "+--------------------------------------------------------------+" << std::endl;
        std::cout << "| Schemes            Iteration            Elapsed time         |" << std::endl;
        std::cout << "| " << std::right << std::setw(7) << W << "            " << std::setw(9) << iteration
                  << "            " << std::setw(12) << fgcli::pretty_time(elapsed) << "         |" << std::endl;
        std::cout << "+--------+--------+--------+-------+------+------+-------------+" << std::endl;
        std::cout << "| run id |  size  |  real  | known | best | curr | flips count |" << std::endl;
        std::cout << "+--------+--------+--------+-------+------+------+-------------+" << std::endl;
        for (auto &kv : by_key) {
            const std::vector<int> &idx = kv.second;
            const int known = fgcli::known_rank(kv.first, r.ring);
            for (int q = 0; q < count && q < (int)idx.size(); ++q) {
                const int w = idx[q];
                std::cout << "| " << std::setw(6) << (w + 1) << " | " << std::setw(6) << kv.first << " | "
                          << std::setw(6) << fg::format_key(n[3 * w], n[3 * w + 1], n[3 * w + 2], false) << " | "
                          << std::setw(5) << known << " | " << std::setw(4) << br[w] << " | " << std::setw(4) << m[w]
                          << " | " << std::setw(11) << fgcli::pretty_int(fc[w]) << " |";
                if (q == 0) {
                    std::cout << " total: " << idx.size();
                    if (br[idx[0]] <= known) std::cout << ", " << (br[idx[0]] == known ? "equal" : "BETTER!!!");
                }
                std::cout << std::endl;
            }
            std::cout << "+--------+--------+--------+-------+------+------+-------------+" << std::endl;
        }
        std::cout << "- iteration time (last / min / max / mean): " << fgcli::pretty_time(last) << " / "
                  << fgcli::pretty_time(mn) << " / " << fgcli::pretty_time(mx) << " / " << fgcli::pretty_time(mean)
                  << std::endl
                  << std::endl;
    }
    auto mk = r.best_by_key.find(r.main_key);
    return mk == r.best_by_key.end() ? -1 : mk->second;
}

}  // namespace

int main(int argc, char *argv[]) {
    fgcli::ArgParser parser("flip_graph", "Find fast matrix multiplication schemes with flip-graph random walks on an AMD MI355X");
    parser.add("-n1", ArgType::Natural, "INT", "number of first matrix rows");
    parser.add("-n2", ArgType::Natural, "INT", "number of first matrix columns and second matrix rows");
    parser.add("-n3", ArgType::Natural, "INT", "number of second matrix columns");
    parser.add("--schemes", ArgType::Natural, "INT", "number of schemes", "1024");
    parser.add("--max-iterations", ArgType::Natural, "INT", "number of flips per iterations", "10000");
    parser.add("--path", ArgType::String, "PATH", "path to save schemes", "schemes");
    parser.add("--input-path", ArgType::String, "PATH", "path to init schemes", "null");
    parser.add("--block-size", ArgType::Natural, "INT", "walkers per workgroup (accepted; the runtime picks 32 or 64 with the LDS layout)", "32");
    parser.add("--expand-probability", ArgType::Real, "REAL", "expand edge probability (divided by max iterations)", "0.01");
    parser.add("--reduce-probability", ArgType::Real, "REAL", "reduce edge probability (divided by max iterations)", "0.0");
    parser.add("--sandwiching-probability", ArgType::Real, "REAL", "sandwiching edge probability (divided by max iterations)", "0.0");
    parser.add("--basis-probability", ArgType::Real, "REAL", "basis edge probability (divided by max iterations)", "0.0");
    parser.add("--resize-probability", ArgType::Real, "REAL", "project/extend edge probability", "0.2");
    parser.add("--plus-iterations", ArgType::Natural, "INT", "number of iterations of plus call", "8000");
    parser.add("--log-period", ArgType::Natural, "INT", "period of report printing", "0");
    parser.add("--seed", ArgType::Natural, "INT", "random seed", "0");
    parser.add("--ring", ArgType::String, "RING", "coefficient ring, Z2 or ZT", "Z2");
    parser.add("--max-launches", ArgType::Natural, "INT", "stop after this many walk launches (0: run forever)", "0");
    parser.add("--target-rank", ArgType::Natural, "INT", "stop once a verified scheme of at most this rank is saved (0: never)", "0");
    parser.add("--max-rank", ArgType::Natural, "INT", "terms a resized scheme may grow to (0: min(255, max(64, 2 n1 n2 n3)))", "0");

    if (!parser.parse(argc, argv)) return 0;

    const int n1 = std::stoi(parser.get("-n1")), n2 = std::stoi(parser.get("-n2")), n3 = std::stoi(parser.get("-n3"));
    const int walkers = std::stoi(parser.get("--schemes"));
    const int max_iterations = std::stoi(parser.get("--max-iterations"));
    const std::string path = parser.get("--path");
    const std::string input_path = parser.get("--input-path");
    const int block_size = std::stoi(parser.get("--block-size"));
    const int plus_iterations = std::stoi(parser.get("--plus-iterations"));
    const int log_period = std::stoi(parser.get("--log-period"));
    long long seed = std::stoll(parser.get("--seed"));
    const double p_expand = std::stod(parser.get("--expand-probability"));
    const double p_reduce = std::stod(parser.get("--reduce-probability"));
    const double p_sandwich = std::stod(parser.get("--sandwiching-probability"));
    const double p_basis = std::stod(parser.get("--basis-probability"));
    const double p_resize = std::stod(parser.get("--resize-probability"));
    const std::string ring_s = parser.get("--ring");
    const int max_launches = std::stoi(parser.get("--max-launches"));
    const int target_rank = std::stoi(parser.get("--target-rank"));
    int max_rank = std::stoi(parser.get("--max-rank"));
    if (ring_s != "Z2" && ring_s != "ZT") {
        std::cout << "Unknown ring \"" << ring_s << "\" (expected Z2 or ZT)" << std::endl;
        return 0;
    }
    const int ring = ring_s == "ZT" ? FG_RING_ZT : FG_RING_Z2;
    if (seed == 0) seed = (long long)time(nullptr);  // main_flip_graph.cu:49-50

    if (input_path == "null" && (n1 * n2 > 32 || n2 * n3 > 32 || n3 * n1 > 32)) {
        std::cout << "Error sizes: factors wider than 32 coefficients are not supported (now: 32)" << std::endl;
        return 0;
    }
    if (input_path == "null" && n1 * n2 * n3 > 255) {
        std::cout << "Error sizes: schemes of more than 255 terms are not supported (now: 255)" << std::endl;
        return 0;
    }

    std::cout << "Config:" << std::endl;
    std::cout << "- working ring: " << fgcli::ring_name(ring) << std::endl;
    std::cout << "- max matrix elements: 32" << std::endl;
    std::cout << "- max scheme rank: 255" << std::endl;
    std::cout << "- walkers per workgroup: 64" << std::endl;
    std::cout << std::endl;
    std::cout << "Start flip graph algorithm" << std::endl;
    std::cout << "- n: " << n1 << " " << n2 << " " << n3 << std::endl;
    std::cout << "- schemes count: " << walkers << std::endl;
    std::cout << "- max iterations: " << max_iterations << std::endl;
    std::cout << "- plus iterations: " << plus_iterations << std::endl;
    std::cout << "- path: " << path << std::endl;
    std::cout << "- block size: " << block_size << std::endl;
    std::cout << "- probabilities:" << std::endl;
    std::cout << "  - expand: " << p_expand << std::endl;
    std::cout << "  - reduce: " << p_reduce << std::endl;
    std::cout << "  - sandwiching: " << p_sandwich << std::endl;
    std::cout << "  - basis: " << p_basis << std::endl;
    std::cout << "  - resize: " << p_resize << std::endl;
    std::cout << "- log period: " << log_period << std::endl;
    std::cout << "- seed: " << seed << std::endl;
    const bool resize = p_resize > 0;
    if (resize) {
        if (max_rank == 0) max_rank = std::min(255, std::max(64, 2 * n1 * n2 * n3));
        if (max_rank > 255) max_rank = 255;
        std::cout << "- resized schemes grow to at most " << max_rank << " terms" << std::endl;
    }

    fg_config cfg{};
    cfg.ring = ring;
    cfg.walkers = walkers;
    cfg.max_iterations = max_iterations;
    cfg.plus_iterations = plus_iterations;
    cfg.random_iterations = p_resize > 0;
    cfg.check_reduce = 1;
    cfg.expand_probability = p_expand;
    cfg.reduce_probability = p_reduce;
    cfg.seed = (uint64_t)seed;
    if (resize) {
        cfg.max_rank = max_rank;
        if (ring == FG_RING_Z2) cfg.word_bits = 32;  // resized Z2 factors may exceed 16 coefficients
    }
    cfg.sandwiching_probability = p_sandwich;
    cfg.basis_probability = p_basis;
    if ((p_sandwich > 0 || p_basis > 0) && ring == FG_RING_Z2) cfg.word_bits = 32;
    Run r;
    r.ring = ring;
    r.walkers = walkers;
    r.path = path;
    r.main_key = fg::format_key(n1, n2, n3);
    if (fg_create(&cfg, &r.ctx) < 0) {
        std::cout << "Error: " << fg_last_error() << std::endl;
        return -1;
    }
    int rc;
    if (input_path != "null") {
        std::vector<int32_t> ints;
        std::string err;
        if (!fg::read_int_file(input_path, &ints, &err)) {
            std::cout << err << std::endl;
            fg_destroy(r.ctx);
            return -1;
        }
        std::cout << "Start reading " << std::min<long long>(ints.empty() ? 0 : ints[0], walkers) << " schemes" << std::endl;
        rc = fg_init_from_schemes(r.ctx, ints.data(), (int64_t)ints.size());
    } else {
        std::cout << "Start initializing with naive schemes" << std::endl;
        rc = fg_init_naive(r.ctx, n1, n2, n3);
    }
    if (rc < 0) {
        std::cout << "Error: " << fg_last_error() << std::endl;
        fg_destroy(r.ctx);
        return -1;
    }
    std::cout << "Start main initialization" << std::endl;
    {
        // updateRanks(0, false): the smallest starting rank per format
        std::vector<int32_t> m(walkers), n((size_t)3 * walkers);
        fg_summary(r.ctx, m.data(), nullptr, nullptr, nullptr, n.data());
        r.best_by_key[fg::format_key(n1, n2, n3)] = n1 * n2 * n3;
        for (int w = 0; w < walkers; ++w) {
            const std::string key = fg::format_key(n[3 * w], n[3 * w + 1], n[3 * w + 2]);
            auto it = r.best_by_key.find(key);
            if (it == r.best_by_key.end() || m[w] < it->second) r.best_by_key[key] = m[w];
        }
    }
    std::cout << "Initialized" << std::endl;

    const auto start = std::chrono::steady_clock::now();
    std::vector<double> times;
    auto t1 = std::chrono::steady_clock::now();
    for (int iteration = 0; max_launches == 0 || iteration < max_launches; ++iteration) {
        const int bad = fg_walk(r.ctx, 0);
        if (bad < 0) {
            std::cout << "Error: " << fg_last_error() << std::endl;
            fg_destroy(r.ctx);
            return -1;
        }
        if (bad > 0) std::cout << "invalid (" << bad << ") schemes (random walk)" << std::endl;
        const auto t2 = std::chrono::steady_clock::now();
        times.push_back(std::chrono::duration<double>(t2 - t1).count());
        const int best = report(r, start, iteration + 1, times, log_period);
        t1 = std::chrono::steady_clock::now();
        if (resize) {  // flip_graph.cu:296-299
            const int rbad = fg_resize(r.ctx, p_resize);
            if (rbad < 0 || update_ranks(r, iteration) < 0) {
                std::cout << "Error: " << fg_last_error() << std::endl;
                fg_destroy(r.ctx);
                return -1;
            }
            if (rbad > 0) std::cout << "invalid (" << rbad << ") schemes (resize)" << std::endl;
        }
        if (target_rank > 0 && best > 0 && best <= target_rank) {
            std::cout << "Target rank " << target_rank << " reached" << std::endl;
            break;
        }
    }
    fg_destroy(r.ctx);
    std::cout << "Success!" << std::endl;
    return 0;
}
