// complexity_minimizer: the reference's complexity minimizer CLI
// (src/main_complexity_minimizer.cu, src/entities/complexity_minimizer.cu)
// on the MI355X engine.  Same flags, defaults, input format, report table,
// re-broadcast and JSON output naming; the minimizer runs through the fg.h
// C ABI.
//
// Extension (not in the reference): --ring picks Z2 or ZT at run time (the
// reference fixes it at compile time, src/config.cuh:6).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <string>
#include <vector>

#include "cli_common.h"
#include "fg.h"
#include "scheme_host.h"

using fgcli::ArgType;

namespace {

constexpr int kTopCount = 10;  // ComplexityMinimizer(..., topCount = 10)

struct Minimizer {
    fg_ctx *ctx = nullptr;
    int ring = 0;
    int walkers = 0;
    std::string path;
    int best_complexity = INT_MAX;
    std::vector<int> indices;
    std::vector<int32_t> best, cur;

    int top() const { return std::min(kTopCount, walkers); }

    bool refresh() {
        if (fg_complexities(ctx, best.data(), cur.data()) < 0) {
            std::cout << "Error: " << fg_last_error() << std::endl;
            return false;
        }
        return true;
    }

    // updateBest (complexity_minimizer.cu:86-101)
    bool update_best(int iteration) {
        std::partial_sort(indices.begin(), indices.begin() + top(), indices.end(),
                          [this](int a, int b) { return best[a] < best[b]; });
        const int w = indices[0];
        if (best[w] >= best_complexity) return false;
        fg::Scheme s;
        if (!fgcli::fetch_scheme(ctx, ring, w, 1, &s) || !fgcli::verified(ring, s)) {
            std::cout << "Best scheme of walker " << w << " failed verification, not saved" << std::endl;
            return false;
        }
        const std::string p = fgcli::save_path(path, ring, s, iteration, w);
        fgcli::make_dirs(path);
        fg::save_json(s, p);
        std::cout << "Best complexity improved from " << best_complexity << " to " << best[w]
                  << "! Scheme saved to \"" << p << "\"" << std::endl;
        best_complexity = best[w];
        return true;
    }

    // report (complexity_minimizer.cu:103-134), including the re-broadcast
    // of the best scheme that ends it
    bool report(std::chrono::steady_clock::time_point start, int iteration, const std::vector<double> &times) {
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        const double last = times.back();
        const double mn = *std::min_element(times.begin(), times.end());
        const double mx = *std::max_element(times.begin(), times.end());
        const double mean = std::accumulate(times.begin(), times.end(), 0.0) / times.size();
        std::cout << "+-----------+-----------+--------+--------+--------+" << std::endl;
        std::cout << "|  elapsed  | iteration | run id |  best  |  curr  |" << std::endl;
        std::cout << "+-----------+-----------+--------+--------+--------+" << std::endl;
        for (int i = 0; i < top(); ++i) {
            const int w = indices[i];
            std::cout << "| " << std::setw(9) << fgcli::pretty_time(elapsed) << " | " << std::setw(9) << iteration
                      << " | " << std::setw(6) << (w + 1) << " | " << std::setw(6) << best[w] << " | "
                      << std::setw(6) << cur[w] << " |" << std::endl;
        }
        std::cout << "+-----------+-----------+--------+--------+--------+" << std::endl;
        std::cout << "- iteration time (last / min / max / mean): " << fgcli::pretty_time(last) << " / "
                  << fgcli::pretty_time(mn) << " / " << fgcli::pretty_time(mx) << " / " << fgcli::pretty_time(mean)
                  << std::endl
                  << std::endl;

        const int period = 1 + rand() % 10;
        std::vector<int32_t> dst;
        for (size_t i = 0; i < indices.size(); ++i)
            if (i % (size_t)(iteration % period + 1) == 0) dst.push_back(indices[i]);
        if (fg_broadcast_best(ctx, indices[0], dst.data(), (int)dst.size()) < 0) {
            std::cout << "Error: " << fg_last_error() << std::endl;
            return false;
        }
        return true;
    }
};

// "n1 n2 n3 m count" followed by count schemes without headers
// (ComplexityMinimizer::read, complexity_minimizer.cu:22-37) -> fg.h blob.
bool read_input(const std::vector<int32_t> &ints, std::vector<int32_t> *blob, int *count, int *n, int *m,
                std::string *err) {
    if (ints.size() < 5) {
        *err = "input has no \"n1 n2 n3 m count\" header";
        return false;
    }
    n[0] = ints[0];
    n[1] = ints[1];
    n[2] = ints[2];
    *m = ints[3];
    *count = ints[4];
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || *m < 1 || *count < 1) {
        *err = "bad header";
        return false;
    }
    const long long per = (long long)*m * (n[0] * n[1] + n[1] * n[2] + n[2] * n[0]);
    if ((long long)ints.size() < 5 + per * *count) {
        *err = "input ends before scheme " + std::to_string(((long long)ints.size() - 5) / std::max(per, 1LL));
        return false;
    }
    blob->assign(1, *count);
    size_t pos = 5;
    for (int q = 0; q < *count; ++q) {
        blob->insert(blob->end(), {n[0], n[1], n[2], *m});
        blob->insert(blob->end(), ints.begin() + pos, ints.begin() + pos + per);
        pos += per;
    }
    return true;
}

}  // namespace

int main(int argc, char *argv[]) {
    fgcli::ArgParser parser("complexity_minimizer",
                            "Find best complexity of the fast matrix multiplication scheme in parallel on an AMD MI355X");
    parser.add("--schemes", ArgType::Natural, "INT", "number of schemes", "1024");
    parser.add("--max-iterations", ArgType::Natural, "INT", "number of flips per iterations", "10000");
    parser.add("--path", ArgType::String, "PATH", "path to save schemes", "schemes");
    parser.add("--input-path", ArgType::String, "PATH", "path to init scheme(s)", "");
    parser.add("--block-size", ArgType::Natural, "INT", "walkers per workgroup (accepted; the runtime picks 32 or 64 with the LDS layout)", "32");
    parser.add("--target-complexity", ArgType::Natural, "INT", "target complexity", "0");
    parser.add("--max-no-improvements", ArgType::Natural, "INT", "max iterations without improvements", "3");
    parser.add("--seed", ArgType::Natural, "INT", "random seed", "0");
    parser.add("--ring", ArgType::String, "RING", "coefficient ring, Z2 or ZT", "Z2");

    if (!parser.parse(argc, argv)) return 0;

    const int walkers = std::stoi(parser.get("--schemes"));
    const int max_iterations = std::stoi(parser.get("--max-iterations"));
    const std::string path = parser.get("--path");
    const std::string input_path = parser.get("--input-path");
    const int block_size = std::stoi(parser.get("--block-size"));
    const int target_complexity = std::stoi(parser.get("--target-complexity"));
    const int max_no_improvements = std::stoi(parser.get("--max-no-improvements"));
    long long seed = std::stoll(parser.get("--seed"));
    const std::string ring_s = parser.get("--ring");
    if (ring_s != "Z2" && ring_s != "ZT") {
        std::cout << "Unknown ring \"" << ring_s << "\" (expected Z2 or ZT)" << std::endl;
        return 0;
    }
    const int ring = ring_s == "ZT" ? FG_RING_ZT : FG_RING_Z2;
    if (seed == 0) seed = (long long)time(nullptr);  // main_complexity_minimizer.cu:30-31

    std::vector<int32_t> ints;
    std::string err;
    if (!fg::read_int_file(input_path, &ints, &err)) {
        std::cout << "Unable to open file \"" << input_path << "\"" << std::endl;
        return -1;
    }

    std::cout << "Start complexity minimization algorithm" << std::endl;
    std::cout << "- working ring: " << fgcli::ring_name(ring) << std::endl;
    std::cout << "- schemes count: " << walkers << std::endl;
    std::cout << "- max iterations: " << max_iterations << std::endl;
    std::cout << "- path: " << path << std::endl;
    std::cout << "- block size: " << block_size << std::endl;
    std::cout << "- target complexity: " << target_complexity << std::endl;
    std::cout << "- max no improvements: " << max_no_improvements << std::endl;
    std::cout << "- seed: " << seed << std::endl;

    std::vector<int32_t> blob;
    int count = 0, n[3], m = 0;
    if (!read_input(ints, &blob, &count, n, &m, &err)) {
        std::cout << err << std::endl << "Read invalid scheme" << std::endl;
        return -1;
    }
    std::cout << "Start reading " << count << " schemes (" << n[0] << "x" << n[1] << "x" << n[2] << ": " << m << ")"
              << std::endl;

    fg_config cfg{};
    cfg.ring = ring;
    cfg.walkers = walkers;
    cfg.max_iterations = max_iterations;
    cfg.check_reduce = 0;
    cfg.seed = (uint64_t)seed;
    Minimizer z;
    z.ring = ring;
    z.walkers = walkers;
    z.path = path;
    if (fg_create(&cfg, &z.ctx) < 0) {
        std::cout << "Error: " << fg_last_error() << std::endl;
        return -1;
    }
    const int rc = fg_init_minimizer(z.ctx, blob.data(), (int64_t)blob.size());
    if (rc < 0) {
        std::cout << fg_last_error() << std::endl << "Read invalid scheme" << std::endl;
        fg_destroy(z.ctx);
        return -1;
    }
    std::cout << "Successfully read " << count << " schemes" << std::endl;

    z.indices.resize(walkers);
    std::iota(z.indices.begin(), z.indices.end(), 0);
    z.best.resize(walkers);
    z.cur.resize(walkers);
    if (!z.refresh()) {
        fg_destroy(z.ctx);
        return -1;
    }
    z.best_complexity = z.best[0];  // every walker starts at the lowest input complexity

    // minimize (complexity_minimizer.cu:39-65)
    const auto start = std::chrono::steady_clock::now();
    std::vector<double> times;
    int no_improvements = 0;
    for (int iteration = 1; z.best_complexity > target_complexity; ++iteration) {
        const auto t1 = std::chrono::steady_clock::now();
        const int bad = fg_minimize(z.ctx);
        if (bad < 0) {
            std::cout << "Error: " << fg_last_error() << std::endl;
            fg_destroy(z.ctx);
            return -1;
        }
        if (bad > 0) std::cout << "invalid (" << bad << ") schemes (minimize)" << std::endl;
        if (!z.refresh()) {
            fg_destroy(z.ctx);
            return -1;
        }
        const bool improved = z.update_best(iteration);
        const auto t2 = std::chrono::steady_clock::now();
        times.push_back(std::chrono::duration<double>(t2 - t1).count());
        if (!z.report(start, iteration, times)) {
            fg_destroy(z.ctx);
            return -1;
        }
        if (improved) {
            no_improvements = 0;
        } else {
            ++no_improvements;
            std::cout << "No improvements for " << no_improvements << " iterations" << std::endl;
            if (no_improvements >= max_no_improvements) break;
        }
    }
    fg_destroy(z.ctx);
    std::cout << "Success!" << std::endl;
    return 0;
}
