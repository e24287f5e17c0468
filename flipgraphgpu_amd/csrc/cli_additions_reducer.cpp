// additions_reducer: the reference's additions reducer CLI
// (src/main_additions_reducer.cu, src/entities/scheme_additions_reducer.cu)
// on the MI355X engine.  Same flags, defaults, input format, report table,
// stop rule and JSON output; the reducers run through the fg.h C ABI.
//
// Not built: --max-flips > 0 (scheme flips between passes) is rejected.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "cli_common.h"
#include "fg.h"

using fgcli::ArgType;

namespace {

struct PartView {
    int count = 0, real = 0, fresh = 0, naive = 0, mode = 0, additions = 0;
    float scale = 0;
    std::vector<int> vars;
    std::vector<std::vector<int>> expr;
};

PartView decode(const std::vector<int32_t> &b) {
    PartView p;
    p.count = b[0];
    p.real = b[1];
    p.fresh = b[2];
    p.naive = b[3];
    p.mode = b[4];
    memcpy(&p.scale, &b[5], 4);
    p.additions = b[6];
    size_t q = 7;
    p.vars.assign(b.begin() + q, b.begin() + q + 2 * p.fresh);
    q += 2 * p.fresh;
    for (int e = 0; e < p.count; ++e) {
        const int n = b[q++];
        p.expr.emplace_back(b.begin() + q, b.begin() + q + n);
        q += n;
    }
    return p;
}

// AdditionsReducer::write (additions_reducer.cuh:296-347)
void write_part(std::ostream &os, const PartView &p, const std::string &name, const std::string &indent) {
    os << indent << "\"" << name << "_fresh\": [" << std::endl;
    for (int i = 0; i < p.fresh; ++i) {
        const int a = p.vars[2 * i], b = p.vars[2 * i + 1];
        os << indent << indent << "[{\"index\": " << (std::abs(a) - 1) << ", \"value\": " << (a > 0 ? 1 : -1)
           << "}, {\"index\": " << (std::abs(b) - 1) << ", \"value\": " << (b > 0 ? 1 : -1) << "}]";
        if (i < p.fresh - 1) os << ",";
        os << std::endl;
    }
    os << indent << "]," << std::endl;
    os << indent << "\"" << name << "\": [" << std::endl;
    for (int i = 0; i < p.count; ++i) {
        os << indent << indent << "[";
        for (size_t j = 0; j < p.expr[i].size(); ++j) {
            const int x = p.expr[i][j];
            os << "{\"index\": " << (std::abs(x) - 1) << ", \"value\": " << (x > 0 ? 1 : -1) << "}";
            if (j + 1 < p.expr[i].size()) os << ", ";
        }
        os << "]";
        if (i < p.count - 1) os << ",";
        os << std::endl;
    }
    os << indent << "]";
}

bool fetch_best(fg_reducer *r, int c, PartView *out) {
    std::vector<int32_t> buf(1 << 20);
    const int64_t n = fg_reducer_best_part(r, c, buf.data(), (int64_t)buf.size());
    if (n < 0) return false;
    buf.resize((size_t)n);
    *out = decode(buf);
    return true;
}

std::string mode_str(int mode, float scale) {
    char buf[16];
    fg_reducer_mode_name(mode, scale, buf, sizeof buf);
    return buf;
}

}  // namespace

int main(int argc, char *argv[]) {
    fgcli::ArgParser parser("additions_reducer",
                            "Find best additions number of the fast matrix multiplication scheme in parallel on an AMD MI355X");
    parser.add("-i", ArgType::String, "PATH", "path to init scheme(s)", "");
    parser.add("-o", ArgType::String, "PATH", "path to save schemes", "schemes/additions_reduced");
    parser.add("--count", ArgType::Natural, "INT", "number of parallel reducers", "32");
    parser.add("--schemes-count", ArgType::Natural, "INT", "number of parallel schemes", "1");
    parser.add("--block-size", ArgType::Natural, "INT", "number of reducers per workgroup request (the kernel runs one wave per reducer component)", "32");
    parser.add("--start-additions", ArgType::Natural, "INT", "upper bound of additions for check optimality", "0");
    parser.add("--max-flips", ArgType::Natural, "INT", "number of scheme flips", "0");
    parser.add("--max-no-improvements", ArgType::Natural, "INT", "max iterations without improvements", "3");
    parser.add("--seed", ArgType::Natural, "INT", "random seed", "0");

    if (!parser.parse(argc, argv)) return 0;

    const std::string input_path = parser.get("-i");
    const std::string output_path = parser.get("-o");
    const int count = std::stoi(parser.get("--count"));
    const int schemes_count = std::stoi(parser.get("--schemes-count"));
    const int block_size = std::stoi(parser.get("--block-size"));
    const int max_no_improvements = std::stoi(parser.get("--max-no-improvements"));
    const int start_additions = std::stoi(parser.get("--start-additions"));
    const int max_flips = std::stoi(parser.get("--max-flips"));
    long long seed = std::stoll(parser.get("--seed"));
    if (seed == 0) seed = (long long)time(nullptr);  // main_additions_reducer.cu:35-36

    std::vector<int32_t> ints;
    std::string err;
    if (!fg::read_int_file(input_path, &ints, &err)) {
        std::cout << "Unable to open file \"" << input_path << "\"" << std::endl;
        return -1;
    }

    std::cout << "Compiled configuration:" << std::endl;
    std::cout << "- max expressions count (U / V / W): 350 / 350 / 64" << std::endl;
    std::cout << "- max real variables (U / V / W): 32 / 32 / 175" << std::endl;
    std::cout << "- max subexpressions (U / V / W): 2016 / 2016 / 61075" << std::endl;
    std::cout << "- max fresh variables (U / V / W): 250 / 250 / 500" << std::endl;
    std::cout << std::endl;
    std::cout << "Start additions reduction algorithm" << std::endl;
    std::cout << "- count: " << count << std::endl;
    std::cout << "- schemes count: " << schemes_count << std::endl;
    std::cout << "- output path: " << output_path << std::endl;
    std::cout << "- block size: " << block_size << std::endl;
    if (start_additions > 0) std::cout << "- start additions: " << start_additions << std::endl;
    if (max_flips > 0) std::cout << "- max flips: " << max_flips << std::endl;
    std::cout << "- max no improvements: " << max_no_improvements << std::endl;
    std::cout << "- seed: " << seed << std::endl;
    std::cout << std::endl;

    if (ints.size() < 4) {
        std::cout << "Error: readed scheme is invalid" << std::endl;
        return -1;
    }
    const int n1 = ints[0], n2 = ints[1], n3 = ints[2], m = ints[3];
    std::cout << "Start reading scheme " << n1 << "x" << n2 << "x" << n3 << " with " << m << " multiplications"
              << std::endl;
    fg_reducer *r = nullptr;
    if (fg_reducer_create(ints.data(), (int64_t)ints.size(), count, schemes_count, max_flips, (uint64_t)seed, &r) < 0) {
        std::cout << fg_last_error() << std::endl;
        return -1;
    }
    const std::string dims = std::to_string(n1) + "x" + std::to_string(n2) + "x" + std::to_string(n3);

    // initialize (scheme_additions_reducer.cu:144-175)
    int32_t best[11], stats[9];
    fg_reducer_best(r, best);
    fg_reducer_scheme_stats(r, stats);
    std::cout << "Readed scheme params:" << std::endl;
    std::cout << "- dimensions: " << dims << std::endl;
    std::cout << "- multiplications (rank): " << m << std::endl;
    std::cout << "- naive additions (U / V / W / total): " << best[0] << " / " << best[1] << " / " << best[2] << " / "
              << best[9] << std::endl;
    std::cout << "- max real variables (U / V / W): " << stats[3] << " / " << stats[4] << " / " << stats[5] << std::endl;
    std::cout << "- max unique subexpressions (U / V / W): " << stats[6] << " / " << stats[7] << " / " << stats[8]
              << std::endl;
    std::cout << std::endl;

    // reduce (:103-130)
    const auto start = std::chrono::steady_clock::now();
    std::vector<double> times;
    int no_improvements = 0;
    for (int iteration = 1; no_improvements < max_no_improvements; ++iteration) {
        const auto t1 = std::chrono::steady_clock::now();
        int32_t improved = 0;
        if (fg_reducer_run(r, iteration) < 0 || fg_reducer_update(r, &improved) < 0) {
            std::cout << "Error: " << fg_last_error() << std::endl;
            fg_reducer_destroy(r);
            return -1;
        }
        int32_t prev[11];
        memcpy(prev, best, sizeof best);
        fg_reducer_best(r, best);
        if (improved) {  // updateBest (:287-307)
            if (improved == 1)
                std::cout << "Reduced scheme improved from " << prev[9] << " to " << best[9]
                          << " additions (fresh vars: " << best[10] << ")" << std::endl;
            else
                std::cout << "Reduced scheme improved from " << prev[10] << " fresh vars to " << best[10]
                          << " fresh vars (additions: " << prev[9] << ")" << std::endl;
            if (best[9] < start_additions || start_additions == 0) {  // save (:365-388)
                PartView parts[3];
                bool ok = true;
                for (int c = 0; c < 3; ++c) ok = ok && fetch_best(r, c, &parts[c]);
                const int naive = parts[0].naive + parts[1].naive + parts[2].naive;
                std::ostringstream path;
                path << output_path << "/" << dims << "_m" << m << "_cr" << best[9] << "_fv" << best[10] << "_cn"
                     << naive << "_ZT_reduced.json";
                fgcli::make_dirs(output_path);
                std::ofstream f(path.str());
                if (!ok || !f) {
                    std::cout << "Unable to save \"" << path.str() << "\"" << std::endl;
                } else {
                    f << "{" << std::endl;
                    f << "    \"n\": [" << n1 << ", " << n2 << ", " << n3 << "]," << std::endl;
                    f << "    \"m\": " << m << "," << std::endl;
                    f << "    \"z2\": false," << std::endl;
                    f << "    \"complexity\": {\"naive\": " << naive << ", \"reduced\": " << best[9] << "},"
                      << std::endl;
                    write_part(f, parts[0], "u", "    ");
                    f << "," << std::endl;
                    write_part(f, parts[1], "v", "    ");
                    f << "," << std::endl;
                    write_part(f, parts[2], "w", "    ");
                    f << std::endl;
                    f << "}" << std::endl;
                    std::cout << "Reduced scheme saved to \"" << path.str() << "\"" << std::endl;
                }
            }
        }
        const auto t2 = std::chrono::steady_clock::now();
        times.push_back(std::chrono::duration<double>(t2 - t1).count());

        // report (:309-363)
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        const double last = times.back();
        const double mn = *std::min_element(times.begin(), times.end());
        const double mx = *std::max_element(times.begin(), times.end());
        const double mean = std::accumulate(times.begin(), times.end(), 0.0) / times.size();
        std::cout << std::endl << std::left;
        std::cout << "+----------------------------------------------------------------------------------------------------------------------------+" << std::endl;
        std::cout << "| Size: " << std::setw(24) << dims << "   Reducers: " << std::setw(20) << count << "   "
                  << "                                 Iteration: " << std::setw(12) << iteration << " |" << std::endl;
        std::cout << "| Rank: " << std::setw(24) << m << "   Schemes: " << std::setw(21) << schemes_count << "   "
                  << "                                 Elapsed: " << std::setw(14) << fgcli::pretty_time(elapsed) << " |"
                  << std::endl;
        std::cout << std::right;
        std::cout << "+================================+================================+================================+=========================+" << std::endl;
        std::cout << "|           Reducers U           |           Reducers V           |           Reducers W           |          Total          |" << std::endl;
        std::cout << "+------+-------+---------+-------+------+-------+---------+-------+------+-------+---------+-------+-------+---------+-------|" << std::endl;
        std::cout << "| mode | naive | reduced | fresh | mode | naive | reduced | fresh | mode | naive | reduced | fresh | naive | reduced | fresh |" << std::endl;
        std::cout << "+------+-------+---------+-------+------+-------+---------+-------+------+-------+---------+-------+-------+---------+-------+" << std::endl;
        std::vector<int32_t> rows((size_t)10 * 3 * 6);
        const int nrows = fg_reducer_rows(r, 10, rows.data());
        for (int i = 0; i < nrows; ++i) {
            int naive = 0, reduced = 0, fresh = 0;
            std::cout << "| ";
            for (int c = 0; c < 3; ++c) {
                const int32_t *o = rows.data() + ((size_t)i * 3 + c) * 6;
                float scale;
                memcpy(&scale, &o[2], 4);
                std::cout << std::setw(4) << mode_str(o[1], scale) << "   " << std::setw(5) << o[3] << "   "
                          << std::setw(7) << o[4] << "   " << std::setw(5) << o[5] << " | ";
                naive += o[3];
                reduced += o[4];
                fresh += o[5];
            }
            std::cout << std::setw(5) << naive << "   " << std::setw(7) << reduced << "   " << std::setw(5) << fresh
                      << " | " << std::endl;
        }
        std::cout << "+--------------------------------+--------------------------------+--------------------------------+-------------------------+" << std::endl;
        std::cout << "- iteration time (last / min / max / mean): " << fgcli::pretty_time(last) << " / "
                  << fgcli::pretty_time(mn) << " / " << fgcli::pretty_time(mx) << " / " << fgcli::pretty_time(mean)
                  << std::endl;
        std::cout << "- best additions (U / V / W / total): " << best[0] << " / " << best[1] << " / " << best[2] << " / "
                  << best[9] << std::endl;
        std::cout << "- best fresh vars (U / V / W / total): " << best[3] << " / " << best[4] << " / " << best[5]
                  << " / " << best[10] << std::endl;
        std::cout << std::endl;

        if (improved) {
            no_improvements = 0;
        } else {
            ++no_improvements;
            std::cout << "No improvements for " << no_improvements << " iterations" << std::endl;
        }
    }
    fg_reducer_destroy(r);
    return 0;
}
