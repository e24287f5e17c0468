#include "cli_common.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace fgcli {

void ArgParser::add(const std::string &flag, ArgType type, const std::string &meta, const std::string &description,
                    const std::string &def) {
    args_.push_back({flag, meta, description, def, type});
}

static bool is_natural(const std::string &v) {  // arg_parser.cu:108-114: digits only
    for (char c : v)
        if (c < '0' || c > '9') return false;
    return true;
}

static bool is_real(const std::string &v) {  // arg_parser.cu:116-135
    if (v.empty()) return false;
    bool point = false;
    for (size_t i = v[0] == '-' ? 1 : 0; i < v.size(); ++i) {
        if (v[i] == '.') {
            if (point) return false;
            point = true;
        } else if (v[i] < '0' || v[i] > '9') {
            return false;
        }
    }
    return true;
}

bool ArgParser::parse(int argc, char **argv) {
    if (argc == 2 && std::string(argv[1]) == "--help") {
        help();
        return false;
    }
    if (argc % 2 != 1) {
        std::cerr << "Invalid number of arguments" << std::endl;
        help();
        return false;
    }
    std::unordered_map<std::string, const Arg *> byname;
    for (const Arg &a : args_) {
        byname[a.name] = &a;
        if (!a.def.empty()) values_[a.name] = a.def;
    }
    for (int i = 1; i < argc; i += 2) {
        const std::string flag = argv[i];
        auto it = byname.find(flag);
        if (it == byname.end()) {
            std::cerr << "unknown argument \"" << flag << "\"" << std::endl;
            return false;
        }
        const std::string value = argv[i + 1];
        const Arg &a = *it->second;
        if (a.type == ArgType::Natural && (value.empty() || !is_natural(value))) {
            std::cerr << "value for arg \"" << a.name << "\" is not natural (" << value << ")" << std::endl;
            return false;
        }
        if (a.type == ArgType::Real && !is_real(value)) {
            std::cerr << "value for arg \"" << a.name << "\" is not real (" << value << ")" << std::endl;
            return false;
        }
        values_[flag] = value;
    }
    for (const Arg &a : args_) {
        if (a.def.empty() && values_.find(a.name) == values_.end()) {
            std::cerr << "no value for argument \"" << a.name << "\"" << std::endl;
            return false;
        }
    }
    return true;
}

void ArgParser::help() const {
    std::cout << description_ << std::endl;
    std::cout << "Usage: ./" << name_;
    for (const Arg &a : args_) {
        if (!a.def.empty()) std::cout << " [" << a.name << " " << a.def << "]";
        else std::cout << " " << a.name << " " << a.meta;
    }
    std::cout << std::endl << std::endl << "Arguments description:" << std::endl;
    for (const Arg &a : args_) {
        std::cout << a.name << ": " << a.description;
        if (!a.def.empty()) std::cout << " (default: " << a.def << ")";
        std::cout << std::endl;
    }
}

std::string ArgParser::get(const std::string &flag) const {
    auto it = values_.find(flag);
    if (it == values_.end()) throw std::runtime_error("no parsed value for argument \"" + flag + "\"");
    return it->second;
}

namespace {

// Best known ranks of n1 x n2 x n3 matrix multiplication, n1 <= n2 <= n3:
// one row per (n1, n2), ranks for n3 = n2, n2 + 1, ...
struct KnownRow {
    int n1, n2;
    std::vector<int> ranks;
};

const std::vector<KnownRow> &known_rows() {
    static const std::vector<KnownRow> rows = {
        {2, 2, {7, 11, 14, 18, 21, 25, 28, 32, 35, 39, 42, 46, 49, 53, 56}},
        {2, 3, {15, 20, 25, 30, 35, 40, 45, 50, 55, 60, 65, 70, 75, 80}},
        {2, 4, {26, 32, 39, 45, 51, 59, 64, 71, 77, 83, 90, 96, 102}},
        {2, 5, {40, 47, 55, 63, 72, 79, 87, 94}},
        {2, 6, {56, 66, 75, 86, 94}},
        {2, 7, {76, 88, 99}},
        {2, 8, {100}},
        {3, 3, {23, 29, 36, 40, 49, 55, 63, 69, 76, 80, 89, 95, 103, 109}},
        {3, 4, {38, 47, 54, 63, 73, 83, 92, 101, 108, 117, 126, 136, 146}},
        {3, 5, {58, 68, 79, 90, 104, 115, 126, 136}},
        {3, 6, {80, 94, 108, 120, 134}},
        {3, 7, {111, 126, 142}},
        {3, 8, {145}},
        {4, 4, {48, 61, 73, 85, 96, 104, 120, 130, 142, 152, 165, 177, 189}},
        {4, 5, {76, 90, 104, 118, 136, 151, 165, 179}},
        {4, 6, {105, 123, 140, 159, 175}},
        {4, 7, {144, 164, 186}},
        {4, 8, {182}},
        {5, 5, {93, 110, 127, 144, 167, 184, 202, 220}},
        {5, 6, {130, 150, 170, 197, 217}},
        {5, 7, {176, 205, 229}},
        {5, 8, {230}},
        {6, 6, {153, 183, 203, 225, 247}},
        {6, 7, {215, 239, 268}},
        {6, 8, {266}},
        {7, 7, {249, 277, 315}},
        {7, 8, {306}},
        {8, 8, {336}},
    };
    return rows;
}

// Ring-specific entries that replace the generic ones.
const char *kIntegerOverrides =
    "2x4x5:33 2x4x10:65 2x4x13:84 2x5x7:57 2x5x8:65 2x5x10:80 2x6x6:57 2x6x7:68 2x6x8:77 2x7x7:77 2x7x8:90 "
    "2x7x9:102 3x3x6:43 3x3x7:51 3x3x8:58 3x3x9:65 3x3x10:72 3x3x11:79 3x3x12:86 3x3x13:94 3x3x14:101 "
    "3x3x15:108 3x3x16:115 3x4x6:56 3x4x7:64 3x4x8:74 3x4x9:84 3x4x10:93 3x4x11:102 3x4x12:111 3x4x13:120 "
    "3x4x14:128 3x4x15:138 3x4x16:148 3x5x6:70 3x5x7:83 3x5x8:94 3x5x9:105 3x5x10:116 3x5x11:128 3x5x12:140 "
    "3x6x6:85 3x6x7:99 3x6x8:112 3x6x9:126 3x6x10:140 3x7x7:115 3x7x8:128 3x7x9:147 3x8x8:148 4x4x4:49 "
    "4x4x9:110 4x4x10:122 4x4x11:134 4x4x12:145 4x4x13:157 4x4x14:169 4x4x15:181 4x4x16:192 4x5x9:137 "
    "4x6x9:160 4x7x7:145 4x7x9:187 5x7x8:206 5x7x9:231 6x6x10:252 7x7x7:250 7x7x8:279 7x7x9:316 7x8x8:310 "
    "8x8x8:343";
const char *kZ2Overrides =
    "2x4x5:33 2x4x10:65 2x4x13:84 2x5x10:80 2x7x9:101 3x3x6:42 3x3x8:56 3x3x9:64 3x3x10:71 3x3x11:78 "
    "3x3x12:84 3x3x13:91 3x3x14:98 3x3x15:105 3x3x16:112 3x4x7:64 3x4x13:118 3x4x14:128 3x4x15:137 "
    "3x6x6:84 3x6x7:96 3x6x9:122 3x6x10:136 3x7x8:128 3x7x9:143 4x4x4:47 4x4x5:60 4x4x8:94 4x4x9:107 "
    "4x4x11:132 4x4x12:141 4x4x13:154 4x4x14:167 4x4x15:179 4x4x16:188 4x5x5:73 4x5x6:89 4x5x9:133 "
    "4x5x10:146 4x5x11:162 4x5x12:177 4x7x9:187 5x5x9:166 5x5x10:183 5x5x11:200 5x5x12:217 6x6x10:252 "
    "7x7x7:248 7x7x8:275 7x7x9:313 7x8x8:302 8x8x8:329";

std::unordered_map<std::string, int> build_table(int ring) {
    std::unordered_map<std::string, int> t;
    for (const KnownRow &r : known_rows())
        for (size_t q = 0; q < r.ranks.size(); ++q)
            t[std::to_string(r.n1) + "x" + std::to_string(r.n2) + "x" + std::to_string(r.n2 + (int)q)] = r.ranks[q];
    std::istringstream is(ring ? kIntegerOverrides : kZ2Overrides);
    std::string item;
    while (is >> item) {
        const size_t colon = item.find(':');
        t[item.substr(0, colon)] = std::stoi(item.substr(colon + 1));
    }
    return t;
}

}  // namespace

int known_rank(const std::string &sorted_key, int ring) {
    static const std::unordered_map<std::string, int> z2 = build_table(0), zt = build_table(1);
    const auto &t = ring ? zt : z2;
    auto it = t.find(sorted_key);
    return it == t.end() ? 0 : it->second;
}

std::string pretty_time(double elapsed) {  // utils.cu:31-48
    std::ostringstream ss;
    if (elapsed < 60) {
        ss << std::setprecision(3) << std::fixed << elapsed;
    } else {
        const int seconds = (int)(elapsed + 0.5);
        ss << std::setw(2) << std::setfill('0') << seconds / 3600 << ":";
        ss << std::setw(2) << std::setfill('0') << (seconds % 3600) / 60 << ":";
        ss << std::setw(2) << std::setfill('0') << seconds % 60;
    }
    return ss.str();
}

std::string pretty_int(long long value) {  // utils.cu:50-61
    std::ostringstream ss;
    if (value < 1000) ss << value;
    else if (value < 1000000) ss << std::setprecision(2) << std::fixed << (value / 1000.0) << "K";
    else ss << std::setprecision(2) << std::fixed << (value / 1000000.0) << "M";
    return ss.str();
}

const char *ring_name(int ring) { return ring ? "ZT" : "Z2"; }

bool make_dirs(const std::string &path) {
    if (path.empty()) return true;
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
        }
        if (i < path.size()) cur += path[i];
    }
    return true;
}

std::string save_path(const std::string &dir, int ring, const fg::Scheme &s, int iteration, int run_id) {
    std::ostringstream ss;
    ss << dir << "/" << fg::format_key(s.n[0], s.n[1], s.n[2]) << "_m" << s.m << "_c" << fg::complexity(s)
       << "_iteration" << iteration << "_run" << run_id << "_" << fg::format_key(s.n[0], s.n[1], s.n[2], false) << "_"
       << ring_name(ring) << ".json";
    return ss.str();
}

bool verified(int ring, const fg::Scheme &s) {
    std::vector<int32_t> blob{1};
    fg::scheme_to_blob(s, &blob);
    int32_t ok = 0;
    const int good = fg_verify(ring, blob.data(), (int64_t)blob.size(), &ok);
    return good == 1 && ok && fg::validate(s);
}

bool fetch_scheme(fg_ctx *ctx, int ring, int walker, int which, fg::Scheme *out) {
    std::vector<int32_t> buf(1 << 20);
    const int n = fg_walker_scheme(ctx, walker, which, buf.data(), (int64_t)buf.size());
    if (n < 0) return false;
    int64_t pos = 0;
    std::string err;
    return fg::scheme_from_blob(ring, buf.data(), n, &pos, out, &err);
}

}  // namespace fgcli
