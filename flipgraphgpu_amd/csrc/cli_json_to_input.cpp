// json_to_input: turns scheme JSON files, as flip_graph and
// complexity_minimizer save them (src/schemes/scheme_z2.cu:1327-1364,
// src/schemes/scheme_integer.cu:1247-1274), back into the input text of the
// three tools, so a saved search can be resumed:
//   flip_graph --input-path:           count, then per scheme n1 n2 n3 m U V W
//                                      (src/entities/flip_graph.cu:201-212)
//   complexity_minimizer --input-path: n1 n2 n3 m count, then U V W per scheme
//                                      (src/entities/complexity_minimizer.cu:22-37)
//   additions_reducer -i:              n1 n2 n3 m, then U V W (one scheme)
//                                      (src/entities/scheme_additions_reducer.cu:29-48)
// No counterpart in the reference.
#include <cctype>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cli_common.h"

using fgcli::ArgType;

namespace {

struct JsonScheme {
    int n[3] = {0, 0, 0}, m = 0;
    std::vector<std::vector<int>> f[3];  // u, v, w: m rows of coefficients
};

// Position just after `"key":`, or npos.
size_t find_key(const std::string &t, const std::string &key) {
    const std::string k = "\"" + key + "\"";
    size_t p = t.find(k);
    if (p == std::string::npos) return p;
    p = t.find(':', p + k.size());
    return p == std::string::npos ? p : p + 1;
}

// Nested integer arrays from position p: returns the rows of the outermost
// array (a flat array is one row).
bool parse_rows(const std::string &t, size_t p, std::vector<std::vector<int>> *rows, bool flat) {
    while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
    if (p >= t.size() || t[p] != '[') return false;
    int depth = 0;
    std::vector<int> cur;
    for (; p < t.size(); ++p) {
        const char c = t[p];
        if (c == '[') {
            ++depth;
            if (depth == 2 || flat) cur.clear();
        } else if (c == ']') {
            if (depth == 2 || (flat && depth == 1)) rows->push_back(cur);
            if (--depth == 0) return true;
        } else if (c == '-' || std::isdigit((unsigned char)c)) {
            size_t q = p;
            if (t[q] == '-') ++q;
            while (q < t.size() && std::isdigit((unsigned char)t[q])) ++q;
            cur.push_back(std::stoi(t.substr(p, q - p)));
            p = q - 1;
        }
    }
    return false;
}

bool read_scheme(const std::string &path, JsonScheme *s, std::string *err) {
    std::ifstream f(path);
    if (!f) {
        *err = "Unable to open file \"" + path + "\"";
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    std::vector<std::vector<int>> n;
    size_t p = find_key(t, "n");
    if (p == std::string::npos || !parse_rows(t, p, &n, true) || n.size() != 1 || n[0].size() != 3) {
        *err = path + ": no \"n\": [n1, n2, n3]";
        return false;
    }
    for (int c = 0; c < 3; ++c) s->n[c] = n[0][c];
    p = find_key(t, "m");
    if (p == std::string::npos) {
        *err = path + ": no \"m\"";
        return false;
    }
    s->m = std::stoi(t.substr(p));
    const char *names[3] = {"u", "v", "w"};
    for (int c = 0; c < 3; ++c) {
        const int nn = s->n[c] * s->n[(c + 1) % 3];
        p = find_key(t, names[c]);
        if (p == std::string::npos || !parse_rows(t, p, &s->f[c], false) || (int)s->f[c].size() != s->m) {
            *err = path + ": \"" + names[c] + "\" is not a list of m factors (the reducer's *_reduced.json holds expressions, not a scheme)";
            return false;
        }
        for (const auto &row : s->f[c])
            if ((int)row.size() != nn) {
                *err = path + ": a factor of \"" + names[c] + "\" does not have " + std::to_string(nn) + " coefficients";
                return false;
            }
    }
    return true;
}

void write_factors(std::ostream &o, const JsonScheme &s) {
    for (int c = 0; c < 3; ++c)
        for (const auto &row : s.f[c]) {
            for (size_t q = 0; q < row.size(); ++q) o << (q ? " " : "") << row[q];
            o << "\n";
        }
}

std::vector<std::string> split(const std::string &s) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == ',') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

}  // namespace

int main(int argc, char *argv[]) {
    fgcli::ArgParser parser("json_to_input", "Convert saved scheme JSON files into the input text of flip_graph, complexity_minimizer or additions_reducer");
    parser.add("-i", ArgType::String, "PATHS", "comma-separated scheme JSON files");
    parser.add("-o", ArgType::String, "PATH", "output text file");
    parser.add("--format", ArgType::String, "TOOL", "flip_graph, complexity_minimizer or additions_reducer", "flip_graph");
    if (!parser.parse(argc, argv)) return 0;
    const std::vector<std::string> inputs = split(parser.get("-i"));
    const std::string out_path = parser.get("-o"), format = parser.get("--format");
    if (format != "flip_graph" && format != "complexity_minimizer" && format != "additions_reducer") {
        std::cout << "Unknown format \"" << format << "\"" << std::endl;
        return -1;
    }
    std::vector<JsonScheme> schemes;
    for (const auto &path : inputs) {
        JsonScheme s;
        std::string err;
        if (!read_scheme(path, &s, &err)) {
            std::cout << err << std::endl;
            return -1;
        }
        schemes.push_back(std::move(s));
    }
    if (schemes.empty()) {
        std::cout << "No input files" << std::endl;
        return -1;
    }
    if (format != "flip_graph") {
        for (const auto &s : schemes)
            if (s.n[0] != schemes[0].n[0] || s.n[1] != schemes[0].n[1] || s.n[2] != schemes[0].n[2] || s.m != schemes[0].m) {
                std::cout << "Error: " << format << " needs schemes of one format and rank" << std::endl;
                return -1;
            }
        if (format == "additions_reducer" && schemes.size() != 1) {
            std::cout << "Error: additions_reducer reads one scheme" << std::endl;
            return -1;
        }
    }
    std::ofstream o(out_path);
    if (!o) {
        std::cout << "Unable to open file \"" << out_path << "\"" << std::endl;
        return -1;
    }
    const JsonScheme &s0 = schemes[0];
    if (format == "flip_graph") {
        o << schemes.size() << "\n";
        for (const auto &s : schemes) {
            o << s.n[0] << " " << s.n[1] << " " << s.n[2] << " " << s.m << "\n";
            write_factors(o, s);
        }
    } else if (format == "complexity_minimizer") {
        o << s0.n[0] << " " << s0.n[1] << " " << s0.n[2] << " " << s0.m << " " << schemes.size() << "\n";
        for (const auto &s : schemes) write_factors(o, s);
    } else {
        o << s0.n[0] << " " << s0.n[1] << " " << s0.n[2] << " " << s0.m << "\n";
        write_factors(o, s0);
    }
    std::cout << "Wrote " << schemes.size() << " scheme(s) to \"" << out_path << "\" for " << format << std::endl;
    return 0;
}
