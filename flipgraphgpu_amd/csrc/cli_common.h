// Shared pieces of the three command-line tools: the flag parser with the
// reference's rules (src/entities/arg_parser.cu), the known-rank table
// (src/entities/flip_graph.cu:21-176) and the report formatting helpers
// (src/utils/utils.cu:31-61).
#pragma once

#include <string>
#include <unordered_map>
#include <vector>

#include "fg.h"
#include "scheme_host.h"

namespace fgcli {

enum class ArgType { String, Natural, Real };

class ArgParser {
public:
    ArgParser(std::string name, std::string description) : name_(std::move(name)), description_(std::move(description)) {}
    // An empty default marks the flag as required.
    void add(const std::string &flag, ArgType type, const std::string &meta, const std::string &description,
             const std::string &def = "");
    // Flags come as `name value` pairs; `--help` alone prints the usage.
    // Returns false (after printing why) when the program should stop.
    bool parse(int argc, char **argv);
    void help() const;
    std::string get(const std::string &flag) const;

private:
    struct Arg {
        std::string name, meta, description, def;
        ArgType type;
    };
    std::string name_, description_;
    std::vector<Arg> args_;
    std::unordered_map<std::string, std::string> values_;
};

// Best known rank for a sorted format key ("3x3x3"), ring 0 = Z2, 1 = ZT;
// 0 when the table has no entry.
int known_rank(const std::string &sorted_key, int ring);

std::string pretty_time(double seconds);
std::string pretty_int(long long value);
const char *ring_name(int ring);
// mkdir -p
bool make_dirs(const std::string &path);

// getSavePath (flip_graph.cu:386-399, complexity_minimizer.cu:136-149):
// {path}/{sorted key}_m{m}_c{complexity}_iteration{i}_run{r}_{real key}_{ring}.json
std::string save_path(const std::string &dir, int ring, const fg::Scheme &s, int iteration, int run_id);
// GPU Brent check plus host re-check of one scheme before it is written.
bool verified(int ring, const fg::Scheme &s);
// Best (which = 1) or current (which = 0) scheme of one walker.
bool fetch_scheme(fg_ctx *ctx, int ring, int walker, int which, fg::Scheme *out);

}  // namespace fgcli
