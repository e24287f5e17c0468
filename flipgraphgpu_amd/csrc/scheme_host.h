// Host-side scheme handling for the CLIs and the runtime: reading the
// reference's text/blob formats, the Brent check that re-verifies GPU
// results, tryReduce after input, complexity and the JSON writer.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fg {

struct Scheme {
    int ring = 0;  // 0 = Z2, 1 = ZT
    int n[3] = {0, 0, 0};
    int nn[3] = {0, 0, 0};
    int m = 0;
    std::vector<uint64_t> v[3];  // Z2 mask / ZT magnitudes
    std::vector<uint64_t> s[3];  // ZT signs

    void resize(int terms);
    int naive_rank() const { return n[0] * n[1] * n[2]; }
};

// Naive scheme (src/schemes/scheme_z2.cu:32-55).
Scheme naive_scheme(int ring, int n1, int n2, int n3);
// One scheme from an int32 blob [n1 n2 n3 m U V W]; advances *pos.  Z2 takes
// |x| mod 2 (scheme_z2.cu:89); ZT rejects |x| > 1 (addition.cu:41-60) and
// applies fixSigns (scheme_integer.cu:398-425).  Returns false on bad input.
bool scheme_from_blob(int ring, const int32_t *blob, int64_t len, int64_t *pos, Scheme *out, std::string *err);
void scheme_to_blob(const Scheme &s, std::vector<int32_t> *out);
// Brent equations (scheme_z2.cu:3-29, scheme_integer.cu:3-38).
bool validate(const Scheme &s);
// getComplexity (scheme_z2.cu:127-139).
int complexity(const Scheme &s);
// tryReduce over the freshly built pair list (scheme_z2.cu:1127-1154,
// scheme_integer.cu:1062-1101); returns true if two terms were merged.
bool try_reduce(Scheme *s);
void fix_signs(Scheme *s);
// JSON writer (scheme_z2.cu:1327-1364, scheme_integer.cu:1247-1274).
std::string to_json(const Scheme &s);
bool save_json(const Scheme &s, const std::string &path);
// "2x2x3" keys (src/utils/utils.cu:3-25).
std::string format_key(int n1, int n2, int n3, bool sorted = true, bool with_zeros = false);
// Text reader for --input-path files: whitespace separated integers.
bool read_int_file(const std::string &path, std::vector<int32_t> *out, std::string *err);

}  // namespace fg
