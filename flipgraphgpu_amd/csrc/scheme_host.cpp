#include "scheme_host.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace fg {

void Scheme::resize(int terms) {
    for (int c = 0; c < 3; ++c) {
        v[c].assign(terms, 0);
        s[c].assign(terms, 0);
    }
}

Scheme naive_scheme(int ring, int n1, int n2, int n3) {
    Scheme s;
    s.ring = ring;
    s.n[0] = n1; s.n[1] = n2; s.n[2] = n3;
    s.nn[0] = n1 * n2; s.nn[1] = n2 * n3; s.nn[2] = n3 * n1;
    s.m = n1 * n2 * n3;
    s.resize(s.m);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n3; ++j)
            for (int k = 0; k < n2; ++k) {
                const int t = (i * n3 + j) * n2 + k;
                s.v[0][t] = 1ull << (i * n2 + k);
                s.v[1][t] = 1ull << (k * n3 + j);
                s.v[2][t] = 1ull << (j * n1 + i);
            }
    return s;
}

static bool first_positive(uint64_t v, uint64_t sg) { return v == 0 || ((v & ~(v & (v - 1))) & ~sg) != 0; }

void fix_signs(Scheme *s) {
    for (int t = 0; t < s->m; ++t) {
        const bool pi = first_positive(s->v[0][t], s->s[0][t]);
        const bool pj = first_positive(s->v[1][t], s->s[1][t]);
        if (pi && pj) continue;
        auto inv = [&](int c) { s->s[c][t] = ~s->s[c][t] & s->v[c][t]; };
        if (!pi && !pj) { inv(0); inv(1); }
        else if (!pi) { inv(0); inv(2); }
        else { inv(1); inv(2); }
    }
}

bool scheme_from_blob(int ring, const int32_t *blob, int64_t len, int64_t *pos, Scheme *out, std::string *err) {
    int64_t p = *pos;
    if (p + 4 > len) { *err = "truncated scheme header"; return false; }
    Scheme s;
    s.ring = ring;
    s.n[0] = blob[p]; s.n[1] = blob[p + 1]; s.n[2] = blob[p + 2]; s.m = blob[p + 3];
    p += 4;
    for (int c = 0; c < 3; ++c) s.nn[c] = s.n[c] * s.n[(c + 1) % 3];
    if (s.n[0] < 1 || s.n[1] < 1 || s.n[2] < 1 || s.nn[0] > 64 || s.nn[1] > 64 || s.nn[2] > 64 || s.m < 0 || s.m > 350) {
        char buf[160];
        snprintf(buf, sizeof buf, "invalid scheme sizes (%d, %d, %d, %d)", s.n[0], s.n[1], s.n[2], s.m);
        *err = buf;
        return false;
    }
    const int64_t need = (int64_t)s.m * (s.nn[0] + s.nn[1] + s.nn[2]);
    if (p + need > len) { *err = "truncated scheme coefficients"; return false; }
    s.resize(s.m);
    for (int c = 0; c < 3; ++c)
        for (int t = 0; t < s.m; ++t)
            for (int q = 0; q < s.nn[c]; ++q) {
                const int x = blob[p++];
                const uint64_t bit = 1ull << q;
                if (ring == 0) {
                    if ((x < 0 ? -x : x) % 2) s.v[c][t] |= bit;
                } else if (x == 1) {
                    s.v[c][t] |= bit;
                } else if (x == -1) {
                    s.v[c][t] |= bit;
                    s.s[c][t] |= bit;
                } else if (x != 0) {
                    *err = "integer coefficient outside {-1, 0, 1}";
                    return false;
                }
            }
    if (ring == 1) fix_signs(&s);
    *pos = p;
    *out = std::move(s);
    return true;
}

void scheme_to_blob(const Scheme &s, std::vector<int32_t> *out) {
    out->push_back(s.n[0]);
    out->push_back(s.n[1]);
    out->push_back(s.n[2]);
    out->push_back(s.m);
    for (int c = 0; c < 3; ++c)
        for (int t = 0; t < s.m; ++t)
            for (int q = 0; q < s.nn[c]; ++q) {
                const int bit = (int)((s.v[c][t] >> q) & 1);
                const int neg = (int)((s.s[c][t] >> q) & 1);
                out->push_back(s.ring == 1 && neg ? -bit : bit);
            }
}

bool validate(const Scheme &s) {
    for (int i = 0; i < s.nn[0]; ++i)
        for (int j = 0; j < s.nn[1]; ++j) {
            // bit-sliced over the W index: one pass over the terms per (i, j)
            std::vector<int> acc(s.nn[2], 0);
            uint64_t accz2 = 0;
            for (int t = 0; t < s.m; ++t) {
                if (!((s.v[0][t] >> i) & (s.v[1][t] >> j) & 1)) continue;
                if (s.ring == 0) {
                    accz2 ^= s.v[2][t];
                } else {
                    const int sg = (int)(((s.s[0][t] >> i) ^ (s.s[1][t] >> j)) & 1);
                    for (int k = 0; k < s.nn[2]; ++k)
                        if ((s.v[2][t] >> k) & 1) acc[k] += (sg ^ (int)((s.s[2][t] >> k) & 1)) ? -1 : 1;
                }
            }
            const int i1 = i / s.n[1], i2 = i % s.n[1], j1 = j / s.n[2], j2 = j % s.n[2];
            for (int k = 0; k < s.nn[2]; ++k) {
                const int k1 = k / s.n[0], k2 = k % s.n[0];
                const int target = (i2 == j1) && (i1 == k2) && (j2 == k1);
                const int got = s.ring == 0 ? (int)((accz2 >> k) & 1) : acc[k];
                if (got != target) return false;
            }
        }
    return true;
}

int complexity(const Scheme &s) {
    int count = 0;
    for (int t = 0; t < s.m; ++t)
        for (int c = 0; c < 3; ++c) count += __builtin_popcountll(s.v[c][t]);
    return count - 2 * s.m - s.nn[2];
}

namespace {

struct Tv {
    uint64_t v, s;
    bool ok;
};
Tv tv(const Scheme &x, int c, int t) { return {x.v[c][t], x.s[c][t], true}; }
Tv tv_add(Tv a, Tv b) {
    Tv r;
    r.v = a.v ^ b.v;
    r.s = ((a.s & a.v) | (b.s & b.v)) & r.v;
    r.ok = !(a.v & b.v & ~(a.s ^ b.s));
    return r;
}
Tv tv_sub(Tv a, Tv b) {
    Tv r;
    r.v = a.v ^ b.v;
    r.s = ((a.s & a.v) | (~b.s & b.v)) & r.v;
    r.ok = !(a.v & b.v & (a.s ^ b.s));
    return r;
}
int tv_cmp(Tv a, Tv b) {
    if (a.v != b.v) return 0;
    if (a.s == b.s) return 1;
    if (a.s == (~b.s & b.v)) return -1;
    return 0;
}
bool limit_sum(Tv a, Tv b, bool fp) {
    if (a.v & b.v & ~(a.s ^ b.s)) return false;
    if (!fp) return true;
    const uint64_t sv = a.v ^ b.v;
    return first_positive(sv, ((a.s & a.v) | (b.s & b.v)) & sv);
}
bool limit_sub(Tv a, Tv b, bool fp) {
    if (a.v & b.v & (a.s ^ b.s)) return false;
    if (!fp) return true;
    const uint64_t sv = a.v ^ b.v;
    return first_positive(sv, ((a.s & a.v) | (~b.s & b.v)) & sv);
}
bool eq(const Scheme &x, int c, int a, int b) { return x.v[c][a] == x.v[c][b] && x.s[c][a] == x.s[c][b]; }

void copy_term(Scheme *s, int dst, int src) {
    for (int c = 0; c < 3; ++c) {
        s->v[c][dst] = s->v[c][src];
        s->s[c][dst] = s->s[c][src];
    }
}

void remove_zeroes(Scheme *s) {
    auto nz = [&](int t) { return s->v[0][t] && s->v[1][t] && s->v[2][t]; };
    if (s->ring == 0) {  // scheme_z2.cu:152-164
        while (s->m > 0 && !nz(s->m - 1)) --s->m;
        for (int t = 0; t < s->m; ++t)
            if (!nz(t)) { --s->m; copy_term(s, t, s->m); }
    } else {  // scheme_integer.cu:237-249
        for (int t = 0; t < s->m; ++t) {
            if (nz(t)) continue;
            --s->m;
            copy_term(s, t, s->m);
            --t;
        }
    }
}

void reduce(Scheme *s, int i, int a, int b, bool sub) {
    bool zero;
    if (s->ring == 0) {
        s->v[i][a] ^= s->v[i][b];
        zero = s->v[i][a] == 0;
    } else {
        Tv r = sub ? tv_sub(tv(*s, i, a), tv(*s, i, b)) : tv_add(tv(*s, i, a), tv(*s, i, b));
        s->v[i][a] = r.v;
        s->s[i][a] = r.s;
        zero = r.v == 0;
    }
    --s->m;
    if (b != s->m) copy_term(s, b, s->m);
    if (zero) remove_zeroes(s);
}

}  // namespace

bool try_reduce(Scheme *s) {
    const int m = s->m;
    if (s->ring == 0) {
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b) {
                if (!eq(*s, 0, a, b)) continue;
                if (s->v[1][a] == s->v[1][b]) { reduce(s, 2, a, b, false); return true; }
                if (s->v[2][a] == s->v[2][b]) { reduce(s, 1, a, b, false); return true; }
            }
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b)
                if (eq(*s, 1, a, b) && s->v[2][a] == s->v[2][b]) { reduce(s, 0, a, b, false); return true; }
        return false;
    }
    for (int a = 0; a < m; ++a)
        for (int b = a + 1; b < m; ++b) {
            if (!eq(*s, 0, a, b)) continue;
            if (eq(*s, 1, a, b) && limit_sum(tv(*s, 2, a), tv(*s, 2, b), false)) { reduce(s, 2, a, b, false); return true; }
            const int c2 = tv_cmp(tv(*s, 2, a), tv(*s, 2, b));
            if (c2 == 1 && limit_sum(tv(*s, 1, a), tv(*s, 1, b), true)) { reduce(s, 1, a, b, false); return true; }
            if (c2 == -1 && limit_sub(tv(*s, 1, a), tv(*s, 1, b), true)) { reduce(s, 1, a, b, true); return true; }
        }
    for (int a = 0; a < m; ++a)
        for (int b = a + 1; b < m; ++b) {
            if (!eq(*s, 1, a, b)) continue;
            const int c2 = tv_cmp(tv(*s, 2, a), tv(*s, 2, b));
            if (c2 == 1 && limit_sum(tv(*s, 0, a), tv(*s, 0, b), true)) { reduce(s, 0, a, b, false); return true; }
            if (c2 == -1 && limit_sub(tv(*s, 0, a), tv(*s, 0, b), true)) { reduce(s, 0, a, b, true); return true; }
        }
    return false;
}

static void json_matrix(std::ostringstream &f, const char *name, const Scheme &s, int c) {
    f << "    \"" << name << "\": [\n";
    for (int t = 0; t < s.m; ++t) {
        f << "        [";
        for (int q = 0; q < s.nn[c]; ++q) {
            if (q) f << ", ";
            const int bit = (int)((s.v[c][t] >> q) & 1);
            const int neg = (int)((s.s[c][t] >> q) & 1);
            f << (s.ring == 1 && neg ? -bit : bit);
        }
        f << "]" << (t < s.m - 1 ? "," : "") << "\n";
    }
    f << "    ]";
}

std::string to_json(const Scheme &s) {
    std::ostringstream f;
    f << "{\n";
    f << "    \"n\": [" << s.n[0] << ", " << s.n[1] << ", " << s.n[2] << "],\n";
    f << "    \"m\": " << s.m << ",\n";
    f << "    \"z2\": " << (s.ring == 0 ? "true" : "false") << ",\n";
    f << "    \"complexity\": " << complexity(s) << ",\n";
    json_matrix(f, "u", s, 0);
    f << ",\n";
    json_matrix(f, "v", s, 1);
    f << ",\n";
    json_matrix(f, "w", s, 2);
    f << "\n}\n";
    return f.str();
}

bool save_json(const Scheme &s, const std::string &path) {
    std::ofstream f(path);
    if (!f) return false;
    f << to_json(s);
    return (bool)f;
}

std::string format_key(int n1, int n2, int n3, bool sorted, bool with_zeros) {
    int n[3] = {n1, n2, n3};
    if (sorted) std::sort(n, n + 3);
    char buf[64];
    if (with_zeros) snprintf(buf, sizeof buf, "%02dx%02dx%02d", n[0], n[1], n[2]);
    else snprintf(buf, sizeof buf, "%dx%dx%d", n[0], n[1], n[2]);
    return buf;
}

bool read_int_file(const std::string &path, std::vector<int32_t> *out, std::string *err) {
    std::ifstream f(path);
    if (!f) { *err = "cannot open " + path; return false; }
    long long x;
    while (f >> x) out->push_back((int32_t)x);
    return true;
}

}  // namespace fg
