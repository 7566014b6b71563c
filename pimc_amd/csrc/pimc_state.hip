// pimc_state.hip -- the reference's text state file: read, write, leftPack and
// the load fix-up, so that a saved configuration (worm link topology
// included) can be scored on the device.
//
// Format (PathIntegralMonteCarlo::saveState, src/pimc.cpp:926-986, with the
// stream operators of include/common.h:244-397):
//   numWorldLines
//   %16d\t%16d              total accepted / attempted, then one line per
//   ...                     move and per estimator (loadState skips them)
//   (0,M-1) x (0,N-1)       beads: "(x,y,z) " per bead, 16 significant
//   [ ... ]                 digits, rows separated by "\n  "
//                           (an empty line after every array)
//   next links, prev links  "(slice,ptcl) " per bead, same layout
//   worm beads              "0 " / "1 " per bead
//   625 words               MTRand::save (state[624] + left), space separated
// The loader (PathIntegralMonteCarlo::loadState, src/pimc.cpp:1108-1258) takes
// the arrays as they are when the file has the run's number of slices, and
// otherwise builds a classical path from slice 0 of every active worldline;
// then it puts every bead inside the box, counts the active beads per slice
// and resets the worm (no head or tail).
#include <cerrno>
#include <cstdarg>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pimc_gpu.h"

struct pimc_state {
    int M = 0, N = 0;            // array extents: slices x columns
    int num_worldlines = 0;      // first line of the file (getNumParticles at save time)
    std::vector<int64_t> counters;   // pairs from the acceptance / estimator lines, file order
    std::vector<double> beads;       // [M][N][3]
    std::vector<int32_t> next, prev; // [M][N][2], {-1,-1} = XXX
    std::vector<uint32_t> worm;      // [M][N]
    std::vector<uint64_t> random;    // 625 words, or empty
};

namespace {

thread_local std::string g_state_err;

int state_fail(int code, const std::string& msg) {
    g_state_err = msg;
    return code;
}

// A cursor over NUL-terminated text with the stream extractors the reference's
// operator>> overloads use: whitespace is skipped before every token.
struct Cursor {
    const char* p;
    const char* end;
    void skip_ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    }
    bool ch(char c) {
        skip_ws();
        if (p < end && *p == c) {
            ++p;
            return true;
        }
        return false;
    }
    bool i64(int64_t* v) {
        skip_ws();
        if (p >= end) return false;
        char* e = nullptr;
        errno = 0;
        const long long x = std::strtoll(p, &e, 10);
        if (e == p || errno) return false;
        p = e;
        *v = x;
        return true;
    }
    bool u64(uint64_t* v) {
        skip_ws();
        if (p >= end || *p == '-') return false;
        char* e = nullptr;
        errno = 0;
        const unsigned long long x = std::strtoull(p, &e, 10);
        if (e == p || errno) return false;
        p = e;
        *v = x;
        return true;
    }
    bool f64(double* v) {
        skip_ws();
        if (p >= end) return false;
        char* e = nullptr;
        const double x = std::strtod(p, &e);
        if (e == p) return false;
        p = e;
        *v = x;
        return true;
    }
    bool at_end() {
        skip_ws();
        return p >= end;
    }
};

// "(r0,r1) x (c0,c1)" then everything up to '[' (include/common.h:307-366).
bool array_header(Cursor& c, int* rows, int* cols) {
    int64_t r0, r1, c0, c1;
    if (!c.ch('(') || !c.i64(&r0) || !c.ch(',') || !c.i64(&r1) || !c.ch(')')) return false;
    if (!c.ch('x')) return false;
    if (!c.ch('(') || !c.i64(&c0) || !c.ch(',') || !c.i64(&c1) || !c.ch(')')) return false;
    if (r1 < r0 || c1 < c0 || r1 - r0 + 1 > (1 << 24) || c1 - c0 + 1 > (1 << 24)) return false;
    *rows = (int)(r1 - r0 + 1);
    *cols = (int)(c1 - c0 + 1);
    c.skip_ws();
    while (c.p < c.end && *c.p != '[') ++c.p;
    if (c.p >= c.end) return false;
    ++c.p;
    return true;
}

bool array_close(Cursor& c) {
    c.skip_ws();
    while (c.p < c.end && *c.p != ']') ++c.p;
    if (c.p >= c.end) return false;
    ++c.p;
    return true;
}

int parse_text(const char* text, size_t len, pimc_state* st) {
    Cursor c{text, text + len};
    int64_t nw;
    if (!c.i64(&nw) || nw < 0) return state_fail(PIMC_E_ARG, "state file: no worldline count on the first line");
    st->num_worldlines = (int)nw;
    // lines up to the first one that starts with '(' (src/pimc.cpp:1136-1144)
    while (c.p < c.end && *c.p != '\n') ++c.p;
    if (c.p < c.end) ++c.p;
    while (c.p < c.end && *c.p != '(') {
        const char* eol = c.p;
        while (eol < c.end && *eol != '\n') ++eol;
        Cursor line{c.p, eol};
        int64_t a, b;
        if (line.i64(&a) && line.i64(&b) && line.at_end()) {
            st->counters.push_back(a);
            st->counters.push_back(b);
        }
        c.p = eol < c.end ? eol + 1 : eol;
    }
    int M, N;
    if (!array_header(c, &M, &N)) return state_fail(PIMC_E_ARG, "state file: bad beads array header");
    st->M = M;
    st->N = N;
    const size_t MN = (size_t)M * N;
    st->beads.resize(3 * MN);
    for (size_t t = 0; t < MN; ++t) {
        double* b = &st->beads[3 * t];
        if (!c.ch('(') || !c.f64(&b[0]) || !c.ch(',') || !c.f64(&b[1]) || !c.ch(',') || !c.f64(&b[2]) || !c.ch(')'))
            return state_fail(PIMC_E_ARG, "state file: bad bead position " + std::to_string(t));
    }
    if (!array_close(c)) return state_fail(PIMC_E_ARG, "state file: beads array not closed");
    for (int which = 0; which < 2; ++which) {
        std::vector<int32_t>& L = which ? st->prev : st->next;
        int m2, n2;
        if (!array_header(c, &m2, &n2) || m2 != M || n2 != N)
            return state_fail(PIMC_E_ARG, which ? "state file: bad prevLink array" : "state file: bad nextLink array");
        L.resize(2 * MN);
        for (size_t t = 0; t < MN; ++t) {
            int64_t a, b;
            if (!c.ch('(') || !c.i64(&a) || !c.ch(',') || !c.i64(&b) || !c.ch(')'))
                return state_fail(PIMC_E_ARG, "state file: bad link " + std::to_string(t));
            L[2 * t] = (int32_t)a;
            L[2 * t + 1] = (int32_t)b;
        }
        if (!array_close(c)) return state_fail(PIMC_E_ARG, "state file: link array not closed");
    }
    int m3, n3;
    if (!array_header(c, &m3, &n3) || m3 != M || n3 != N) return state_fail(PIMC_E_ARG, "state file: bad worm array");
    st->worm.resize(MN);
    for (size_t t = 0; t < MN; ++t) {
        uint64_t w;
        if (!c.u64(&w)) return state_fail(PIMC_E_ARG, "state file: bad worm entry " + std::to_string(t));
        st->worm[t] = (uint32_t)w;
    }
    if (!array_close(c)) return state_fail(PIMC_E_ARG, "state file: worm array not closed");
    // MTRand::save words (include/MersenneTwister.h:370-377), when present
    std::vector<uint64_t> words;
    uint64_t w;
    while (words.size() < PIMC_MT_SAVE && c.u64(&w)) words.push_back(w);
    if (!words.empty() && words.size() != PIMC_MT_SAVE)
        return state_fail(PIMC_E_ARG, "state file: random state has " + std::to_string(words.size()) + " words");
    st->random = words;
    // links must point inside the arrays or be XXX
    for (size_t t = 0; t < 2 * MN; t += 2)
        for (const std::vector<int32_t>* L : {&st->next, &st->prev}) {
            const int a = (*L)[t], b = (*L)[t + 1];
            if (!(a == -1 && b == -1) && (a < 0 || a >= M || b < 0 || b >= N))
                return state_fail(PIMC_E_ARG, "state file: link out of range at bead " + std::to_string(t / 2));
        }
    return PIMC_OK;
}

void put(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void put(std::string& s, const char* fmt, ...) {
    char buf[160];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    s.append(buf, (size_t)(n < (int)sizeof(buf) ? n : (int)sizeof(buf) - 1));
}

// operator<<(ostream, DynamicArray<T,2>) (include/common.h:259-272) followed by
// the std::endl saveState adds.
template <typename F>
void put_array(std::string& s, int M, int N, F elem) {
    put(s, "(0,%d) x (0,%d)\n[ ", M - 1, N - 1);
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j < N; ++j) {
            elem(s, (size_t)i * N + j);
            s += ' ';
        }
        if (i < M - 1) s += "\n  ";
    }
    s += "]\n\n";
}

std::string format_text(const pimc_state* st) {
    std::string s;
    s.reserve((size_t)st->M * st->N * 80 + 8192);
    put(s, "%d\n", st->num_worldlines);
    for (size_t k = 0; k + 1 < st->counters.size(); k += 2)
        put(s, "%16" PRId64 "\t%16" PRId64 "\n", st->counters[k], st->counters[k + 1]);
    // std::setprecision(16) with the default float field: %.16g
    put_array(s, st->M, st->N, [&](std::string& o, size_t t) {
        put(o, "(%.16g,%.16g,%.16g)", st->beads[3 * t], st->beads[3 * t + 1], st->beads[3 * t + 2]);
    });
    put_array(s, st->M, st->N, [&](std::string& o, size_t t) { put(o, "(%d,%d)", st->next[2 * t], st->next[2 * t + 1]); });
    put_array(s, st->M, st->N, [&](std::string& o, size_t t) { put(o, "(%d,%d)", st->prev[2 * t], st->prev[2 * t + 1]); });
    put_array(s, st->M, st->N, [&](std::string& o, size_t t) { put(o, "%u", st->worm[t]); });
    for (uint64_t w : st->random) put(s, "%" PRIu64 " ", w);
    s += '\n';
    return s;
}

}  // namespace

extern "C" {

const char* pimc_state_last_error(void) { return g_state_err.c_str(); }

int pimc_state_parse(const char* text, size_t len, pimc_state** out) {
    if (!text || !out) return state_fail(PIMC_E_ARG, "null argument");
    *out = nullptr;
    pimc_state* st = new pimc_state();
    const int r = parse_text(text, len, st);
    if (r) {
        delete st;
        return r;
    }
    *out = st;
    return PIMC_OK;
}

int pimc_state_read(const char* path, pimc_state** out) {
    if (!path || !out) return state_fail(PIMC_E_ARG, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return state_fail(PIMC_E_ARG, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    return pimc_state_parse(text.data(), text.size(), out);
}

int pimc_state_create(int num_slices, int num_alloc, pimc_state** out) {
    if (!out || num_slices < 1 || num_alloc < 1) return state_fail(PIMC_E_ARG, "bad state dimensions");
    pimc_state* st = new pimc_state();
    st->M = num_slices;
    st->N = num_alloc;
    st->num_worldlines = num_alloc;
    const size_t MN = (size_t)num_slices * num_alloc;
    st->beads.assign(3 * MN, 0.0);
    st->next.assign(2 * MN, -1);
    st->prev.assign(2 * MN, -1);
    st->worm.assign(MN, 0u);
    *out = st;
    return PIMC_OK;
}

void pimc_state_destroy(pimc_state* st) { delete st; }

int pimc_state_dims(const pimc_state* st, int* num_slices, int* num_alloc, int* num_worldlines) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    if (num_slices) *num_slices = st->M;
    if (num_alloc) *num_alloc = st->N;
    if (num_worldlines) *num_worldlines = st->num_worldlines;
    return PIMC_OK;
}

int pimc_state_get(const pimc_state* st, double* beads, int32_t* next, int32_t* prev, uint32_t* worm) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    const size_t MN = (size_t)st->M * st->N;
    if (beads) std::memcpy(beads, st->beads.data(), sizeof(double) * 3 * MN);
    if (next) std::memcpy(next, st->next.data(), sizeof(int32_t) * 2 * MN);
    if (prev) std::memcpy(prev, st->prev.data(), sizeof(int32_t) * 2 * MN);
    if (worm) std::memcpy(worm, st->worm.data(), sizeof(uint32_t) * MN);
    return PIMC_OK;
}

int pimc_state_set(pimc_state* st, const double* beads, const int32_t* next, const int32_t* prev,
                   const uint32_t* worm) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    const size_t MN = (size_t)st->M * st->N;
    for (const int32_t* L : {next, prev})
        if (L)
            for (size_t t = 0; t < MN; ++t) {
                const int a = L[2 * t], b = L[2 * t + 1];
                if (!(a == -1 && b == -1) && (a < 0 || a >= st->M || b < 0 || b >= st->N))
                    return state_fail(PIMC_E_ARG, "link out of range at bead " + std::to_string(t));
            }
    if (beads) std::memcpy(st->beads.data(), beads, sizeof(double) * 3 * MN);
    if (next) std::memcpy(st->next.data(), next, sizeof(int32_t) * 2 * MN);
    if (prev) std::memcpy(st->prev.data(), prev, sizeof(int32_t) * 2 * MN);
    if (worm) std::memcpy(st->worm.data(), worm, sizeof(uint32_t) * MN);
    return PIMC_OK;
}

int pimc_state_counters(pimc_state* st, int64_t* pairs, int cap, int* count, const int64_t* set_pairs, int set_count) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    if (count) *count = (int)(st->counters.size() / 2);
    if (pairs)
        for (int k = 0; k < cap && 2 * (size_t)k + 1 < st->counters.size(); ++k) {
            pairs[2 * k] = st->counters[2 * k];
            pairs[2 * k + 1] = st->counters[2 * k + 1];
        }
    if (set_pairs && set_count >= 0) st->counters.assign(set_pairs, set_pairs + 2 * (size_t)set_count);
    return PIMC_OK;
}

int pimc_state_random(pimc_state* st, uint64_t* words, int* have, const uint64_t* set_words) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    if (have) *have = st->random.empty() ? 0 : 1;
    if (words && !st->random.empty()) std::memcpy(words, st->random.data(), sizeof(uint64_t) * PIMC_MT_SAVE);
    if (set_words) st->random.assign(set_words, set_words + PIMC_MT_SAVE);
    return PIMC_OK;
}

int pimc_state_left_pack(pimc_state* st) {
    // Path::leftPack (src/path.cpp:145-189)
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    const int M = st->M, N = st->N;
    auto idx = [N](int m, int i) { return (size_t)m * N + i; };
    for (int m = 0; m < M; ++m)
        for (int i = 0; i < N; ++i) {
            if (st->worm[idx(m, i)]) continue;
            int j = i + 1;
            while (j < N && !st->worm[idx(m, j)]) ++j;
            if (j == N) continue;
            const size_t b1 = idx(m, i), b2 = idx(m, j);
            for (int d = 0; d < 3; ++d) st->beads[3 * b1 + d] = st->beads[3 * b2 + d];
            st->worm[b1] = 1;
            st->worm[b2] = 0;
            const int32_t n0 = st->next[2 * b2], n1 = st->next[2 * b2 + 1];
            const int32_t p0 = st->prev[2 * b2], p1 = st->prev[2 * b2 + 1];
            st->next[2 * b1] = n0;
            st->next[2 * b1 + 1] = n1;
            st->prev[2 * b1] = p0;
            st->prev[2 * b1 + 1] = p1;
            // the neighbours now link to the new place (a worm end has no
            // neighbour on that side: its link is XXX and stays untouched)
            if (!(p0 == -1 && p1 == -1)) {
                st->next[2 * idx(p0, p1)] = m;
                st->next[2 * idx(p0, p1) + 1] = i;
            }
            if (!(n0 == -1 && n1 == -1)) {
                st->prev[2 * idx(n0, n1)] = m;
                st->prev[2 * idx(n0, n1) + 1] = i;
            }
            st->next[2 * b2] = st->next[2 * b2 + 1] = -1;
            st->prev[2 * b2] = st->prev[2 * b2 + 1] = -1;
        }
    return PIMC_OK;
}

int pimc_state_format(const pimc_state* st, char* buf, size_t cap, size_t* len) {
    if (!st) return state_fail(PIMC_E_ARG, "null state");
    const std::string s = format_text(st);
    if (len) *len = s.size();
    if (buf) {
        if (cap < s.size() + 1) return state_fail(PIMC_E_ARG, "buffer too small for the state text");
        std::memcpy(buf, s.data(), s.size());
        buf[s.size()] = '\0';
    }
    return PIMC_OK;
}

int pimc_state_write(const pimc_state* st, const char* path) {
    if (!st || !path) return state_fail(PIMC_E_ARG, "null argument");
    const std::string s = format_text(st);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return state_fail(PIMC_E_ARG, std::string("cannot write ") + path);
    const size_t w = std::fwrite(s.data(), 1, s.size(), f);
    const int e = std::fclose(f);
    if (w != s.size() || e != 0) return state_fail(PIMC_E_ARG, std::string("short write to ") + path);
    return PIMC_OK;
}

int pimc_state_to_path(const pimc_state* st, int num_slices, double side, double* beads, uint8_t* bead_on,
                       int32_t* num_beads, int32_t* next, int32_t* prev) {
    // loadState after the arrays are read (src/pimc.cpp:1150-1253)
    if (!st || !beads || !bead_on || !num_beads || !next || !prev) return state_fail(PIMC_E_ARG, "null argument");
    if (!(side > 0.0)) return state_fail(PIMC_E_ARG, "box side must be positive");
    const int M = num_slices, NW = st->num_worldlines;
    if (NW < 1) return state_fail(PIMC_E_ARG, "state holds no worldlines");
    if (st->M == M) {
        if (st->N != NW)
            return state_fail(PIMC_E_ARG, "state arrays have " + std::to_string(st->N) + " columns, the file says " +
                                              std::to_string(NW) + " worldlines");
        std::memcpy(beads, st->beads.data(), sizeof(double) * 3 * (size_t)M * NW);
        std::memcpy(next, st->next.data(), sizeof(int32_t) * 2 * (size_t)M * NW);
        std::memcpy(prev, st->prev.data(), sizeof(int32_t) * 2 * (size_t)M * NW);
        for (size_t t = 0; t < (size_t)M * NW; ++t) bead_on[t] = st->worm[t] != 0;
    } else {
        // loadClassicalState (src/pimc.cpp:1006-1027): every worldline active on
        // slice 0 of the file, at its slice-0 position on all M slices
        if (st->N < NW) return state_fail(PIMC_E_ARG, "state arrays narrower than the worldline count");
        std::memset(beads, 0, sizeof(double) * 3 * (size_t)M * NW);
        std::memset(bead_on, 0, (size_t)M * NW);
        int ptcl = 0;
        for (int n = 0; n < NW; ++n) {
            if (!st->worm[n]) continue;
            for (int m = 0; m < M; ++m) {
                const size_t t = (size_t)m * NW + ptcl;
                for (int d = 0; d < 3; ++d) beads[3 * t + d] = st->beads[3 * (size_t)n + d];
                bead_on[t] = 1;
            }
            ++ptcl;
        }
        for (int m = 0; m < M; ++m)
            for (int i = 0; i < NW; ++i) {
                const size_t t = (size_t)m * NW + i;
                if (bead_on[t]) {
                    next[2 * t] = (m + 1 == M) ? 0 : m + 1;
                    next[2 * t + 1] = i;
                    prev[2 * t] = (m == 0) ? M - 1 : m - 1;
                    prev[2 * t + 1] = i;
                } else {
                    next[2 * t] = next[2 * t + 1] = prev[2 * t] = prev[2 * t + 1] = -1;
                }
            }
    }
    // putInside (a fully periodic cube: putInBC, include/container.h:50-59,
    // :118-132) and numBeadsAtSlice
    const double inv = 1.0 / side;
    for (int m = 0; m < M; ++m) {
        int on = 0;
        for (int i = 0; i < NW; ++i) {
            const size_t t = (size_t)m * NW + i;
            for (int d = 0; d < 3; ++d) {
#pragma clang fp contract(off)
                double& x = beads[3 * t + d];
                x = x - side * std::floor(x * inv + 0.5);
            }
            on += bead_on[t] ? 1 : 0;
        }
        num_beads[m] = on;
    }
    return PIMC_OK;
}

}  // extern "C"
