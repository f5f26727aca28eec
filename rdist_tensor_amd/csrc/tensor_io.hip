// FROSTT .tns text ingest (host code).  A .tns file holds one nonzero per
// line: N 1-based coordinates, then the value, separated by blanks; blank
// lines and lines starting with '#' are skipped.  The reference converts
// such a file to its HDF5 layout (tensor_io/process_frostt_tensor.cpp:19-172:
// N = tokens of the first line - 1, MODE_i / VALUES columns, per-mode minimum
// and maximum) and its reader shifts every coordinate by the mode's minimum
// (sparse_tensor_e.py:68-103).  Here the file is streamed in 16 MiB blocks
// and parsed in place; the caller receives the raw coordinates with the
// per-mode minimum and maximum, as the HDF5 file stores them.
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>

#include "common.hpp"

namespace exafac {
namespace {

constexpr size_t kBlock = (size_t)16 << 20;

struct LineReader {
    FILE* f = nullptr;
    std::unique_ptr<char[]> buf;
    size_t len = 0, pos = 0;
    bool eof = false;
    int64_t lineno = 0;

    explicit LineReader(const char* path) : buf(new char[2 * kBlock + 2]) {
        f = std::fopen(path, "rb");
        EX_CHECK(f, 1, std::string("cannot open ") + path + ": " + std::strerror(errno));
    }
    ~LineReader() {
        if (f) std::fclose(f);
    }
    // Next line as [b, e) (no newline); false at the end of the file.
    bool next(const char*& b, const char*& e) {
        while (true) {
            char* nl = static_cast<char*>(std::memchr(buf.get() + pos, '\n', len - pos));
            if (nl) {
                b = buf.get() + pos;
                e = nl;
                pos = (size_t)(nl - buf.get()) + 1;
                lineno++;
                return true;
            }
            if (eof) {
                if (pos < len) {   // last line without a newline
                    b = buf.get() + pos;
                    e = buf.get() + len;
                    pos = len;
                    lineno++;
                    return true;
                }
                return false;
            }
            // keep the partial line, refill behind it
            const size_t rest = len - pos;
            EX_CHECK(rest <= kBlock, 1, "line " + std::to_string(lineno + 1) + " longer than 16 MiB");
            std::memmove(buf.get(), buf.get() + pos, rest);
            const size_t got = std::fread(buf.get() + rest, 1, kBlock, f);
            EX_CHECK(got == kBlock || !std::ferror(f), 1, "read error");
            if (got < kBlock) eof = true;
            len = rest + got;
            pos = 0;
        }
    }
};

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// A data line's tokens: true for a data line, false for a blank or comment line.
bool split(const char* b, const char* e, std::vector<std::pair<const char*, const char*>>& tok) {
    tok.clear();
    const char* p = b;
    while (p < e && blank(*p)) p++;
    if (p == e || *p == '#') return false;
    while (p < e) {
        const char* s = p;
        while (p < e && !blank(*p)) p++;
        tok.emplace_back(s, p);
        while (p < e && blank(*p)) p++;
    }
    return true;
}

uint64_t parse_index(const char* s, const char* e, int64_t lineno) {
    EX_CHECK(s < e, 1, "empty coordinate on line " + std::to_string(lineno));
    uint64_t v = 0;
    for (const char* p = s; p < e; p++) {
        EX_CHECK(*p >= '0' && *p <= '9', 1, "coordinate is not a non-negative integer on line " + std::to_string(lineno));
        const uint64_t d = (uint64_t)(*p - '0');
        EX_CHECK(v <= (UINT64_MAX - d) / 10, 1, "coordinate overflows 64 bits on line " + std::to_string(lineno));
        v = v * 10 + d;
    }
    return v;
}

// Locale-independent (std::from_chars); a leading '+' is accepted as well.
double parse_value(const char* s, const char* e, int64_t lineno) {
    if (s < e && *s == '+') s++;
    double v = 0.0;
    const auto r = std::from_chars(s, e, v);
    EX_CHECK(s < e && r.ec == std::errc() && r.ptr == e, 1, "malformed value on line " + std::to_string(lineno));
    return v;
}

}  // namespace

void tns_scan(const char* path, int64_t& nnz, int& N) {
    LineReader r(path);
    std::vector<std::pair<const char*, const char*>> tok;
    const char *b, *e;
    nnz = 0;
    N = 0;
    while (r.next(b, e)) {
        if (!split(b, e, tok)) continue;
        if (nnz == 0) {
            EX_CHECK(tok.size() >= 2, 1, "line " + std::to_string(r.lineno) + ": need coordinates and a value");
            N = (int)tok.size() - 1;
            EX_CHECK(N <= kMaxModes, 1, "more than " + std::to_string(kMaxModes) + " modes");
        } else {
            EX_CHECK((int)tok.size() == N + 1, 1,
                     "line " + std::to_string(r.lineno) + " has " + std::to_string(tok.size()) + " fields, expected " +
                         std::to_string(N + 1));
        }
        nnz++;
    }
}

void tns_read(const char* path, int64_t nnz, int N, uint64_t* coords, double* vals, uint64_t* mins,
              uint64_t* maxs) {
    LineReader r(path);
    std::vector<std::pair<const char*, const char*>> tok;
    const char *b, *e;
    int64_t i = 0;
    for (int k = 0; k < N; k++) {
        mins[k] = UINT64_MAX;
        maxs[k] = 0;
    }
    while (r.next(b, e)) {
        if (!split(b, e, tok)) continue;
        EX_CHECK(i < nnz, 1, "file holds more nonzeros than scanned");
        EX_CHECK((int)tok.size() == N + 1, 1,
                 "line " + std::to_string(r.lineno) + " has " + std::to_string(tok.size()) + " fields, expected " +
                     std::to_string(N + 1));
        for (int k = 0; k < N; k++) {
            const uint64_t v = parse_index(tok[k].first, tok[k].second, r.lineno);
            coords[i * N + k] = v;
            if (v < mins[k]) mins[k] = v;
            if (v > maxs[k]) maxs[k] = v;
        }
        vals[i] = parse_value(tok[N].first, tok[N].second, r.lineno);
        i++;
    }
    EX_CHECK(i == nnz, 1, "file holds fewer nonzeros than scanned");
    if (nnz == 0)
        for (int k = 0; k < N; k++) mins[k] = 0;
}

}  // namespace exafac
