// libkiwi_io: KIWI's SPH snapshot readers (src/snapshot.cpp:17-522) as a native host library.  The
// file is read once into memory, cut into lines with getline's rules, and the lines are parsed by a
// pool of threads, each into its own column runs, which are joined in file order.  Every line is
// parsed with the same library routine the reference uses for it (strtod for plain_ascii, a string
// stream's double extraction for splash_ascii), so the values, the rows kept, the first error and its
// message are the reference's.
//
//   kiwi_io_detect_format   phantom_binary (a Phantom capture record first), splash_ascii (a first
//                           line "#...SPLASH..."), else plain_ascii (detect_snapshot_format, :495-509)
//   kiwi_io_load_snapshot   plain_ascii (load_plain_ascii, :32-84), splash_ascii (load_splash_ascii,
//                           :101-491) and phantom_binary (load_phantom_dump, src/phantom_dump.cpp:156-317),
//                           "auto" resolved as above
#include "kiwi_io.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;
int g_threads = 0;  // 0: the hardware's count, at most 16

int worker_count() {
    int n = g_threads > 0 ? g_threads : static_cast<int>(std::thread::hardware_concurrency());
    return std::max(1, std::min(n, 16));
}

// The whole file plus a terminating '\0'.  Lines follow getline: '\n' ends a line (each '\n' is turned
// into '\0' by whoever parses that line, so a line is a C string for strtod); a last line without '\n'
// counts when it is not empty.
bool read_file(const std::string& path, std::vector<char>& buf) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    buf.clear();
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long size = std::ftell(f);
        if (size > 0) buf.reserve(static_cast<size_t>(size) + 1);
        std::rewind(f);
    }
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    std::fclose(f);
    buf.push_back('\0');
    return true;
}

// The next line of buf from byte `at` (below `end`, which is a line boundary or the data's end):
// returns false at the end; otherwise its start and length, and `at` moves past it.
inline bool next_line(char* buf, size_t& at, size_t end, char*& line, size_t& len) {
    if (at >= end) return false;
    line = buf + at;
    char* nl = static_cast<char*>(std::memchr(line, '\n', end - at));
    if (nl) {
        *nl = '\0';
        len = static_cast<size_t>(nl - line);
        at += len + 1;
    } else {
        len = end - at;
        at = end;
    }
    return true;
}

bool blank(const char* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (p[i] != ' ' && p[i] != '\t' && p[i] != '\r') return false;
    return true;
}

bool comment(const char* p, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        if (p[i] == ' ' || p[i] == '\t' || p[i] == '\r') continue;
        return p[i] == '#';
    }
    return false;
}

struct Columns {
    std::vector<double> c[9];  // x y z vx vy vz m h rho
};

// The first error of a chunk: its line within the chunk (0-based) and what the message needs.
struct ChunkError {
    size_t line = SIZE_MAX;
    int kind = 0, column = 0;
    std::string msg;
};

// Splits bytes [first, last) of buf into contiguous runs of whole lines, one per worker (at most 16, in
// file order), and runs parse(chunk, a, b, cols, err, nlines) on each; joins the columns in order.  On
// failure returns false with the earliest chunk's error, its line made global (lines before the chunk
// + its line within it + line0) through report(err, global_line).
template <class F, class R>
bool parallel_rows(std::vector<char>& buf, size_t first, size_t last, F parse, R report, Columns& out) {
    const size_t n = last > first ? last - first : 0;
    const int T = static_cast<int>(std::min<size_t>(worker_count(), std::max<size_t>(1, n / (1 << 20))));
    std::vector<size_t> cut(T + 1);
    cut[0] = first;
    cut[T] = last;
    for (int t = 1; t < T; ++t) {
        size_t c = std::max(first + n * t / T, cut[t - 1]);
        const void* nl = c < last ? std::memchr(buf.data() + c, '\n', last - c) : nullptr;
        cut[t] = nl ? static_cast<size_t>(static_cast<const char*>(nl) - buf.data()) + 1 : last;
    }
    std::vector<Columns> part(T);
    std::vector<ChunkError> perr(T);
    std::vector<size_t> nlines(T, 0);
    auto work = [&](int t) {
        for (auto& c : part[t].c) c.reserve((cut[t + 1] - cut[t]) / 96 + 16);  // rows are rarely shorter
        parse(t, cut[t], cut[t + 1], part[t], perr[t], nlines[t]);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    size_t before = 0;
    for (int t = 0; t < T; ++t) {
        if (perr[t].line != SIZE_MAX) {
            report(perr[t], before + perr[t].line);
            return false;
        }
        before += nlines[t];
    }
    for (int k = 0; k < 9; ++k) {
        size_t total = 0;
        for (int t = 0; t < T; ++t) total += part[t].c[k].size();
        out.c[k].reserve(total);
        for (int t = 0; t < T; ++t) out.c[k].insert(out.c[k].end(), part[t].c[k].begin(), part[t].c[k].end());
    }
    return true;
}

// ---------------------------------------------------------------- plain_ascii (:32-84)
bool load_plain(const std::string& path, std::vector<char>& buf, Columns& out, std::string& err) {
    auto parse = [&](int, size_t a, size_t b, Columns& cols, ChunkError& e, size_t& nlines) {
        char* line;
        size_t len;
        for (size_t at = a; next_line(buf.data(), at, b, line, len); ++nlines) {
            if (blank(line, len)) continue;
            double v[9];
            const char* p = line;
            for (int k = 0; k < 9; ++k) {
                char* end = nullptr;
                v[k] = std::strtod(p, &end);
                if (end == p || !std::isfinite(v[k])) {
                    e.line = nlines;
                    e.kind = end == p ? 0 : 1;
                    e.column = k + 1;
                    return;
                }
                p = end;
            }
            while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
            if (*p != '\0') {
                e.line = nlines;
                e.kind = 2;
                return;
            }
            for (int k = 0; k < 9; ++k) cols.c[k].push_back(v[k]);
        }
    };
    auto report = [&](const ChunkError& e, size_t line) {
        const std::string at = "SPH snapshot file (plain_ascii format), line " + std::to_string(line + 1) + ": ";
        if (e.kind == 0)
            err = at + "expected 9 numeric columns (x y z vx vy vz m h rho), could not read column " +
                  std::to_string(e.column) + ": " + path;
        else if (e.kind == 1)
            err = at + "column " + std::to_string(e.column) + " is not finite: " + path;
        else
            err = at + "more than 9 columns (expected x y z vx vy vz m h rho): " + path;
    };
    if (!parallel_rows(buf, 0, buf.size() - 1, parse, report, out)) return false;
    if (out.c[0].empty()) {
        err = "SPH snapshot file (plain_ascii format) contained no rows -- expected 9 columns "
              "(x y z vx vy vz m h rho) per line: " + path;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- splash_ascii (:101-491)
int whitespace_tokens(const std::string& s) {
    std::istringstream in(s);
    std::string tok;
    int n = 0;
    while (in >> tok) ++n;
    return n;
}

// the label with any [unit] removed, without \ _ space tab, lower case
std::string label_key(const std::string& raw) {
    std::string k;
    int depth = 0;
    for (char c : raw) {
        if (c == '[') { depth = 1; continue; }
        if (c == ']') { depth = 0; continue; }
        if (depth || c == '\\' || c == '_' || c == ' ' || c == '\t') continue;
        k += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    }
    return k;
}

// the text inside the last [...] of a label, or ""
std::string label_unit(const std::string& raw) {
    const size_t close = raw.rfind(']');
    if (close == std::string::npos) return "";
    const size_t open = raw.rfind('[', close);
    return open == std::string::npos ? "" : raw.substr(open + 1, close - open - 1);
}

std::string unit_key(const std::string& raw) {
    std::string k;
    for (char c : raw)
        if (c != '{' && c != '}' && c != '_' && c != ' ' && c != '\t')
            k += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return k;
}

// g/cm^3 in one Msun/pc^3, from the same two constants and in the same order as the reference
double gcm3_per_msun_pc3() {
    const double msun_g = 1.9891e33, pc_cm = 3.08567758e18;
    return msun_g / (pc_cm * pc_cm * pc_cm);
}

const char* column_role(const std::string& key) {
    static const std::unordered_map<std::string, const char*> roles = {
        {"x", "x"}, {"y", "y"}, {"z", "z"}, {"particlemass", "m"}, {"mass", "m"}, {"m", "m"}, {"h", "h"},
        {"smoothinglength", "h"}, {"density", "rho"}, {"rho", "rho"}, {"vx", "vx"}, {"vy", "vy"}, {"vz", "vz"},
        {"itype", "itype"}};
    auto it = roles.find(key);
    return it == roles.end() ? nullptr : it->second;
}

// The header line cut into one label per data column: fixed fields of 24 or 16 characters, and a bare
// "itype" at the end when present (split_splash_ascii_labels, :243-311).
bool header_labels(const std::string& header, int ncols, const std::string& path, std::vector<std::string>& labels,
                   std::string& err) {
    if (header.size() < 2 || header[0] != '#') {
        err = "splash-ascii snapshot: expected the line immediately before the particle data to be the "
              "'#'-prefixed column-header line, but it wasn't: " + path;
        return false;
    }
    std::string body = header.substr(1);
    while (!body.empty() && (body.back() == '\r' || body.back() == '\n')) body.pop_back();
    if (!body.empty() && body.front() == ' ') body.erase(0, 1);
    const bool itype_tail = body.size() >= 5 && body.compare(body.size() - 5, 5, "itype") == 0;
    const int nfixed = itype_tail ? ncols - 1 : ncols;
    if (nfixed <= 0) {
        err = "splash-ascii snapshot: column-header line doesn't have enough fields for the " + std::to_string(ncols) +
              " column(s) seen in the particle data: " + path;
        return false;
    }
    const std::string fixed = itype_tail ? body.substr(0, body.size() - 5) : body;
    int w = -1;
    for (int cand : {24, 16}) {
        const size_t full = static_cast<size_t>(nfixed) * cand;
        const bool fits = itype_tail ? fixed.size() == full
                                     : fixed.size() >= full - cand + 1 && fixed.size() <= full;
        if (fits) {
            w = cand;
            break;
        }
    }
    if (w < 0) {
        err = "splash-ascii snapshot: could not reliably split the column-header line into " + std::to_string(nfixed) +
              " field(s) -- its length doesn't match either of SPLASH's known label field widths (23 or 15 "
              "characters): " + path;
        return false;
    }
    labels.clear();
    for (int i = 0; i < nfixed; ++i) {
        const size_t at = static_cast<size_t>(i) * w;
        const std::string f = (!itype_tail && i == nfixed - 1) ? fixed.substr(at) : fixed.substr(at, w);
        const size_t a = f.find_first_not_of(" \t"), b = f.find_last_not_of(" \t");
        labels.push_back(a == std::string::npos ? std::string() : f.substr(a, b - a + 1));
    }
    if (itype_tail) labels.push_back("itype");
    return true;
}

bool load_splash(const std::string& path, double hfact, std::vector<char>& buf, Columns& out, std::string& err) {
    // the data block: the first uncommented, non-blank line followed by another (:101-133); the header
    // lines are scanned in place, nothing is written to the buffer before the rows are parsed
    const size_t size = buf.size() - 1;
    struct Span {
        size_t at, len;
    };
    auto line_at = [&](size_t at, Span& sp) {
        if (at >= size) return false;
        const void* nl = std::memchr(buf.data() + at, '\n', size - at);
        sp = {at, nl ? static_cast<size_t>(static_cast<const char*>(nl) - buf.data()) - at : size - at};
        return true;
    };
    auto after = [&](const Span& sp) { return sp.at + sp.len + 1; };
    auto meta = [&](const Span& sp) { return comment(buf.data() + sp.at, sp.len) || blank(buf.data() + sp.at, sp.len); };
    auto text = [&](const Span& sp) { return std::string(buf.data() + sp.at, sp.len); };
    Span cur, nxt, prev{0, 0};
    if (!line_at(0, cur)) {
        err = "splash-ascii snapshot file is empty: " + path;
        return false;
    }
    const std::string eof_msg = "splash-ascii snapshot: reached end of file while still in the header preamble, "
                                "no particle data found: " + path;
    size_t data = 0;  // the data block's first line number
    for (;;) {
        if (meta(cur)) {
            prev = cur;
            ++data;
            if (!line_at(after(cur), cur)) {
                err = eof_msg;
                return false;
            }
            continue;
        }
        const bool have_next = line_at(after(cur), nxt);
        if (have_next && !meta(nxt)) break;
        ++data;
        if (!have_next) {
            err = eof_msg;
            return false;
        }
        prev = cur;
        cur = nxt;
    }
    if (data == 0) {
        err = "splash-ascii snapshot: no header preamble found before the particle data -- expected at least the "
              "column-header comment line: " + path;
        return false;
    }
    const int ncols = whitespace_tokens(text(cur));
    if (ncols < 8) {
        err = "splash-ascii snapshot: first particle-data row has only " + std::to_string(ncols) +
              " column(s), too few to be a real splash_ascii file: " + path;
        return false;
    }
    std::vector<std::string> labels;
    if (!header_labels(text(prev), ncols, path, labels, err)) return false;
    std::unordered_map<std::string, int> at;
    for (int i = 0; i < static_cast<int>(labels.size()); ++i) {
        const char* role = column_role(label_key(labels[i]));
        if (!role) continue;
        auto ins = at.emplace(role, i);
        if (!ins.second) {
            err = "splash-ascii snapshot: the column-header line appears to name '" + std::string(role) +
                  "' more than once (columns " + std::to_string(ins.first->second) + " and " + std::to_string(i) +
                  ") -- the header likely failed to parse correctly: " + path;
            return false;
        }
    }
    for (const char* req : {"x", "y", "z", "m", "vx", "vy", "vz"})
        if (!at.count(req)) {
            err = "splash-ascii snapshot: could not find a '" + std::string(req) + "' column in the column-header line: " +
                  path;
            return false;
        }
    const bool has_h = at.count("h") != 0, has_rho = at.count("rho") != 0, has_itype = at.count("itype") != 0;
    if (!has_h && !has_rho) {
        err = "splash-ascii snapshot: could not find either an 'h' (smoothing length) or a 'density' column in the "
              "column-header line -- need at least one, to derive the other from it: " + path;
        return false;
    }
    if (!has_itype)
        std::cerr << "[kiwi] splash-ascii snapshot: no 'itype' column in the header -- this file has only one particle "
                     "type (SPLASH omits itype entirely in that case); treating every row as gas: "
                  << path << "\n";
    if (has_h != has_rho) {
        const std::string which = has_h ? "h" : "density";
        const std::string mu = label_unit(labels[at.at("m")]);
        const std::string pu = label_unit(labels[at.at(has_h ? "h" : "rho")]);
        const std::string pk = unit_key(pu);
        const bool ok = unit_key(mu) == "msun" &&
                        (has_h ? pk == "pc" : (pk == "g/cm^3" || pk == "g/cm3" || pk == "gcm^-3" || pk == "gcm-3"));
        if (!ok) {
            err = "splash-ascii snapshot: only a '" + which + "' column was found, and this tool only knows how to "
                  "derive the other from it when mass is in M_Sun, h is in pc, and density is in g/cm^3 (the h = "
                  "hfact*(m/rho)^(1/3) relation needs m, h, and rho in a mutually consistent unit system) -- got "
                  "mass unit '" + mu + "' and " + which + " unit '" + pu + "' instead. Extend this derivation "
                  "before using it on a file in different units: " + path;
            return false;
        }
        std::cerr << "[kiwi] splash-ascii snapshot: only a '" << which << "' column was found in the header -- deriving "
                  << (has_h ? "density from h" : "h from density") << " via h = hfact*(m/rho)^(1/3) with hfact = "
                  << hfact << " (Price et al. 2018, eq. 6), converting through Msun/pc^3 since density is in g/cm^3 "
                              "but mass and h aren't: "
                  << path << "\n";
    }
    const int ix = at.at("x"), iy = at.at("y"), iz = at.at("z"), im = at.at("m"), ivx = at.at("vx"),
              ivy = at.at("vy"), ivz = at.at("vz");
    const int ih = has_h ? at.at("h") : -1, irho = has_rho ? at.at("rho") : -1, it = has_itype ? at.at("itype") : -1;
    const double conv = gcm3_per_msun_pc3();
    long long seen[16] = {0}, dropped[16] = {0};  // data rows and non-gas rows, per chunk
    auto parse = [&](int t, size_t a, size_t b, Columns& cols, ChunkError& e, size_t& nlines) {
        std::vector<double> v(static_cast<size_t>(ncols));
        long long s = 0, d = 0;
        char* line;
        size_t len;
        for (size_t at = a; next_line(buf.data(), at, b, line, len); ++nlines) {
            if (comment(line, len) || blank(line, len)) continue;
            const std::string row(line, len);
            std::istringstream in(row);
            for (int k = 0; k < ncols; ++k)
                if (!(in >> v[static_cast<size_t>(k)])) {
                    e.line = nlines;
                    e.msg = "splash-ascii snapshot: could not parse a particle row (expected " + std::to_string(ncols) +
                            " columns, matching the column-header line): " + path + ": " + row;
                    return;
                }
            ++s;
            if ((has_itype ? std::lround(v[static_cast<size_t>(it)]) : 1L) != 1) {
                ++d;
                continue;
            }
            const double m = v[static_cast<size_t>(im)];
            double h, rho;
            if (has_h && has_rho) {
                h = v[static_cast<size_t>(ih)];
                rho = v[static_cast<size_t>(irho)];
            } else if (has_h) {
                h = v[static_cast<size_t>(ih)];
                const double q = h / hfact;
                rho = (q > 0.0 ? m / (q * q * q) : 0.0) * conv;
            } else {
                rho = v[static_cast<size_t>(irho)];
                const double r = rho / conv;
                h = (m > 0.0 && r > 0.0) ? hfact * std::cbrt(m / r) : 0.0;
            }
            const double row9[9] = {v[static_cast<size_t>(ix)], v[static_cast<size_t>(iy)], v[static_cast<size_t>(iz)],
                                    v[static_cast<size_t>(ivx)], v[static_cast<size_t>(ivy)],
                                    v[static_cast<size_t>(ivz)], m, h, rho};
            for (int k = 0; k < 9; ++k) cols.c[k].push_back(row9[k]);
        }
        seen[t] = s;
        dropped[t] = d;
    };
    auto report = [&](const ChunkError& e, size_t) { err = e.msg; };
    if (!parallel_rows(buf, cur.at, size, parse, report, out)) return false;
    long long n_seen = 0, n_dropped = 0;
    for (int t = 0; t < 16; ++t) {
        n_seen += seen[t];
        n_dropped += dropped[t];
    }
    if (n_dropped > 0)
        std::cerr << "[kiwi] splash-ascii snapshot: dropped " << n_dropped << " of " << n_seen
                  << " particle row(s) with itype != 1 (non-gas -- sink/dust/star/etc.); only gas particles are used "
                     "for the grid: "
                  << path << "\n";
    if (out.c[0].empty()) {
        err = "splash-ascii snapshot file contained no gas (itype == 1) particle rows: " + path;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- phantom_binary (src/phantom_dump.cpp)
// Fortran unformatted sequential records ([4-byte length][payload][4-byte length]) walked with a cursor
// over the file in memory; failures carry the reference's "phantom-binary snapshot: <why> while
// reading <what>: <path>" text (RecordReader, :50-132).
struct Records {
    const char* p;
    size_t size, at = 0;
    bool swap = false;
    std::string path;
    std::string err;

    template <class T>
    T get(const char* q) const {
        unsigned char b[sizeof(T)];
        for (size_t i = 0; i < sizeof(T); ++i) b[i] = static_cast<unsigned char>(swap ? q[sizeof(T) - 1 - i] : q[i]);
        T v;
        std::memcpy(&v, b, sizeof(T));
        return v;
    }
    bool fail(const std::string& why, const std::string& what) {
        err = "phantom-binary snapshot: " + why + " while reading " + what + ": " + path;
        return false;
    }
    bool marker(const std::string& what, int32_t& n) {
        if (at + 4 > size) {
            at = size;
            return fail("unexpected end of file", what);
        }
        n = get<int32_t>(p + at);
        at += 4;
        if (n < 0) return fail("negative record length (records over 2 GB are not supported)", what);
        return true;
    }
    // the next record's payload (pointer into the file) and length
    bool record(const std::string& what, const char*& q, size_t& len) {
        int32_t n = 0, m = 0;
        if (!marker(what, n)) return false;
        if (at + static_cast<size_t>(n) > size) {
            at = size;
            return fail("unexpected end of file", what);
        }
        q = p + at;
        len = static_cast<size_t>(n);
        at += len;
        if (!marker(what, m)) return false;
        if (m != n) return fail("record markers disagree (" + std::to_string(n) + " vs " + std::to_string(m) + ")", what);
        return true;
    }
    bool skip(const std::string& what) {
        int32_t n = 0, m = 0;
        if (!marker(what, n)) return false;
        at = std::min(size, at + static_cast<size_t>(n));  // a seek past the end fails at the next marker
        if (!marker(what, m)) return false;
        if (m != n) return fail("record markers disagree while skipping", what);
        return true;
    }
};

struct Kind {
    int size;
    bool real;
};

double decode(const Records& R, const char* q, Kind k) {
    if (k.real) return k.size == 4 ? static_cast<double>(R.get<float>(q)) : R.get<double>(q);
    switch (k.size) {
        case 1: return static_cast<double>(R.get<int8_t>(q));
        case 2: return static_cast<double>(R.get<int16_t>(q));
        case 4: return static_cast<double>(R.get<int32_t>(q));
        default: return static_cast<double>(R.get<int64_t>(q));
    }
}

// the capture record (i1 r1 i2 iversion i3): byte order and default int / real widths (:60-95)
bool capture(const char* p, size_t size, bool& swap, int& ib, int& rb) {
    if (size < 4) return false;
    for (bool sw : {false, true}) {
        Records R{p, size};
        R.swap = sw;
        const int32_t len = R.get<int32_t>(p);
        if (len == 20) ib = 4, rb = 4;
        else if (len == 24) ib = 4, rb = 8;
        else if (len == 36) ib = 8, rb = 4;
        else if (len == 40) ib = 8, rb = 8;
        else continue;
        if (size < static_cast<size_t>(len) + 8) return false;
        const char* q = p + 4;
        if (R.get<int32_t>(q + len) != len) return false;
        auto gi = [&](int off) { return ib == 4 ? static_cast<int64_t>(R.get<int32_t>(q + off)) : R.get<int64_t>(q + off); };
        const double r1 = rb == 4 ? static_cast<double>(R.get<float>(q + ib)) : R.get<double>(q + ib);
        if (gi(0) != 60769 || gi(ib + rb) != 60878 || gi(3 * ib + rb) != 690706 || r1 != 60878.0) return false;
        swap = sw;
        return true;
    }
    return false;
}

std::string trim_spaces(const char* q, size_t n) {
    size_t a = 0, b = n;
    while (a < b && q[a] == ' ') ++a;
    while (b > a && q[b - 1] == ' ') --b;
    return std::string(q + a, b - a);
}

bool load_phantom(const std::string& path, const std::vector<char>& buf, const kiwi_phantom_units& units,
                  bool allow_missing_velocity, Columns& out, std::string& err) {
    const size_t size = buf.size() - 1;
    Records R{buf.data(), size};
    R.path = path;
    int ib = 0, rb = 0;
    if (!capture(buf.data(), size, R.swap, ib, rb)) {
        err = "phantom-binary snapshot: the file does not start with the Phantom capture pattern (is it a Phantom "
              "dump?): " + path;
        return false;
    }
    R.at = 4 + static_cast<size_t>(4 * ib + rb) + 4;
    auto bad = [&]() {
        err = R.err;
        return false;
    };
    const char* q;
    size_t len;
    if (!R.record("the file identifier", q, len)) return bad();
    if (len < 2 || (q[0] != 'F' && q[0] != 'S') || q[1] != 'T') {
        R.fail("only tagged full dumps (file identifier 'FT:...') are supported, got '" +
                   std::string(q, std::min<size_t>(len, 20)) + "'", "the file identifier");
        return bad();
    }
    const Kind kinds[8] = {{ib, false}, {1, false}, {2, false}, {4, false}, {8, false}, {rb, true}, {4, true}, {8, true}};
    constexpr int TAG = 16;
    // the global header: eight typed sections of (tag, value) pairs, a tag may repeat
    std::unordered_map<std::string, std::vector<double>> hdr;
    for (int k = 0; k < 8; ++k) {
        if (!R.record("a header section size", q, len)) return bad();
        if (len != 4) {
            R.fail("malformed header section size", "the global header");
            return bad();
        }
        const int32_t nv = R.get<int32_t>(q);
        if (nv <= 0) continue;
        const char *tq, *vq;
        size_t tl, vl;
        if (!R.record("header tags", tq, tl) || !R.record("header values", vq, vl)) return bad();
        if (tl != static_cast<size_t>(nv) * TAG || vl != static_cast<size_t>(nv) * static_cast<size_t>(kinds[k].size)) {
            R.fail("header section " + std::to_string(k) + " has inconsistent sizes", "the global header");
            return bad();
        }
        for (int i = 0; i < nv; ++i)
            hdr[trim_spaces(tq + static_cast<size_t>(i) * TAG, TAG)].push_back(
                decode(R, vq + static_cast<size_t>(i) * kinds[k].size, kinds[k]));
    }
    const std::vector<double>* need[5];
    const char* keys[5] = {"massoftype", "hfact", "udist", "umass", "utime"};
    for (int i = 0; i < 5; ++i) {
        auto it = hdr.find(keys[i]);
        if (it == hdr.end() || it->second.empty()) {
            R.fail(std::string("required header entry '") + keys[i] + "' is missing", "the global header");
            return bad();
        }
        need[i] = &it->second;
    }
    const std::vector<double>& massoftype = *need[0];
    const double hfact = (*need[1])[0], udist = (*need[2])[0], umass = (*need[3])[0], utime = (*need[4])[0];
    if (hfact <= 0.0 || udist <= 0.0 || umass <= 0.0 || utime <= 0.0) {
        R.fail("header hfact/udist/umass/utime must be > 0", "the global header");
        return bad();
    }
    auto first_of = [&](const char* key, bool& has) -> double {
        auto it = hdr.find(key);
        has = it != hdr.end();
        return has && !it->second.empty() ? it->second[0] : 0.0;
    };
    bool has;
    const double hb = first_of("nblocks", has);
    const int hdr_blocks = has ? std::max(1, static_cast<int>(hb)) : 1;
    // the array blocks
    if (!R.record("the number of array blocks", q, len)) return bad();
    if (len != 4) {
        R.fail("malformed block count", "the array section");
        return bad();
    }
    const int nblocks = R.get<int32_t>(q);
    if (nblocks <= 0 || nblocks % hdr_blocks != 0) {
        R.fail("array block count " + std::to_string(nblocks) + " is not a multiple of the header's nblocks " +
                   std::to_string(hdr_blocks), "the array section");
        return bad();
    }
    struct Block {
        int64_t n;
        int32_t nums[8];
    };
    std::vector<Block> blocks(static_cast<size_t>(nblocks));
    for (auto& b : blocks) {
        if (!R.record("a block header", q, len)) return bad();
        if (len != 8 + 4 * 8) {
            R.fail("malformed block header", "the array section");
            return bad();
        }
        b.n = R.get<int64_t>(q);
        for (int k = 0; k < 8; ++k) b.nums[k] = R.get<int32_t>(q + 8 + 4 * k);
        if (b.n < 0) {
            R.fail("negative particle count", "the array section");
            return bad();
        }
    }
    const double nt = first_of("ntypes", has);
    const int ntypes = has ? static_cast<int>(nt) : static_cast<int>(massoftype.size());
    bool several = false;
    if (hdr.count("npartoftype")) {
        const auto& npt = hdr["npartoftype"];
        for (int t = 1; t < ntypes && t < static_cast<int>(npt.size()); ++t)
            if (npt[t] > 0.0) several = true;
    }
    static const char* wanted[8] = {"x", "y", "z", "vx", "vy", "vz", "h", "itype"};
    const double lf = udist / units.length_cm, mf = umass / units.mass_g, vf = (udist / utime) / units.velocity_cms;
    const double rho_cgs = umass / (udist * udist * udist);
    const double m_gas = massoftype[0];
    long long n_seen = 0, n_dead = 0, n_other = 0;
    bool zeroed = false;
    const int nsub = nblocks / hdr_blocks;
    for (int bi = 0; bi < nblocks; ++bi) {
        const Block& bh = blocks[static_cast<size_t>(bi)];
        const bool sink = bi % nsub == 1;
        struct Arr {
            const char* q = nullptr;
            Kind k{0, false};
        };
        Arr arr[8];  // wanted order
        for (int k = 0; k < 8; ++k)
            for (int a = 0; a < bh.nums[k]; ++a) {
                if (!R.record("an array tag", q, len)) return bad();
                if (len != static_cast<size_t>(TAG)) {
                    R.fail("malformed array tag", "the array section");
                    return bad();
                }
                const std::string name = trim_spaces(q, TAG);
                int w = -1;
                for (int j = 0; j < 8; ++j)
                    if (name == wanted[j]) w = j;
                if (sink || w < 0 || arr[w].q) {
                    if (!R.skip("array '" + name + "'")) return bad();
                    continue;
                }
                if (!R.record("array '" + name + "'", q, len)) return bad();
                if (len != static_cast<size_t>(bh.n) * static_cast<size_t>(kinds[k].size)) {
                    R.fail("array '" + name + "' has " + std::to_string(len) + " bytes, expected " +
                               std::to_string(bh.n * kinds[k].size), "the array section");
                    return bad();
                }
                arr[w] = Arr{q, kinds[k]};
            }
        if (sink || bh.n == 0) continue;
        for (int j : {0, 1, 2, 6})
            if (!arr[j].q) {
                R.fail(std::string("array '") + wanted[j] + "' not found in a particle block", "the array section");
                return bad();
            }
        std::string missing;
        for (int j : {3, 4, 5})
            if (!arr[j].q) missing += std::string(missing.empty() ? "" : ", ") + wanted[j];
        if (!missing.empty()) {
            if (!allow_missing_velocity) {
                R.fail("velocity array(s) " + missing + " not found in a particle block (a Phantom small dump has "
                       "none). Set \"allow_missing_velocity\": true in the config to use the dump for density only, "
                       "with all velocities set to zero", "the array section");
                return bad();
            }
            zeroed = true;
        }
        const bool has_itype = arr[7].q != nullptr;
        if (!has_itype && several) {
            R.fail("the header lists particles of more than one type but the dump has no itype array",
                   "the array section");
            return bad();
        }
        auto val = [&](int j, int64_t i) {
            return arr[j].q ? decode(R, arr[j].q + static_cast<size_t>(i) * arr[j].k.size, arr[j].k) : 0.0;
        };
        for (int64_t i = 0; i < bh.n; ++i) {
            ++n_seen;
            const double h = val(6, i);
            if (!(h > 0.0)) {
                ++n_dead;
                continue;
            }
            if (has_itype && std::lround(val(7, i)) != 1) {
                ++n_other;
                continue;
            }
            const double qh = hfact / h;
            const double row[9] = {val(0, i) * lf, val(1, i) * lf, val(2, i) * lf, val(3, i) * vf, val(4, i) * vf,
                                   val(5, i) * vf, m_gas * mf, h * lf, m_gas * qh * qh * qh * rho_cgs};
            for (int k = 0; k < 9; ++k) out.c[k].push_back(row[k]);
        }
    }
    if (out.c[0].empty()) {
        R.fail("no gas particles found", "the array section");
        return bad();
    }
    std::cerr << "[kiwi] phantom-binary snapshot: kept " << out.c[0].size() << " of " << n_seen << " particle(s) (dropped "
              << n_dead << " dead with h <= 0 and " << n_other << " of another type; sink particles are ignored); gas "
              << "mass " << m_gas * mf << " each, density from the header's hfact = " << hfact << ": " << path << "\n";
    if (zeroed)
        std::cerr << "[kiwi] WARNING: this dump has no velocity arrays (a Phantom small dump?); all velocities are set "
                     "to zero because allow_missing_velocity is true. Density is unaffected, but every velocity output "
                     "(vx/vy/vz grids, POLARIS velocities, ...) is meaningless.\n";
    return true;
}

// ---------------------------------------------------------------- detection (:495-509)
// A Phantom dump opens with the capture record (is_phantom_dump, src/phantom_dump.cpp:145-154)
bool phantom_capture(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char head[4 + 40 + 4];
    const size_t got = std::fread(head, 1, sizeof(head), f);
    std::fclose(f);
    bool swap;
    int ib, rb;
    return capture(head, got, swap, ib, rb);
}

bool detect(const std::string& path, std::string& fmt, std::string& err) {
    if (phantom_capture(path)) {
        fmt = "phantom_binary";
        return true;
    }
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "could not open SPH snapshot file: " + path;
        return false;
    }
    std::string first;
    int c;
    while ((c = std::fgetc(f)) != EOF && c != '\n') first += static_cast<char>(c);
    std::fclose(f);
    fmt = (!first.empty() && first[0] == '#' && first.find("SPLASH") != std::string::npos) ? "splash_ascii"
                                                                                           : "plain_ascii";
    return true;
}

}  // namespace

extern "C" {

int kiwi_io_abi_version(void) { return KIWI_IO_ABI_VERSION; }

const char* kiwi_io_last_error(void) { return g_err.c_str(); }

int kiwi_io_set_threads(int n) {
    g_threads = n < 0 ? 0 : n;
    return worker_count();
}

int kiwi_io_detect_format(const char* path, char* format, int capacity) {
    g_err.clear();
    if (!path || !format || capacity < 16) {
        g_err = "null argument or format buffer under 16 bytes";
        return KIWI_IO_ERR_INVALID_ARG;
    }
    std::string fmt;
    if (!detect(path, fmt, g_err)) return KIWI_IO_ERR_READ;
    std::snprintf(format, static_cast<size_t>(capacity), "%s", fmt.c_str());
    return KIWI_IO_OK;
}

int kiwi_io_load_snapshot(const char* path, const char* format, double hfact, kiwi_snapshot* out) {
    return kiwi_io_load_snapshot_ex(path, format, hfact, nullptr, 0, out);
}

int kiwi_io_load_snapshot_ex(const char* path_c, const char* format_c, double hfact, const kiwi_phantom_units* units_c,
                             int allow_missing_velocity, kiwi_snapshot* out) {
    g_err.clear();
    if (!path_c || !out) {
        g_err = "null argument";
        return KIWI_IO_ERR_INVALID_ARG;
    }
    *out = kiwi_snapshot{};
    const std::string path = path_c, format = format_c ? format_c : "auto";
    std::string fmt = format;
    if (fmt == "auto" && !detect(path, fmt, g_err)) return KIWI_IO_ERR_READ;
    if (fmt != "plain_ascii" && fmt != "splash_ascii" && fmt != "phantom_binary") {
        g_err = "unknown SPH snapshot format '" + format +
                "' (expected auto, plain_ascii, splash_ascii, or phantom_binary)";
        return KIWI_IO_ERR_INVALID_ARG;
    }
    std::vector<char> buf;
    if (!read_file(path, buf)) {
        g_err = "could not open SPH snapshot file: " + path;
        return KIWI_IO_ERR_READ;
    }
    Columns cols;
    const kiwi_phantom_units units = units_c ? *units_c : kiwi_phantom_units{3.086e18, 1.989e33, 1.0e5};
    const bool ok = fmt == "plain_ascii"    ? load_plain(path, buf, cols, g_err)
                    : fmt == "splash_ascii" ? load_splash(path, hfact, buf, cols, g_err)
                                            : load_phantom(path, buf, units, allow_missing_velocity != 0, cols, g_err);
    if (!ok) return KIWI_IO_ERR_FORMAT;
    const size_t n = cols.c[0].size();
    if (n > 0x7fffffffu) {
        g_err = "more than 2^31-1 particles";
        return KIWI_IO_ERR_FORMAT;
    }
    double* arrays[9];
    for (int k = 0; k < 9; ++k) {
        arrays[k] = static_cast<double*>(std::malloc(sizeof(double) * std::max<size_t>(n, 1)));
        if (arrays[k]) std::memcpy(arrays[k], cols.c[k].data(), sizeof(double) * n);
    }
    int* id = static_cast<int*>(std::malloc(sizeof(int) * std::max<size_t>(n, 1)));
    bool alloc_ok = id != nullptr;
    for (int k = 0; k < 9; ++k) alloc_ok = alloc_ok && arrays[k];
    if (!alloc_ok) {
        for (int k = 0; k < 9; ++k) std::free(arrays[k]);
        std::free(id);
        g_err = "out of host memory";
        return KIWI_IO_ERR_READ;
    }
    for (size_t i = 0; i < n; ++i) id[i] = static_cast<int>(i + 1);  // 1-based among the rows kept
    *out = kiwi_snapshot{static_cast<int64_t>(n), id, arrays[0], arrays[1], arrays[2], arrays[3], arrays[4], arrays[5],
                         arrays[6], arrays[7], arrays[8]};
    return KIWI_IO_OK;
}

void kiwi_io_free_snapshot(kiwi_snapshot* s) {
    if (!s) return;
    double* arrays[9] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m, s->h, s->rho};
    for (double* a : arrays) std::free(a);
    std::free(s->id);
    *s = kiwi_snapshot{};
}

}  // extern "C"
