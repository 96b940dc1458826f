// libkiwi_io: KIWI's SPH snapshot readers (src/snapshot.cpp:17-522) as a native host library.  The
// file is read once into memory, cut into lines with getline's rules, and the lines are parsed by a
// pool of threads, each into its own column runs, which are joined in file order.  Every line is
// parsed with the same library routine the reference uses for it (strtod for plain_ascii, a string
// stream's double extraction for splash_ascii), so the values, the rows kept, the first error and its
// message are the reference's.
//
//   kiwi_io_detect_format   phantom_binary (a Phantom capture record first), splash_ascii (a first
//                           line "#...SPLASH..."), else plain_ascii (detect_snapshot_format, :495-509)
//   kiwi_io_load_snapshot   plain_ascii (load_plain_ascii, :32-84) and splash_ascii (load_splash_ascii,
//                           :101-491), "auto" resolved as above; phantom_binary is not read here
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

// ---------------------------------------------------------------- detection (:495-509)
// A Phantom dump opens with the capture record (i1 r1 i2 iversion i3 = 60769, 60878.0, 60878, v, 690706)
// framed by 4-byte lengths of 20, 24, 36 or 40 bytes in either byte order (src/phantom_dump.cpp:60-95).
bool phantom_capture(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    unsigned char head[4 + 40 + 4];
    const size_t got = std::fread(head, 1, sizeof(head), f);
    std::fclose(f);
    if (got < 4) return false;
    auto load = [](const unsigned char* p, int n, bool swap, unsigned char* dst) {
        for (int i = 0; i < n; ++i) dst[i] = swap ? p[n - 1 - i] : p[i];
    };
    for (bool swap : {false, true}) {
        int32_t len;
        load(head, 4, swap, reinterpret_cast<unsigned char*>(&len));
        int ib, rb;
        if (len == 20) ib = 4, rb = 4;
        else if (len == 24) ib = 4, rb = 8;
        else if (len == 36) ib = 8, rb = 4;
        else if (len == 40) ib = 8, rb = 8;
        else continue;
        if (got < static_cast<size_t>(len) + 8) return false;
        const unsigned char* p = head + 4;
        int32_t tail;
        load(p + len, 4, swap, reinterpret_cast<unsigned char*>(&tail));
        if (tail != len) return false;
        auto get_int = [&](int off) -> int64_t {
            if (ib == 4) {
                int32_t v;
                load(p + off, 4, swap, reinterpret_cast<unsigned char*>(&v));
                return v;
            }
            int64_t v;
            load(p + off, 8, swap, reinterpret_cast<unsigned char*>(&v));
            return v;
        };
        double r1;
        if (rb == 4) {
            float v;
            load(p + ib, 4, swap, reinterpret_cast<unsigned char*>(&v));
            r1 = v;
        } else {
            load(p + ib, 8, swap, reinterpret_cast<unsigned char*>(&r1));
        }
        return get_int(0) == 60769 && r1 == 60878.0 && get_int(ib + rb) == 60878 && get_int(3 * ib + rb) == 690706;
    }
    return false;
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

int kiwi_io_load_snapshot(const char* path_c, const char* format_c, double hfact, kiwi_snapshot* out) {
    g_err.clear();
    if (!path_c || !out) {
        g_err = "null argument";
        return KIWI_IO_ERR_INVALID_ARG;
    }
    *out = kiwi_snapshot{};
    const std::string path = path_c, format = format_c ? format_c : "auto";
    std::string fmt = format;
    if (fmt == "auto" && !detect(path, fmt, g_err)) return KIWI_IO_ERR_READ;
    if (fmt == "phantom_binary") {
        g_err = "phantom_binary snapshots are not read by libkiwi_io (use the reference's load_phantom_dump): " + path;
        return KIWI_IO_ERR_UNSUPPORTED;
    }
    if (fmt != "plain_ascii" && fmt != "splash_ascii") {
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
    const bool ok = fmt == "plain_ascii" ? load_plain(path, buf, cols, g_err) : load_splash(path, hfact, buf, cols, g_err);
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
