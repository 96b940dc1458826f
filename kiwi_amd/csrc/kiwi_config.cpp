// The run configuration and the provenance manifest of KIWI's stages on the host: the JSON config
// read with the reference's keys, defaults and error messages (Config::load_from_file,
// src/config.cpp:59-206), the run directory (:279-287), and manifest.json written byte for byte as the
// reference writes it (build_manifest / write_manifest / check_manifest_consistency,
// src/manifest.cpp:36-199).
//
// The reference reads JSON with nlohmann::json; its exceptions' texts are part of the messages the
// reference prints (a missing key: "[json.exception.out_of_range.403] key 'k' not found", a wrong type:
// "[json.exception.type_error.302] type must be number, but is string"), so this reader produces the
// same texts for those.  A syntax error is reported with the library's prefix and the line and column
// of the offending character.  Objects keep their keys sorted, as nlohmann::json's default std::map
// does, which is the key order of the manifest; numbers keep the library's three kinds (integer,
// unsigned, floating) because the manifest prints each kind its own way.
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kiwi_io.h"
#include "kernel_tables.h"

namespace {

// ------------------------------------------------------------------------------------------------
// JSON values
// ------------------------------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Object, Array, String, Boolean, Integer, Unsigned, Float } kind = Null;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;
    double d = 0.0;
    std::string s;
    std::vector<Json> a;
    std::map<std::string, Json> o;

    static Json obj() {
        Json j;
        j.kind = Object;
        return j;
    }
    static Json str(const std::string& v) {
        Json j;
        j.kind = String;
        j.s = v;
        return j;
    }
    static Json boolean(bool v) {
        Json j;
        j.kind = Boolean;
        j.b = v;
        return j;
    }
    static Json integer(int64_t v) {
        Json j;
        j.kind = Integer;
        j.i = v;
        return j;
    }
    static Json unsig(uint64_t v) {
        Json j;
        j.kind = Unsigned;
        j.u = v;
        return j;
    }
    static Json real(double v) {
        Json j;
        j.kind = Float;
        j.d = v;
        return j;
    }
    bool is_number() const { return kind == Integer || kind == Unsigned || kind == Float; }
};

// the type names nlohmann::json puts into its messages
const char* type_name(const Json& j) {
    switch (j.kind) {
        case Json::Null: return "null";
        case Json::Object: return "object";
        case Json::Array: return "array";
        case Json::String: return "string";
        case Json::Boolean: return "boolean";
        default: return "number";
    }
}

struct OutOfRange : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct TypeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// j.at(key) of an object
const Json& at(const Json& j, const std::string& key) {
    if (j.kind != Json::Object)
        throw TypeError(std::string("[json.exception.type_error.304] cannot use at() with ") + type_name(j));
    auto it = j.o.find(key);
    if (it == j.o.end()) throw OutOfRange("[json.exception.out_of_range.403] key '" + key + "' not found");
    return it->second;
}

bool contains(const Json& j, const std::string& key) { return j.kind == Json::Object && j.o.count(key) > 0; }

double get_double(const Json& j) {
    switch (j.kind) {
        case Json::Integer: return static_cast<double>(j.i);
        case Json::Unsigned: return static_cast<double>(j.u);
        case Json::Float: return j.d;
        default: throw TypeError(std::string("[json.exception.type_error.302] type must be number, but is ") + type_name(j));
    }
}

int64_t get_int(const Json& j) {
    switch (j.kind) {
        case Json::Integer: return j.i;
        case Json::Unsigned: return static_cast<int64_t>(j.u);
        case Json::Float: return static_cast<int64_t>(j.d);
        default: throw TypeError(std::string("[json.exception.type_error.302] type must be number, but is ") + type_name(j));
    }
}

bool get_bool(const Json& j) {
    if (j.kind != Json::Boolean)
        throw TypeError(std::string("[json.exception.type_error.302] type must be boolean, but is ") + type_name(j));
    return j.b;
}

std::string get_string(const Json& j) {
    if (j.kind != Json::String)
        throw TypeError(std::string("[json.exception.type_error.302] type must be string, but is ") + type_name(j));
    return j.s;
}

// j.value(key, default): the default when the key is absent; a present key must convert
void value_check(const Json& j) {
    if (j.kind != Json::Object)
        throw TypeError(std::string("[json.exception.type_error.306] cannot use value() with ") + type_name(j));
}
double value_double(const Json& j, const std::string& k, double def) {
    value_check(j);
    auto it = j.o.find(k);
    return it == j.o.end() ? def : get_double(it->second);
}
int64_t value_int(const Json& j, const std::string& k, int64_t def) {
    value_check(j);
    auto it = j.o.find(k);
    return it == j.o.end() ? def : get_int(it->second);
}
bool value_bool(const Json& j, const std::string& k, bool def) {
    value_check(j);
    auto it = j.o.find(k);
    return it == j.o.end() ? def : get_bool(it->second);
}
std::string value_string(const Json& j, const std::string& k, const std::string& def) {
    value_check(j);
    auto it = j.o.find(k);
    return it == j.o.end() ? def : get_string(it->second);
}

// ------------------------------------------------------------------------------------------------
// Parser (RFC 8259), reporting a syntax error with its line and column
// ------------------------------------------------------------------------------------------------
struct Parser {
    const std::string& t;
    size_t p = 0;
    explicit Parser(const std::string& text) : t(text) {}

    [[noreturn]] void fail(const std::string& what) {
        size_t line = 1, col = 0;
        for (size_t k = 0; k < std::min(p + 1, t.size()); ++k) {
            if (t[k] == '\n') {
                ++line;
                col = 0;
            } else {
                ++col;
            }
        }
        if (p >= t.size()) ++col;
        std::ostringstream m;
        m << "[json.exception.parse_error.101] parse error at line " << line << ", column " << col << ": " << what;
        throw std::runtime_error(m.str());
    }
    void ws() {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
    }
    Json parse_document() {
        ws();
        if (p >= t.size()) fail("attempting to parse an empty input; check that your input string or stream contains the expected JSON");
        Json v = value();
        ws();
        if (p < t.size()) fail("syntax error while parsing value - unexpected trailing characters; expected end of input");
        return v;
    }
    Json value() {
        ws();
        if (p >= t.size()) fail("syntax error while parsing value - unexpected end of input; expected '[', '{', or a literal");
        const char c = t[p];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Json::str(string());
        if (c == 't' || c == 'f' || c == 'n') return literal();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("syntax error while parsing value - invalid literal; expected '[', '{', or a literal");
    }
    Json object() {
        Json j = Json::obj();
        ++p;
        ws();
        if (p < t.size() && t[p] == '}') {
            ++p;
            return j;
        }
        for (;;) {
            ws();
            if (p >= t.size() || t[p] != '"') fail("syntax error while parsing object key - invalid literal; expected string literal");
            std::string k = string();
            ws();
            if (p >= t.size() || t[p] != ':') fail("syntax error while parsing object separator - invalid literal; expected ':'");
            ++p;
            j.o[k] = value();  // a repeated key keeps the last value, as nlohmann::json does
            ws();
            if (p < t.size() && t[p] == ',') {
                ++p;
                continue;
            }
            if (p < t.size() && t[p] == '}') {
                ++p;
                return j;
            }
            fail("syntax error while parsing object - invalid literal; expected '}'");
        }
    }
    Json array() {
        Json j;
        j.kind = Json::Array;
        ++p;
        ws();
        if (p < t.size() && t[p] == ']') {
            ++p;
            return j;
        }
        for (;;) {
            j.a.push_back(value());
            ws();
            if (p < t.size() && t[p] == ',') {
                ++p;
                continue;
            }
            if (p < t.size() && t[p] == ']') {
                ++p;
                return j;
            }
            fail("syntax error while parsing array - invalid literal; expected ']'");
        }
    }
    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) {
            o += static_cast<char>(cp);
        } else if (cp < 0x800) {
            o += static_cast<char>(0xC0 | (cp >> 6));
            o += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += static_cast<char>(0xE0 | (cp >> 12));
            o += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            o += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            o += static_cast<char>(0xF0 | (cp >> 18));
            o += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            o += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            o += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (p + 4 > t.size()) fail("syntax error while parsing value - invalid string: '\\u' must be followed by 4 hex digits");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = t[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
            else fail("syntax error while parsing value - invalid string: '\\u' must be followed by 4 hex digits");
        }
        return v;
    }
    std::string string() {
        std::string o;
        ++p;  // opening quote
        for (;;) {
            if (p >= t.size()) fail("syntax error while parsing value - invalid string: missing closing quote");
            const unsigned char c = static_cast<unsigned char>(t[p++]);
            if (c == '"') return o;
            if (c < 0x20) fail("syntax error while parsing value - invalid string: control character must be escaped");
            if (c != '\\') {
                o += static_cast<char>(c);
                continue;
            }
            if (p >= t.size()) fail("syntax error while parsing value - invalid string: missing closing quote");
            const char e = t[p++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (p + 2 > t.size() || t[p] != '\\' || t[p + 1] != 'u')
                            fail("syntax error while parsing value - invalid string: surrogate U+D800..U+DBFF must be followed by U+DC00..U+DFFF");
                        p += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF)
                            fail("syntax error while parsing value - invalid string: surrogate U+D800..U+DBFF must be followed by U+DC00..U+DFFF");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("syntax error while parsing value - invalid string: forbidden character after backslash");
            }
        }
    }
    Json literal() {
        static const char* words[3] = {"true", "false", "null"};
        for (int w = 0; w < 3; ++w) {
            const size_t n = std::strlen(words[w]);
            if (t.compare(p, n, words[w]) == 0) {
                p += n;
                if (w == 2) return Json();
                return Json::boolean(w == 0);
            }
        }
        fail("syntax error while parsing value - invalid literal; expected '[', '{', or a literal");
    }
    Json number() {
        const size_t s = p;
        bool neg = false, frac = false;
        if (t[p] == '-') {
            neg = true;
            ++p;
        }
        if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9')) fail("syntax error while parsing value - invalid number; expected digit after '-'");
        if (t[p] == '0') {
            ++p;
        } else {
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        if (p < t.size() && t[p] == '.') {
            frac = true;
            ++p;
            if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9')) fail("syntax error while parsing value - invalid number; expected digit after '.'");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            frac = true;
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
            if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9')) fail("syntax error while parsing value - invalid number; expected digit after exponent sign");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        const std::string tok = t.substr(s, p - s);
        if (!frac) {  // integers: unsigned when non-negative, signed otherwise; floating when out of range
            errno = 0;
            if (neg) {
                const long long v = std::strtoll(tok.c_str(), nullptr, 10);
                if (errno == 0) return Json::integer(v);
            } else {
                const unsigned long long v = std::strtoull(tok.c_str(), nullptr, 10);
                if (errno == 0) return Json::unsig(v);
            }
        }
        return Json::real(std::strtod(tok.c_str(), nullptr));
    }
};

// ------------------------------------------------------------------------------------------------
// Serializer: nlohmann::json's dump(indent) layout
// ------------------------------------------------------------------------------------------------
// The shortest decimal digits that read back as v, printed the way the library prints a double: plain
// digits with a trailing ".0" when the value is integral, a decimal point inside the digits, "0.000ddd"
// for small values (down to three zeros), otherwise d.ddde+XX with at least two exponent digits.
std::string format_double(double v) {
    if (std::isnan(v) || std::isinf(v)) return "null";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    // buf: [-]d.ddde[+-]XX -> digits and decimal exponent
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t e = s.find('e');
    std::string digits = s.substr(0, e);
    const int exp10 = std::atoi(s.c_str() + e + 1);
    digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int k = static_cast<int>(digits.size());
    const int n = exp10 + 1;  // position of the decimal point after the first n digits
    std::string out;
    if (k <= n && n <= 15) {
        out = digits + std::string(static_cast<size_t>(n - k), '0') + ".0";
    } else if (0 < n && n <= 15) {
        out = digits.substr(0, static_cast<size_t>(n)) + "." + digits.substr(static_cast<size_t>(n));
    } else if (-4 < n && n <= 0) {
        out = "0." + std::string(static_cast<size_t>(-n), '0') + digits;
    } else {
        out = digits.substr(0, 1);
        if (k > 1) out += "." + digits.substr(1);
        const int x = n - 1;
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", x < 0 ? '-' : '+', x < 0 ? -x : x);
        out += eb;
    }
    return neg ? "-" + out : out;
}

void dump_string(std::string& o, const std::string& s) {
    o += '"';
    for (const unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04x", c);
                    o += b;
                } else {
                    o += static_cast<char>(c);
                }
        }
    }
    o += '"';
}

// indent < 0: compact (no spaces); otherwise pretty with `indent` spaces per level
void dump(std::string& o, const Json& j, int indent, int level) {
    switch (j.kind) {
        case Json::Null: o += "null"; return;
        case Json::Boolean: o += j.b ? "true" : "false"; return;
        case Json::Integer: o += std::to_string(j.i); return;
        case Json::Unsigned: o += std::to_string(j.u); return;
        case Json::Float: o += format_double(j.d); return;
        case Json::String: dump_string(o, j.s); return;
        case Json::Array: {
            if (j.a.empty()) {
                o += "[]";
                return;
            }
            o += '[';
            for (size_t k = 0; k < j.a.size(); ++k) {
                if (indent >= 0) o += '\n' + std::string(static_cast<size_t>(indent * (level + 1)), ' ');
                dump(o, j.a[k], indent, level + 1);
                if (k + 1 < j.a.size()) o += ',';
            }
            if (indent >= 0) o += '\n' + std::string(static_cast<size_t>(indent * level), ' ');
            o += ']';
            return;
        }
        case Json::Object: {
            if (j.o.empty()) {
                o += "{}";
                return;
            }
            o += '{';
            size_t k = 0;
            for (const auto& kv : j.o) {
                if (indent >= 0) o += '\n' + std::string(static_cast<size_t>(indent * (level + 1)), ' ');
                dump_string(o, kv.first);
                o += indent >= 0 ? ": " : ":";
                dump(o, kv.second, indent, level + 1);
                if (++k < j.o.size()) o += ',';
            }
            if (indent >= 0) o += '\n' + std::string(static_cast<size_t>(indent * level), ' ');
            o += '}';
            return;
        }
    }
}

std::string dumps(const Json& j, int indent) {
    std::string o;
    dump(o, j, indent, 0);
    return o;
}

bool equal(const Json& a, const Json& b) {
    if (a.is_number() && b.is_number()) {  // the library compares numbers of mixed kinds by value
        if (a.kind == Json::Float || b.kind == Json::Float) return get_double(a) == get_double(b);
        if (a.kind == b.kind) return a.kind == Json::Integer ? a.i == b.i : a.u == b.u;
        const int64_t si = a.kind == Json::Integer ? a.i : b.i;
        const uint64_t ui = a.kind == Json::Unsigned ? a.u : b.u;
        return si >= 0 && static_cast<uint64_t>(si) == ui;
    }
    if (a.kind != b.kind) return false;
    switch (a.kind) {
        case Json::Null: return true;
        case Json::Boolean: return a.b == b.b;
        case Json::String: return a.s == b.s;
        case Json::Array:
            if (a.a.size() != b.a.size()) return false;
            for (size_t k = 0; k < a.a.size(); ++k)
                if (!equal(a.a[k], b.a[k])) return false;
            return true;
        case Json::Object:
            if (a.o.size() != b.o.size()) return false;
            for (auto ia = a.o.begin(), ib = b.o.begin(); ia != a.o.end(); ++ia, ++ib)
                if (ia->first != ib->first || !equal(ia->second, ib->second)) return false;
            return true;
        default: return false;
    }
}

// ------------------------------------------------------------------------------------------------
// The configuration (src/config.cpp:59-206), its JSON form (:208-265) and its run directory
// ------------------------------------------------------------------------------------------------
struct Cfg {
    std::string run_name, input_snapshot, input_snapshot_format = "auto", work_folder;
    double dom[6] = {0, 0, 0, 0, 0, 0};
    double block_scale_factor = 3.0;
    bool buffer_enabled = false;
    std::string buffer_strategy = "coverage_grid";
    double gs_dx = 0, gs_dy = 0, gs_dz = 0, gs_s1 = 3.0, gs_s2 = 4.0, gs_s3 = 5.0;
    double cg_spacing_fraction = 0.05, cg_radius_factor = 2.0;
    int lloyd_iterations = 0;
    bool lloyd_save_each = false;
    bool mapping_tab = false;
    std::string mapping_table;
    bool diag_enabled = false;
    std::string diag_python = "python3";
    int cmap_resolution = 128;
    int cg_res = 128, cg_rx = 0, cg_ry = 0, cg_rz = 0;
    bool cg_tab = false;
    std::string cg_table;
    std::string kernel_name = "cubic_spline";
    double kernel_support = 0.0, kernel_hfact = 0.0;
    bool kernel_generic = false;
    double pu_length_cm = 3.086e18, pu_mass_g = 1.989e33, pu_velocity_cms = 1.0e5;
    bool allow_missing_velocity = false;
    unsigned pw_sim_id = 50;
    double pw_length_scale_m = 3.08567758e16, pw_velocity_scale = 1000.0,
           pw_density_scale = 1.9891e30 / (3.08567758e16 * 3.08567758e16 * 3.08567758e16), pw_temperature = 65.0;
};

// The kernel a name selects, as get_kernel reads names (src/kernel_spec_data.cpp:175-187): lower case,
// '-' and ' ' read as '_', three spellings per kernel; -1 for an unknown name.
const KiwiKernelTable kTables[KIWI_TAB_NKERNELS] = {KIWI_KERNEL_TABLE_DATA};

int kernel_id_of(const std::string& name) {
    std::string n = name;
    for (char& ch : n) {
        ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        if (ch == '-' || ch == ' ') ch = '_';
    }
    static const char* spellings[6][3] = {{"cubic", "cubic_spline", "m4"},         {"quartic", "quartic_spline", "m5"},
                                          {"quintic", "quintic_spline", "m6"},     {"wendland_c2", "wendlandc2", "wc2"},
                                          {"wendland_c4", "wendlandc4", "wc4"},    {"wendland_c6", "wendlandc6", "wc6"}};
    for (int k = 0; k < 6; ++k)
        for (const char* s : spellings[k])
            if (n == s) return k;
    return -1;
}

bool kernel_is_default(const Cfg& c) {
    return c.kernel_name == "cubic_spline" && c.kernel_support == 0.0 && c.kernel_hfact == 0.0 && !c.kernel_generic;
}
bool units_default(const Cfg& c) {
    return c.pu_length_cm == 3.086e18 && c.pu_mass_g == 1.989e33 && c.pu_velocity_cms == 1.0e5;
}
double support_in_h(const Cfg& c) {
    return c.kernel_support > 0.0 ? c.kernel_support : kTables[kernel_id_of(c.kernel_name)].radkern;
}
bool generic_path(const Cfg& c) {
    return c.kernel_generic || kernel_id_of(c.kernel_name) != 0 || support_in_h(c) != 2.0;
}

std::string read_file(const std::string& path, bool* ok) {
    std::ifstream f(path, std::ios::binary);
    *ok = static_cast<bool>(f);
    if (!*ok) return {};
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

Cfg load_config(const std::string& path) {
    bool ok = false;
    const std::string text = read_file(path, &ok);
    if (!ok) throw ConfigError("could not open config file: " + path);
    Json j;
    try {
        Parser ps(text);
        j = ps.parse_document();
    } catch (const std::runtime_error& e) {
        throw ConfigError("config file is not valid JSON (" + path + "): " + e.what());
    }
    Cfg c;
    try {
        c.run_name = get_string(at(j, "run_name"));
        const Json& d = at(j, "domain");
        static const char* dk[6] = {"x_min", "x_max", "y_min", "y_max", "z_min", "z_max"};
        for (int k = 0; k < 6; ++k) c.dom[k] = get_double(at(d, dk[k]));
        if (!(c.dom[0] < c.dom[1])) throw ConfigError("domain.x_min must be < domain.x_max");
        if (!(c.dom[2] < c.dom[3])) throw ConfigError("domain.y_min must be < domain.y_max");
        if (!(c.dom[4] < c.dom[5])) throw ConfigError("domain.z_min must be < domain.z_max");
        if (contains(j, "grid_blocks")) {
            c.block_scale_factor = value_double(at(j, "grid_blocks"), "block_scale_factor", 3.0);
            if (c.block_scale_factor <= 0.0) throw ConfigError("grid_blocks.block_scale_factor must be > 0");
        }
        c.input_snapshot = get_string(at(j, "input_snapshot"));
        c.input_snapshot_format = value_string(j, "input_snapshot_format", "auto");
        if (c.input_snapshot_format != "auto" && c.input_snapshot_format != "plain_ascii" &&
            c.input_snapshot_format != "splash_ascii" && c.input_snapshot_format != "phantom_binary")
            throw ConfigError("input_snapshot_format must be 'auto', 'plain_ascii', 'splash_ascii', or "
                              "'phantom_binary' (got '" + c.input_snapshot_format + "')");
        c.work_folder = get_string(at(j, "work_folder"));
        if (contains(j, "buffer")) {
            const Json& b = at(j, "buffer");
            c.buffer_enabled = value_bool(b, "enabled", false);
            c.buffer_strategy = value_string(b, "strategy", "coverage_grid");
            if (c.buffer_enabled && c.buffer_strategy == "graded_shells") {
                if (!contains(b, "graded_shells"))
                    throw ConfigError("buffer.strategy is graded_shells but buffer.graded_shells params are missing");
                const Json& g = at(b, "graded_shells");
                c.gs_dx = get_double(at(g, "dx"));
                c.gs_dy = get_double(at(g, "dy"));
                c.gs_dz = get_double(at(g, "dz"));
                c.gs_s1 = value_double(g, "shell_spacing_1", 3.0);
                c.gs_s2 = value_double(g, "shell_spacing_2", 4.0);
                c.gs_s3 = value_double(g, "shell_spacing_3", 5.0);
            }
            if (c.buffer_enabled && c.buffer_strategy == "coverage_grid") {
                value_check(b);
                const Json cg = contains(b, "coverage_grid") ? at(b, "coverage_grid") : Json::obj();
                c.cg_spacing_fraction = value_double(cg, "spacing_fraction", 0.05);
                c.cg_radius_factor = value_double(cg, "coverage_radius_factor", 2.0);
                if (c.cg_spacing_fraction <= 0.0) throw ConfigError("buffer.coverage_grid.spacing_fraction must be > 0");
                if (c.cg_radius_factor <= 0.0) throw ConfigError("buffer.coverage_grid.coverage_radius_factor must be > 0");
            }
        }
        if (contains(j, "lloyd")) {
            const Json& l = at(j, "lloyd");
            c.lloyd_iterations = static_cast<int>(value_int(l, "iterations", 0));
            c.lloyd_save_each = value_bool(l, "save_each_iteration", false);
            if (c.lloyd_iterations < 0) throw ConfigError("lloyd.iterations must be >= 0");
        }
        if (contains(j, "mapping")) {
            const Json& m = at(j, "mapping");
            c.mapping_tab = value_bool(m, "use_tabulated_integral", false);
            c.mapping_table = value_string(m, "density_table", std::string());
        }
        if (contains(j, "diagnostics")) {
            const Json& dg = at(j, "diagnostics");
            c.diag_enabled = value_bool(dg, "enabled", false);
            c.diag_python = value_string(dg, "python_executable", "python3");
            if (c.diag_python.empty()) throw ConfigError("diagnostics.python_executable must not be empty");
        }
        if (contains(j, "cartesian_map")) {
            c.cmap_resolution = static_cast<int>(value_int(at(j, "cartesian_map"), "resolution", 128));
            if (c.cmap_resolution <= 0) throw ConfigError("cartesian_map.resolution must be > 0");
        }
        if (contains(j, "cartesian_grid")) {
            const Json& g = at(j, "cartesian_grid");
            c.cg_res = static_cast<int>(value_int(g, "resolution", 128));
            c.cg_rx = static_cast<int>(value_int(g, "resolution_x", 0));
            c.cg_ry = static_cast<int>(value_int(g, "resolution_y", 0));
            c.cg_rz = static_cast<int>(value_int(g, "resolution_z", 0));
            c.cg_tab = value_bool(g, "use_tabulated_integral", false);
            c.cg_table = value_string(g, "density_table", std::string());
            if (c.cg_res <= 0) throw ConfigError("cartesian_grid.resolution must be > 0");
            if (c.cg_rx < 0 || c.cg_ry < 0 || c.cg_rz < 0)
                throw ConfigError("cartesian_grid.resolution_x/resolution_y/resolution_z must be >= 0 "
                                  "(0 means: fall back to resolution)");
        }
        if (contains(j, "kernel")) {
            const Json& k = at(j, "kernel");
            c.kernel_name = value_string(k, "name", "cubic_spline");
            c.kernel_support = value_double(k, "support_radius_in_h", 0.0);
            c.kernel_hfact = value_double(k, "hfact", 0.0);
            c.kernel_generic = value_bool(k, "generic_path", false);
            if (kernel_id_of(c.kernel_name) < 0)
                throw ConfigError("kernel.name: unknown kernel '" + c.kernel_name +
                                  "' (expected one of: cubic_spline, quartic_spline, quintic_spline, wendland_c2, "
                                  "wendland_c4, wendland_c6)");
            if (c.kernel_support < 0.0)
                throw ConfigError("kernel.support_radius_in_h must be > 0 (or 0 / unset for the kernel's default)");
            if (c.kernel_hfact < 0.0)
                throw ConfigError("kernel.hfact must be > 0 (or 0 / unset for the kernel's default)");
        }
        if (generic_path(c) && (c.mapping_tab || c.cg_tab))
            throw ConfigError("use_tabulated_integral is only available for the cubic spline kernel with its "
                              "default support radius (and kernel.generic_path unset); the selected kernel "
                              "settings need the analytic integral -- set use_tabulated_integral to false");
        if (contains(j, "phantom_units")) {
            const Json& u = at(j, "phantom_units");
            c.pu_length_cm = value_double(u, "length_cm", c.pu_length_cm);
            c.pu_mass_g = value_double(u, "mass_g", c.pu_mass_g);
            c.pu_velocity_cms = value_double(u, "velocity_cms", c.pu_velocity_cms);
            if (c.pu_length_cm <= 0.0 || c.pu_mass_g <= 0.0 || c.pu_velocity_cms <= 0.0)
                throw ConfigError("phantom_units.length_cm / mass_g / velocity_cms must all be > 0");
        }
        c.allow_missing_velocity = value_bool(j, "allow_missing_velocity", false);
        if (contains(j, "polaris_writer")) {
            const Json& pw = at(j, "polaris_writer");
            c.pw_sim_id = static_cast<unsigned short>(value_int(pw, "sim_id", 50));
            c.pw_length_scale_m = value_double(pw, "length_scale_m", 3.08567758e16);
            c.pw_velocity_scale = value_double(pw, "velocity_scale", 1000.0);
            c.pw_density_scale =
                value_double(pw, "density_scale", 1.9891e30 / (3.08567758e16 * 3.08567758e16 * 3.08567758e16));
            c.pw_temperature = value_double(pw, "gas_temperature_k", 65.0);
        }
    } catch (const OutOfRange& e) {
        throw ConfigError(std::string("config missing required field: ") + e.what());
    }
    return c;
}

Json cfg_json(const Cfg& c) {
    Json j = Json::obj();
    j.o["run_name"] = Json::str(c.run_name);
    Json d = Json::obj();
    static const char* dk[6] = {"x_min", "x_max", "y_min", "y_max", "z_min", "z_max"};
    for (int k = 0; k < 6; ++k) d.o[dk[k]] = Json::real(c.dom[k]);
    j.o["domain"] = d;
    Json gb = Json::obj();
    gb.o["block_scale_factor"] = Json::real(c.block_scale_factor);
    j.o["grid_blocks"] = gb;
    j.o["input_snapshot"] = Json::str(c.input_snapshot);
    j.o["input_snapshot_format"] = Json::str(c.input_snapshot_format);
    j.o["work_folder"] = Json::str(c.work_folder);
    Json b = Json::obj();
    b.o["enabled"] = Json::boolean(c.buffer_enabled);
    b.o["strategy"] = Json::str(c.buffer_strategy);
    if (c.buffer_strategy == "graded_shells") {
        Json g = Json::obj();
        g.o["dx"] = Json::real(c.gs_dx);
        g.o["dy"] = Json::real(c.gs_dy);
        g.o["dz"] = Json::real(c.gs_dz);
        g.o["shell_spacing_1"] = Json::real(c.gs_s1);
        g.o["shell_spacing_2"] = Json::real(c.gs_s2);
        g.o["shell_spacing_3"] = Json::real(c.gs_s3);
        b.o["graded_shells"] = g;
    } else if (c.buffer_strategy == "coverage_grid") {
        Json g = Json::obj();
        g.o["spacing_fraction"] = Json::real(c.cg_spacing_fraction);
        g.o["coverage_radius_factor"] = Json::real(c.cg_radius_factor);
        b.o["coverage_grid"] = g;
    }
    j.o["buffer"] = b;
    Json l = Json::obj();
    l.o["iterations"] = Json::integer(c.lloyd_iterations);
    l.o["save_each_iteration"] = Json::boolean(c.lloyd_save_each);
    j.o["lloyd"] = l;
    Json m = Json::obj();
    m.o["use_tabulated_integral"] = Json::boolean(c.mapping_tab);
    m.o["density_table"] = Json::str(c.mapping_table);
    j.o["mapping"] = m;
    Json pw = Json::obj();
    pw.o["sim_id"] = Json::unsig(c.pw_sim_id);
    pw.o["length_scale_m"] = Json::real(c.pw_length_scale_m);
    pw.o["velocity_scale"] = Json::real(c.pw_velocity_scale);
    pw.o["density_scale"] = Json::real(c.pw_density_scale);
    pw.o["gas_temperature_k"] = Json::real(c.pw_temperature);
    j.o["polaris_writer"] = pw;
    Json dg = Json::obj();
    dg.o["enabled"] = Json::boolean(c.diag_enabled);
    dg.o["python_executable"] = Json::str(c.diag_python);
    j.o["diagnostics"] = dg;
    Json cm = Json::obj();
    cm.o["resolution"] = Json::integer(c.cmap_resolution);
    j.o["cartesian_map"] = cm;
    Json cg = Json::obj();
    cg.o["resolution"] = Json::integer(c.cg_res);
    cg.o["resolution_x"] = Json::integer(c.cg_rx);
    cg.o["resolution_y"] = Json::integer(c.cg_ry);
    cg.o["resolution_z"] = Json::integer(c.cg_rz);
    cg.o["use_tabulated_integral"] = Json::boolean(c.cg_tab);
    cg.o["density_table"] = Json::str(c.cg_table);
    j.o["cartesian_grid"] = cg;
    if (!kernel_is_default(c)) {
        Json k = Json::obj();
        k.o["name"] = Json::str(c.kernel_name);
        k.o["support_radius_in_h"] = Json::real(c.kernel_support);
        k.o["hfact"] = Json::real(c.kernel_hfact);
        k.o["generic_path"] = Json::boolean(c.kernel_generic);
        j.o["kernel"] = k;
    }
    if (c.allow_missing_velocity) j.o["allow_missing_velocity"] = Json::boolean(true);
    if (!units_default(c)) {
        Json u = Json::obj();
        u.o["length_cm"] = Json::real(c.pu_length_cm);
        u.o["mass_g"] = Json::real(c.pu_mass_g);
        u.o["velocity_cms"] = Json::real(c.pu_velocity_cms);
        j.o["phantom_units"] = u;
    }
    return j;
}

// directory components and the last extension stripped (src/config.cpp:13-20)
std::string trimmed_basename(const std::string& path) {
    std::string base = path;
    const auto slash = base.find_last_of('/');
    if (slash != std::string::npos) base = base.substr(slash + 1);
    const auto dot = base.find_last_of('.');
    if (dot != std::string::npos) base = base.substr(0, dot);
    return base;
}

void mkdir_p(const std::string& path) {
    std::string partial;
    for (size_t k = 0; k < path.size(); ++k) {
        if (path[k] == '/' && !partial.empty()) mkdir(partial.c_str(), 0755);
        partial += path[k];
    }
    mkdir(partial.c_str(), 0755);
}

std::string run_dir(const Cfg& c) { return c.work_folder + "/" + trimmed_basename(c.input_snapshot); }

std::string stage_path(const Cfg& c, const std::string& name) {
    const std::string d = run_dir(c);
    mkdir_p(d);
    return d + "/" + name;
}

// ------------------------------------------------------------------------------------------------
// The manifest (src/manifest.cpp:36-199)
// ------------------------------------------------------------------------------------------------
constexpr size_t kWholeFileLimit = 64ull << 20;  // files up to this size are hashed whole
constexpr size_t kEndSample = 32ull << 20;       // larger ones: this much from each end, and the size

// 64-bit FNV-1a with the prime 1099511628211 and the reference's offset basis 1469598103934665603
// (src/manifest.cpp:53; the published basis with its last digit dropped), kept so that fingerprints
// equal the reference's
uint64_t fnv1a(uint64_t h, const char* p, size_t n) {
    for (size_t k = 0; k < n; ++k) {
        h ^= static_cast<unsigned char>(p[k]);
        h *= 1099511628211ull;
    }
    return h;
}

Json fingerprint(const std::string& path) {
    Json j = Json::obj();
    struct stat st;
    if (path.empty() || stat(path.c_str(), &st) != 0) {
        j.o["exists"] = Json::boolean(false);
        return j;
    }
    j.o["exists"] = Json::boolean(true);
    const size_t size = static_cast<size_t>(st.st_size);
    j.o["size_bytes"] = Json::unsig(size);
    j.o["mtime_unix"] = Json::integer(static_cast<long>(st.st_mtime));
    std::ifstream f(path, std::ios::binary);
    if (!f) return j;  // metadata without a hash when the file cannot be read
    uint64_t h = 1469598103934665603ull;
    std::vector<char> buf(1 << 20);
    auto hash_run = [&](size_t want) {
        while (want > 0 && f) {
            f.read(buf.data(), static_cast<std::streamsize>(std::min(want, buf.size())));
            const std::streamsize got = f.gcount();
            if (got <= 0) break;
            h = fnv1a(h, buf.data(), static_cast<size_t>(got));
            want -= static_cast<size_t>(got);
        }
    };
    if (size <= kWholeFileLimit) {
        hash_run(SIZE_MAX);
    } else {
        hash_run(kEndSample);
        f.clear();
        f.seekg(static_cast<std::streamoff>(size - kEndSample), std::ios::beg);
        hash_run(kEndSample);
        h = fnv1a(h, reinterpret_cast<const char*>(&size), sizeof(size));  // the size's bytes, host order
    }
    char hex[32];
    std::snprintf(hex, sizeof(hex), "%" PRIx64, h);
    j.o["fingerprint"] = Json::str(hex);
    return j;
}

Json manifest(const Cfg& c) {
    Json j = Json::obj();
    j.o["manifest_schema_version"] = Json::integer(1);
    j.o["written_at_unix"] = Json::integer(static_cast<long>(std::time(nullptr)));
    j.o["config"] = cfg_json(c);
    Json in = Json::obj();
    in.o["input_snapshot"] = fingerprint(c.input_snapshot);
    if (c.mapping_tab && !c.mapping_table.empty()) in.o["density_table"] = fingerprint(c.mapping_table);
    if (c.cg_tab && !c.cg_table.empty()) in.o["cartesian_grid_density_table"] = fingerprint(c.cg_table);
    j.o["inputs"] = in;
    Json out = Json::obj();
    static const char* names[] = {"points-final.dat", "tessellation.dat", "mapping.dat", "grid-polaris.dat",
                                  "grid-voronoi.txt", "cartesian_map.dat", "cartesian_grid.dat",
                                  "buffer_diagnostic.png", "lloyd_diagnostic.png"};
    for (const char* n : names) {
        Json fp = fingerprint(stage_path(c, n));
        if (fp.o["exists"].b) out.o[n] = fp;
    }
    j.o["outputs"] = out;
    return j;
}

thread_local std::string g_cfg_err;

void copy_str(char* dst, size_t cap, const std::string& s) {
    const size_t n = std::min(s.size(), cap - 1);
    std::memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

int set_err(const std::string& m) {
    g_cfg_err = m;
    return KIWI_IO_ERR_FORMAT;
}

}  // namespace

extern "C" {

const char* kiwi_config_last_error(void) { return g_cfg_err.c_str(); }

int kiwi_config_load(const char* path, kiwi_config* out) {
    g_cfg_err.clear();
    if (!path || !out) return set_err("null argument"), KIWI_IO_ERR_INVALID_ARG;
    try {
        const Cfg c = load_config(path);
        std::memset(out, 0, sizeof(*out));
        copy_str(out->run_name, sizeof(out->run_name), c.run_name);
        copy_str(out->input_snapshot, sizeof(out->input_snapshot), c.input_snapshot);
        copy_str(out->input_snapshot_format, sizeof(out->input_snapshot_format), c.input_snapshot_format);
        copy_str(out->work_folder, sizeof(out->work_folder), c.work_folder);
        copy_str(out->run_dir, sizeof(out->run_dir), run_dir(c));
        for (int k = 0; k < 6; ++k) out->domain[k] = c.dom[k];
        out->kernel_id = kernel_id_of(c.kernel_name);
        out->support_in_h = support_in_h(c);
        out->hfact = c.kernel_hfact > 0.0 ? c.kernel_hfact : kTables[out->kernel_id].hfact;
        out->generic_path = generic_path(c) ? 1 : 0;
        out->grid_nx = c.cg_rx > 0 ? c.cg_rx : c.cg_res;
        out->grid_ny = c.cg_ry > 0 ? c.cg_ry : c.cg_res;
        out->grid_nz = c.cg_rz > 0 ? c.cg_rz : c.cg_res;
        out->grid_tabulated = c.cg_tab;
        out->mapping_tabulated = c.mapping_tab;
        out->buffer_enabled = c.buffer_enabled;
        copy_str(out->buffer_strategy, sizeof(out->buffer_strategy), c.buffer_strategy);
        out->coverage_spacing_fraction = c.cg_spacing_fraction;
        out->coverage_radius_factor = c.cg_radius_factor;
        out->lloyd_iterations = c.lloyd_iterations;
        out->cartesian_map_resolution = c.cmap_resolution;
        out->block_scale_factor = c.block_scale_factor;
        out->phantom_units.length_cm = c.pu_length_cm;
        out->phantom_units.mass_g = c.pu_mass_g;
        out->phantom_units.velocity_cms = c.pu_velocity_cms;
        out->allow_missing_velocity = c.allow_missing_velocity;
        return KIWI_IO_OK;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int kiwi_config_json(const char* path, char* buf, int64_t capacity, int64_t* length) {
    g_cfg_err.clear();
    if (!path) return set_err("null argument"), KIWI_IO_ERR_INVALID_ARG;
    try {
        const std::string s = dumps(cfg_json(load_config(path)), -1);
        if (length) *length = static_cast<int64_t>(s.size());
        if (buf && capacity > 0) copy_str(buf, static_cast<size_t>(capacity), s);
        return KIWI_IO_OK;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int kiwi_manifest_write(const char* config_path) {
    g_cfg_err.clear();
    if (!config_path) return set_err("null argument"), KIWI_IO_ERR_INVALID_ARG;
    try {
        const Cfg c = load_config(config_path);
        const std::string path = stage_path(c, "manifest.json");
        std::ofstream f(path);
        if (!f) {
            std::cerr << "[kiwi] warning: could not write manifest to " << path << "\n";
            return KIWI_IO_OK;  // a warning, as in the reference
        }
        f << dumps(manifest(c), 2) << "\n";
        return KIWI_IO_OK;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int kiwi_manifest_check(const char* config_path) {
    g_cfg_err.clear();
    if (!config_path) return set_err("null argument"), -static_cast<int>(KIWI_IO_ERR_INVALID_ARG);
    try {
        const Cfg c = load_config(config_path);
        const std::string path = stage_path(c, "manifest.json");
        bool ok = false;
        const std::string text = read_file(path, &ok);
        if (!ok) return 0;  // no prior manifest: the first run in this directory
        Json prior;
        try {
            Parser ps(text);
            prior = ps.parse_document();
        } catch (const std::runtime_error&) {
            std::cerr << "[kiwi] warning: existing manifest.json is not valid JSON, ignoring: " << path << "\n";
            return 0;
        }
        if (!contains(prior, "config")) return 0;
        const Json& old_cfg = at(prior, "config");
        const Json cur = cfg_json(c);
        std::vector<std::string> changed;
        for (const auto& kv : cur.o)
            if (!contains(old_cfg, kv.first) || !equal(at(old_cfg, kv.first), kv.second)) changed.push_back(kv.first);
        if (changed.empty()) return 0;
        std::cerr << "[kiwi] WARNING: this run directory (" << run_dir(c) << ") was previously "
                  << "used with a DIFFERENT config. Cached stage output here may have been computed "
                  << "under the old settings. Changed section(s): ";
        for (size_t k = 0; k < changed.size(); ++k) std::cerr << changed[k] << (k + 1 < changed.size() ? ", " : "");
        std::cerr << "\n[kiwi]   old: " << dumps(old_cfg, -1) << "\n";
        std::cerr << "[kiwi]   new: " << dumps(cur, -1) << "\n";
        std::cerr << "[kiwi] Delete the stale cached file(s) under that directory (or use a "
                     "different work_folder/run_name) if you intended this change to take effect.\n";
        return 1;
    } catch (const std::exception& e) {
        set_err(e.what());
        return -static_cast<int>(KIWI_IO_ERR_FORMAT);
    }
}

}  // extern "C"
