// TOML reader for scene files (see toml.h).
#include "toml.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace arp_toml {

const char* kind_name(Kind k) {
    switch (k) {
        case Kind::String: return "string";
        case Kind::Integer: return "integer";
        case Kind::Float: return "float";
        case Kind::Boolean: return "boolean";
        case Kind::Array: return "array";
        case Kind::Table: return "table";
    }
    return "?";
}

namespace {

struct Fail {};

class Parser {
   public:
    Parser(const char* t, size_t n) : s(t), n(n) {}

    VPtr run(std::string& err) {
        root = std::make_shared<Value>();
        root->defined = true;
        cur = root.get();
        try {
            while (true) {
                skip_ws_nl_comments();
                if (eof()) break;
                if (peek() == '[') {
                    header();
                } else {
                    keyval(cur, false);
                }
                eol();
            }
        } catch (Fail&) {
            err = msg;
            return nullptr;
        }
        return root;
    }

   private:
    const char* s;
    size_t n;
    size_t p = 0;
    int line = 1;
    std::string msg;
    VPtr root;
    Value* cur = nullptr;

    [[noreturn]] void fail(const std::string& m) {
        msg = "TOML line " + std::to_string(line) + ": " + m;
        throw Fail{};
    }
    bool eof() const { return p >= n; }
    char peek(size_t k = 0) const { return p + k < n ? s[p + k] : '\0'; }
    char get() {
        char c = s[p++];
        if (c == '\n') line++;
        return c;
    }
    bool starts(const char* lit) const {
        size_t m = std::strlen(lit);
        return p + m <= n && std::memcmp(s + p, lit, m) == 0;
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t')) p++;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') {
                unsigned char c = (unsigned char)peek();
                if (c < 0x20 && c != '\t' && c != '\r') fail("control character in a comment");
                p++;
            }
    }
    bool newline() {
        if (peek() == '\n') {
            get();
            return true;
        }
        if (peek() == '\r' && peek(1) == '\n') {
            p++;
            get();
            return true;
        }
        return false;
    }
    void skip_ws_nl_comments() {
        while (true) {
            skip_ws();
            skip_comment();
            if (!newline()) break;
        }
    }
    void eol() {
        skip_ws();
        skip_comment();
        if (eof()) return;
        if (!newline()) fail("expected the end of the line");
    }

    static bool bare_char(char c) {
        return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' ||
               c == '-';
    }

    std::vector<std::string> key() {
        std::vector<std::string> parts;
        while (true) {
            skip_ws();
            if (peek() == '"') {
                if (starts("\"\"\"")) fail("multi-line string used as a key");
                parts.push_back(basic_string());
            } else if (peek() == '\'') {
                if (starts("'''")) fail("multi-line string used as a key");
                parts.push_back(literal_string());
            } else {
                size_t b = p;
                while (!eof() && bare_char(peek())) p++;
                if (p == b) fail("expected a key");
                parts.emplace_back(s + b, p - b);
            }
            skip_ws();
            if (peek() != '.') break;
            p++;
        }
        return parts;
    }

    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }

    void escape(std::string& o) {
        char c = get();
        switch (c) {
            case 'b': o += '\b'; return;
            case 't': o += '\t'; return;
            case 'n': o += '\n'; return;
            case 'f': o += '\f'; return;
            case 'r': o += '\r'; return;
            case '"': o += '"'; return;
            case '\\': o += '\\'; return;
            case 'u':
            case 'U': {
                int nd = c == 'u' ? 4 : 8;
                uint32_t cp = 0;
                for (int k = 0; k < nd; k++) {
                    char h = eof() ? '\0' : get();
                    int v = (h >= '0' && h <= '9') ? h - '0'
                            : (h >= 'a' && h <= 'f') ? h - 'a' + 10
                            : (h >= 'A' && h <= 'F') ? h - 'A' + 10
                                                     : -1;
                    if (v < 0) fail("bad unicode escape");
                    cp = cp * 16 + (uint32_t)v;
                }
                if (cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) fail("unicode escape is not a scalar value");
                put_utf8(o, cp);
                return;
            }
            default: fail(std::string("bad escape \\") + c);
        }
    }

    static bool bad_ctrl(char c) {
        unsigned char u = (unsigned char)c;
        return (u < 0x20 && c != '\t') || u == 0x7F;
    }

    std::string basic_string() {
        std::string o;
        if (starts("\"\"\"")) {
            p += 3;
            newline();  // a newline right after the opening is trimmed
            while (true) {
                if (eof()) fail("unterminated multi-line string");
                if (starts("\"\"\"")) {
                    p += 3;
                    // up to two more quotes belong to the content
                    int extra = 0;
                    while (peek() == '"' && extra < 2) {
                        o += '"';
                        p++;
                        extra++;
                    }
                    return o;
                }
                char c = peek();
                if (c == '\\') {
                    p++;
                    // line-ending backslash: trim the newline and the whitespace after it
                    size_t q = p;
                    while (q < n && (s[q] == ' ' || s[q] == '\t')) q++;
                    if (q < n && (s[q] == '\n' || (s[q] == '\r' && q + 1 < n && s[q + 1] == '\n'))) {
                        p = q;
                        while (true) {
                            skip_ws();
                            if (!newline()) break;
                        }
                        continue;
                    }
                    escape(o);
                } else if (newline()) {
                    o += '\n';
                } else {
                    if (bad_ctrl(c)) fail("control character in a string");
                    o += c;
                    p++;
                }
            }
        }
        p++;
        while (true) {
            if (eof() || peek() == '\n') fail("unterminated string");
            char c = get();
            if (c == '"') return o;
            if (c == '\\') {
                escape(o);
            } else {
                if (bad_ctrl(c)) fail("control character in a string");
                o += c;
            }
        }
    }

    std::string literal_string() {
        std::string o;
        if (starts("'''")) {
            p += 3;
            newline();
            while (true) {
                if (eof()) fail("unterminated multi-line literal string");
                if (starts("'''")) {
                    p += 3;
                    int extra = 0;
                    while (peek() == '\'' && extra < 2) {
                        o += '\'';
                        p++;
                        extra++;
                    }
                    return o;
                }
                if (newline()) {
                    o += '\n';
                } else {
                    char c = get();
                    if (bad_ctrl(c)) fail("control character in a string");
                    o += c;
                }
            }
        }
        p++;
        while (true) {
            if (eof() || peek() == '\n') fail("unterminated literal string");
            char c = get();
            if (c == '\'') return o;
            if (bad_ctrl(c)) fail("control character in a string");
            o += c;
        }
    }

    // digits with single underscores between digits
    static bool digits_ok(const std::string& t, bool (*isd)(char)) {
        if (t.empty()) return false;
        bool prev_digit = false;
        for (char c : t) {
            if (c == '_') {
                if (!prev_digit) return false;
                prev_digit = false;
            } else if (isd(c)) {
                prev_digit = true;
            } else {
                return false;
            }
        }
        return prev_digit;
    }
    static bool isdec(char c) { return c >= '0' && c <= '9'; }
    static bool ishex(char c) { return isdec(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); }
    static bool isoct(char c) { return c >= '0' && c <= '7'; }
    static bool isbin(char c) { return c == '0' || c == '1'; }
    static std::string strip(const std::string& t) {
        std::string o;
        for (char c : t)
            if (c != '_') o += c;
        return o;
    }

    void number(Value& v) {
        size_t b = p;
        while (!eof()) {
            char c = peek();
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == ',' || c == ']' || c == '}' || c == '#') break;
            p++;
        }
        std::string t(s + b, p - b);
        if (t.empty()) fail("expected a value");
        std::string body = t;
        char sign = 0;
        if (body[0] == '+' || body[0] == '-') {
            sign = body[0];
            body = body.substr(1);
        }
        if (body == "inf" || body == "nan") {
            v.kind = Kind::Float;
            v.f = body == "inf" ? HUGE_VAL : std::nan("");
            if (sign == '-') v.f = -v.f;
            return;
        }
        if (body.find(':') != std::string::npos ||
            (body.size() > 4 && body[4] == '-' && isdec(body[0]) && isdec(body[1]) && isdec(body[2]) && isdec(body[3])))
            fail("dates and times are not supported");
        if (body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b')) {
            if (sign) fail("a sign before a hexadecimal, octal or binary integer");
            std::string d = body.substr(2);
            int base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
            bool ok = base == 16 ? digits_ok(d, ishex) : base == 8 ? digits_ok(d, isoct) : digits_ok(d, isbin);
            if (!ok) fail("malformed integer '" + t + "'");
            errno = 0;
            char* e = nullptr;
            std::string c = strip(d);
            unsigned long long u = std::strtoull(c.c_str(), &e, base);
            if (errno == ERANGE || u > (unsigned long long)INT64_MAX) fail("integer out of range '" + t + "'");
            v.kind = Kind::Integer;
            v.i = (int64_t)u;
            return;
        }
        size_t ep = body.find_first_of("eE");
        size_t dp = body.find('.');
        if (ep != std::string::npos || dp != std::string::npos) {
            // float: int part [. frac] [e exp]
            std::string ip = body.substr(0, std::min(dp, ep));
            std::string frac, ex;
            if (dp != std::string::npos) {
                if (ep != std::string::npos && ep < dp) fail("malformed float '" + t + "'");
                frac = body.substr(dp + 1, (ep == std::string::npos ? body.size() : ep) - dp - 1);
                if (!digits_ok(frac, isdec)) fail("malformed float '" + t + "'");
            }
            if (ep != std::string::npos) {
                ex = body.substr(ep + 1);
                std::string exd = ex;
                if (!exd.empty() && (exd[0] == '+' || exd[0] == '-')) exd = exd.substr(1);
                if (!digits_ok(exd, isdec)) fail("malformed float '" + t + "'");
            }
            if (!digits_ok(ip, isdec) || (ip.size() > 1 && ip[0] == '0')) fail("malformed float '" + t + "'");
            std::string c = (sign ? std::string(1, sign) : std::string()) + strip(body);
            errno = 0;
            char* e = nullptr;
            double f = std::strtod(c.c_str(), &e);
            if (*e != '\0') fail("malformed float '" + t + "'");
            v.kind = Kind::Float;
            v.f = f;
            return;
        }
        if (!digits_ok(body, isdec) || (body.size() > 1 && body[0] == '0'))
            fail("malformed value '" + t + "'");
        std::string c = (sign ? std::string(1, sign) : std::string()) + strip(body);
        errno = 0;
        char* e = nullptr;
        long long x = std::strtoll(c.c_str(), &e, 10);
        if (errno == ERANGE) fail("integer out of range '" + t + "'");
        v.kind = Kind::Integer;
        v.i = (int64_t)x;
    }

    int depth = 0;  // nesting of arrays and inline tables (bounded: the parser recurses)
    struct Nest {
        Parser& p;
        explicit Nest(Parser& q) : p(q) {
            if (++p.depth > 128) p.fail("arrays or inline tables nested more than 128 deep");
        }
        ~Nest() { --p.depth; }
    };

    VPtr value() {
        Nest nest(*this);
        auto v = std::make_shared<Value>();
        v->defined = true;
        char c = peek();
        if (c == '"') {
            v->kind = Kind::String;
            v->str = basic_string();
        } else if (c == '\'') {
            v->kind = Kind::String;
            v->str = literal_string();
        } else if (c == '[') {
            p++;
            v->kind = Kind::Array;
            v->closed = true;
            while (true) {
                skip_ws_nl_comments();
                if (peek() == ']') {
                    p++;
                    break;
                }
                v->arr.push_back(value());
                skip_ws_nl_comments();
                if (peek() == ',') {
                    p++;
                    continue;
                }
                if (peek() == ']') {
                    p++;
                    break;
                }
                fail("expected ',' or ']' in an array");
            }
        } else if (c == '{') {
            p++;
            v->kind = Kind::Table;
            skip_ws();
            if (peek() == '}') {
                p++;
            } else {
                while (true) {
                    keyval(v.get(), true);
                    skip_ws();
                    if (peek() == ',') {
                        p++;
                        continue;
                    }
                    if (peek() == '}') {
                        p++;
                        break;
                    }
                    fail("expected ',' or '}' in an inline table");
                }
            }
            close_all(v.get());
        } else if (starts("true") && !bare_char(peek(4))) {
            p += 4;
            v->kind = Kind::Boolean;
            v->b = true;
        } else if (starts("false") && !bare_char(peek(5))) {
            p += 5;
            v->kind = Kind::Boolean;
            v->b = false;
        } else {
            number(*v);
        }
        return v;
    }

    static void close_all(Value* t) {
        t->closed = true;
        for (auto& kv : t->tab)
            if (kv.second->kind == Kind::Table) close_all(kv.second.get());
    }

    void keyval(Value* table, bool /*in_inline*/) {
        std::vector<std::string> k = key();
        skip_ws();
        if (peek() != '=') fail("expected '=' after a key");
        p++;
        skip_ws();
        if (eof() || peek() == '\n' || peek() == '#') fail("missing value");
        Value* t = table;
        for (size_t q = 0; q + 1 < k.size(); q++) {
            auto it = t->tab.find(k[q]);
            if (it == t->tab.end()) {
                auto nt = std::make_shared<Value>();
                nt->kind = Kind::Table;
                nt->dotted = true;
                nt->defined = true;
                t->tab[k[q]] = nt;
                t = nt.get();
            } else {
                Value* x = it->second.get();
                if (x->kind != Kind::Table || x->closed || !x->dotted) fail("key '" + k[q] + "' is already defined");
                t = x;
            }
        }
        if (t->tab.count(k.back())) fail("duplicate key '" + k.back() + "'");
        t->tab[k.back()] = value();
    }

    void header() {
        bool aot = starts("[[");
        p += aot ? 2 : 1;
        std::vector<std::string> k = key();
        if (aot) {
            if (!starts("]]")) fail("expected ']]'");
            p += 2;
        } else {
            if (peek() != ']') fail("expected ']'");
            p++;
        }
        Value* t = root.get();
        for (size_t q = 0; q + 1 < k.size(); q++) {
            auto it = t->tab.find(k[q]);
            if (it == t->tab.end()) {
                auto nt = std::make_shared<Value>();
                nt->kind = Kind::Table;
                t->tab[k[q]] = nt;
                t = nt.get();
            } else {
                Value* x = it->second.get();
                if (x->kind == Kind::Array && x->aot) {
                    t = x->arr.back().get();
                } else if (x->kind == Kind::Table && !x->closed) {
                    t = x;
                } else {
                    fail("'" + k[q] + "' is not a table");
                }
            }
        }
        const std::string& last = k.back();
        auto it = t->tab.find(last);
        if (aot) {
            Value* arr;
            if (it == t->tab.end()) {
                auto na = std::make_shared<Value>();
                na->kind = Kind::Array;
                na->aot = true;
                na->defined = true;
                t->tab[last] = na;
                arr = na.get();
            } else {
                arr = it->second.get();
                if (arr->kind != Kind::Array || !arr->aot) fail("'" + last + "' is not an array of tables");
            }
            auto nt = std::make_shared<Value>();
            nt->kind = Kind::Table;
            nt->defined = true;
            arr->arr.push_back(nt);
            cur = nt.get();
            return;
        }
        if (it == t->tab.end()) {
            auto nt = std::make_shared<Value>();
            nt->kind = Kind::Table;
            nt->defined = true;
            t->tab[last] = nt;
            cur = nt.get();
            return;
        }
        Value* x = it->second.get();
        if (x->kind != Kind::Table || x->closed || x->defined) fail("table '" + last + "' is already defined");
        x->defined = true;
        cur = x;
    }
};

void json_str(std::string& o, const std::string& s) {
    o += '"';
    for (unsigned char c : s) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            o += b;
        } else {
            o += (char)c;
        }
    }
    o += '"';
}

void json(std::string& o, const Value& v) {
    switch (v.kind) {
        case Kind::String: json_str(o, v.str); break;
        case Kind::Integer: o += std::to_string(v.i); break;
        case Kind::Float: {
            if (std::isnan(v.f)) {
                o += "NaN";
            } else if (std::isinf(v.f)) {
                o += v.f > 0 ? "Infinity" : "-Infinity";
            } else {
                char b[40];
                std::snprintf(b, sizeof b, "%.17g", v.f);
                std::string t(b);
                if (t.find_first_of(".eE") == std::string::npos) t += ".0";
                o += t;
            }
            break;
        }
        case Kind::Boolean: o += v.b ? "true" : "false"; break;
        case Kind::Array: {
            o += '[';
            for (size_t k = 0; k < v.arr.size(); k++) {
                if (k) o += ',';
                json(o, *v.arr[k]);
            }
            o += ']';
            break;
        }
        case Kind::Table: {
            o += '{';
            bool first = true;
            for (auto& kv : v.tab) {
                if (!first) o += ',';
                first = false;
                json_str(o, kv.first);
                o += ':';
                json(o, *kv.second);
            }
            o += '}';
            break;
        }
    }
}

}  // namespace

VPtr parse(const char* text, size_t len, std::string& err) {
    Parser ps(text, len);
    return ps.run(err);
}

std::string to_json(const Value& v) {
    std::string o;
    json(o, v);
    return o;
}

}  // namespace arp_toml
