#include "toml.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "common.hpp"

namespace mlr::toml {

const Value* Value::get(const std::string& key) const {
    if (kind != Table) return nullptr;
    for (auto& kv : tab)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
Value* Value::get_mut(const std::string& key) {
    if (kind != Table) return nullptr;
    for (auto& kv : tab)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
Value& Value::set(const std::string& key, Value v) {
    if (Value* e = get_mut(key)) { *e = std::move(v); return *e; }
    tab.emplace_back(key, std::move(v));
    return tab.back().second;
}
double Value::as_float() const {
    if (kind == Float) return f;
    if (kind == Int) return (double)i;
    fail("TOML value is not a number");
}
int64_t Value::as_int() const {
    if (kind == Int) return i;
    fail("TOML value is not an integer");
}
bool Value::as_bool() const {
    if (kind == Bool) return b;
    fail("TOML value is not a boolean");
}
const std::string& Value::as_string() const {
    if (kind == String) return s;
    fail("TOML value is not a string");
}

std::string Value::dump(int ind) const {
    std::ostringstream o;
    o.precision(17);
    const std::string pad(ind, ' ');
    switch (kind) {
    case Null: o << "null"; break;
    case Bool: o << (b ? "true" : "false"); break;
    case Int: o << i; break;
    case Float: o << f; break;
    case String: o << '"' << s << '"'; break;
    case Array: {
        o << "[";
        for (size_t k = 0; k < arr.size(); ++k) o << (k ? ", " : "") << arr[k].dump(ind + 2);
        o << "]";
        break;
    }
    case Table: {
        o << "{\n";
        for (auto& kv : tab) o << pad << "  " << kv.first << " = " << kv.second.dump(ind + 2) << "\n";
        o << pad << "}";
        break;
    }
    }
    return o.str();
}

namespace {

struct Parser {
    const std::string& t;
    std::string origin;
    size_t p = 0;
    int line = 1;

    [[noreturn]] void err(const std::string& m) {
        fail("TOML parse error (" + origin + ":" + std::to_string(line) + "): " + m);
    }
    bool eof() const { return p >= t.size(); }
    char peek(size_t o = 0) const { return p + o < t.size() ? t[p + o] : '\0'; }
    char get() {
        char c = t[p++];
        if (c == '\n') ++line;
        return c;
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t')) get();
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') get();
    }
    // whitespace, newlines and comments (inside arrays and between lines)
    void skip_all() {
        for (;;) {
            skip_ws();
            if (peek() == '#') { skip_comment(); continue; }
            if (peek() == '\n' || peek() == '\r') { get(); continue; }
            break;
        }
    }
    void expect_eol() {
        skip_ws();
        skip_comment();
        if (peek() == '\r') get();
        if (!eof() && peek() != '\n') err(std::string("unexpected character '") + peek() + "'");
        if (!eof()) get();
    }
    static bool bare(char c) {
        return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' ||
               c == '-';
    }
    std::string key_part() {
        skip_ws();
        if (peek() == '"') return basic_string();
        if (peek() == '\'') return literal_string();
        std::string k;
        while (!eof() && bare(peek())) k += get();
        if (k.empty()) err("expected a key");
        return k;
    }
    std::vector<std::string> dotted_key() {
        std::vector<std::string> ks{key_part()};
        for (;;) {
            skip_ws();
            if (peek() != '.') break;
            get();
            ks.push_back(key_part());
        }
        return ks;
    }
    void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    void escape(std::string& out) {
        char c = get();
        switch (c) {
        case 'b': out += '\b'; break;
        case 't': out += '\t'; break;
        case 'n': out += '\n'; break;
        case 'f': out += '\f'; break;
        case 'r': out += '\r'; break;
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case 'u':
        case 'U': {
            const int n = c == 'u' ? 4 : 8;
            uint32_t cp = 0;
            for (int k = 0; k < n; ++k) {
                char h = get();
                cp = cp * 16 + (uint32_t)(std::isdigit((unsigned char)h) ? h - '0' : (std::tolower(h) - 'a' + 10));
            }
            utf8(out, cp);
            break;
        }
        default: err(std::string("bad escape \\") + c);
        }
    }
    std::string basic_string() {
        if (peek() == '"' && peek(1) == '"' && peek(2) == '"') return ml_basic_string();
        get();
        std::string s;
        for (;;) {
            if (eof() || peek() == '\n') err("unterminated string");
            char c = get();
            if (c == '"') break;
            if (c == '\\') escape(s);
            else s += c;
        }
        return s;
    }
    std::string ml_basic_string() {
        p += 3;
        if (peek() == '\r') get();
        if (peek() == '\n') get();  // newline right after the delimiter is trimmed
        std::string s;
        for (;;) {
            if (eof()) err("unterminated multi-line string");
            if (peek() == '"' && peek(1) == '"' && peek(2) == '"') {
                p += 3;
                while (peek() == '"') { s += '"'; ++p; }  // up to two quotes may close the body
                break;
            }
            char c = get();
            if (c == '\\') {
                // line-ending backslash: drop whitespace and newlines that follow
                size_t q = p;
                while (q < t.size() && (t[q] == ' ' || t[q] == '\t')) ++q;
                if (q < t.size() && (t[q] == '\n' || t[q] == '\r')) {
                    while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\n' || peek() == '\r')) get();
                    continue;
                }
                escape(s);
            } else {
                s += c;
            }
        }
        return s;
    }
    std::string literal_string() {
        if (peek(1) == '\'' && peek(2) == '\'') {
            p += 3;
            if (peek() == '\r') get();
            if (peek() == '\n') get();
            std::string s;
            for (;;) {
                if (eof()) err("unterminated literal string");
                if (peek() == '\'' && peek(1) == '\'' && peek(2) == '\'') { p += 3; break; }
                s += get();
            }
            return s;
        }
        get();
        std::string s;
        for (;;) {
            if (eof() || peek() == '\n') err("unterminated literal string");
            char c = get();
            if (c == '\'') break;
            s += c;
        }
        return s;
    }
    Value number_or_bool() {
        std::string w;
        while (!eof() && (bare(peek()) || peek() == '.' || peek() == '+' || peek() == ':')) w += get();
        Value v;
        if (w == "true") { v.kind = Value::Bool; v.b = true; return v; }
        if (w == "false") { v.kind = Value::Bool; v.b = false; return v; }
        std::string c;
        for (char ch : w) if (ch != '_') c += ch;
        if (c.empty()) err("expected a value");
        std::string body = c;
        if (body[0] == '+' || body[0] == '-') body = body.substr(1);
        if (body == "inf" || body == "nan") {
            v.kind = Value::Float;
            v.f = body == "inf" ? INFINITY : NAN;
            if (c[0] == '-') v.f = -v.f;
            return v;
        }
        const bool is_hex = body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b');
        if (is_hex) {
            const int base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
            v.kind = Value::Int;
            v.i = (int64_t)std::strtoull(body.c_str() + 2, nullptr, base);
            return v;
        }
        const bool is_float = c.find_first_of(".eE") != std::string::npos;
        errno = 0;
        char* end = nullptr;
        if (is_float) {
            v.kind = Value::Float;
            v.f = std::strtod(c.c_str(), &end);
        } else {
            v.kind = Value::Int;
            v.i = std::strtoll(c.c_str(), &end, 10);
        }
        if (end == c.c_str() || *end != '\0' || errno == ERANGE) err("bad number '" + w + "'");
        return v;
    }
    Value array() {
        get();
        Value v;
        v.kind = Value::Array;
        for (;;) {
            skip_all();
            if (peek() == ']') { get(); break; }
            v.arr.push_back(value());
            skip_all();
            if (peek() == ',') { get(); continue; }
            if (peek() == ']') { get(); break; }
            err("expected ',' or ']' in array");
        }
        return v;
    }
    Value inline_table() {
        get();
        Value v;
        v.kind = Value::Table;
        skip_ws();
        if (peek() == '}') { get(); return v; }
        for (;;) {
            auto ks = dotted_key();
            skip_ws();
            if (get() != '=') err("expected '=' in inline table");
            skip_ws();
            insert(v, ks, value());
            skip_ws();
            if (peek() == ',') { get(); skip_ws(); continue; }
            if (peek() == '}') { get(); break; }
            err("expected ',' or '}' in inline table");
        }
        return v;
    }
    Value value() {
        skip_ws();
        char c = peek();
        if (c == '"') { Value v; v.kind = Value::String; v.s = basic_string(); return v; }
        if (c == '\'') { Value v; v.kind = Value::String; v.s = literal_string(); return v; }
        if (c == '[') return array();
        if (c == '{') return inline_table();
        return number_or_bool();
    }
    Value& descend(Value& root, const std::vector<std::string>& ks, size_t n) {
        Value* cur = &root;
        for (size_t k = 0; k < n; ++k) {
            Value* nx = cur->get_mut(ks[k]);
            if (!nx) {
                Value t2;
                t2.kind = Value::Table;
                nx = &cur->set(ks[k], t2);
            }
            if (nx->kind == Value::Array && !nx->arr.empty() && nx->arr.back().kind == Value::Table)
                nx = &nx->arr.back();  // dotted path through an array of tables
            if (nx->kind != Value::Table) err("key '" + ks[k] + "' is not a table");
            cur = nx;
        }
        return *cur;
    }
    void insert(Value& tbl, const std::vector<std::string>& ks, Value v) {
        Value& parent = descend(tbl, ks, ks.size() - 1);
        if (parent.contains(ks.back())) err("duplicate key '" + ks.back() + "'");
        parent.set(ks.back(), std::move(v));
    }
    Value document() {
        Value root;
        root.kind = Value::Table;
        Value* cur = &root;
        for (;;) {
            skip_all();
            if (eof()) break;
            if (peek() == '[') {
                const bool aot = peek(1) == '[';
                p += aot ? 2 : 1;
                auto ks = dotted_key();
                skip_ws();
                if (get() != ']' || (aot && get() != ']')) err("bad table header");
                expect_eol();
                if (aot) {
                    Value& parent = descend(root, ks, ks.size() - 1);
                    Value* a = parent.get_mut(ks.back());
                    if (!a) {
                        Value arr;
                        arr.kind = Value::Array;
                        a = &parent.set(ks.back(), arr);
                    }
                    if (a->kind != Value::Array) err("key '" + ks.back() + "' is not an array of tables");
                    Value t2;
                    t2.kind = Value::Table;
                    a->arr.push_back(t2);
                    cur = &a->arr.back();
                } else {
                    cur = &descend(root, ks, ks.size());
                }
                continue;
            }
            auto ks = dotted_key();
            skip_ws();
            if (get() != '=') err("expected '='");
            Value v = value();
            insert(*cur, ks, std::move(v));
            expect_eol();
        }
        return root;
    }
};

}  // namespace

Value parse(const std::string& text, const std::string& origin) {
    Parser ps{text, origin};
    return ps.document();
}

Value parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("cannot open input file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), path);
}

const Value& find(const Value& v, const std::string& k1) {
    const Value* r = v.get(k1);
    if (!r) fail("missing key '" + k1 + "'");
    return *r;
}
const Value& find(const Value& v, const std::string& k1, const std::string& k2) {
    return find(find(v, k1), k2);
}
const Value& find(const Value& v, const std::string& k1, const std::string& k2,
                  const std::string& k3) {
    return find(find(find(v, k1), k2), k3);
}
const Value* lookup(const Value& v, std::initializer_list<const char*> path) {
    const Value* cur = &v;
    for (const char* k : path) {
        cur = cur->get(k);
        if (!cur) return nullptr;
    }
    return cur;
}
double find_or(const Value& v, std::initializer_list<const char*> path, double def) {
    const Value* r = lookup(v, path);
    return r ? r->as_float() : def;
}
int64_t find_or_int(const Value& v, std::initializer_list<const char*> path, int64_t def) {
    const Value* r = lookup(v, path);
    return r ? r->as_int() : def;
}
bool find_or_bool(const Value& v, std::initializer_list<const char*> path, bool def) {
    const Value* r = lookup(v, path);
    return r ? r->as_bool() : def;
}
std::string find_or_str(const Value& v, std::initializer_list<const char*> path,
                        const std::string& def) {
    const Value* r = lookup(v, path);
    return r ? r->as_string() : def;
}

}  // namespace mlr::toml
