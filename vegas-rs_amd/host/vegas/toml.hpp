// A small TOML reader and writer for the simulation input files
// (reference src/input.rs, read with the `toml` crate in src/main.rs:73-85).
//
// Reads TOML 1.0 documents made of: comments; bare, quoted and dotted keys;
// [tables] and [[arrays of tables]]; basic and literal strings (single
// line); integers (decimal, with '_' separators); floats (with exponents,
// inf and nan); booleans; arrays (over several lines) and inline tables.
// Multi-line strings and date-times are not supported and are reported as
// errors with their line number.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <locale>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "error.hpp"

namespace vegas {
namespace toml {

struct Value;
using Table = std::map<std::string, Value>;
using Array = std::vector<Value>;

struct Value {
    enum class Type { String, Integer, Float, Boolean, Array, Table };
    Type type = Type::Table;
    std::string s;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    std::shared_ptr<Array> arr;
    std::shared_ptr<Table> tab;
    bool defined_by_header = false;  // a [table] header made it (may not repeat)
    bool is_table_array = false;     // made by [[header]]

    static Value table() {
        Value v;
        v.type = Type::Table;
        v.tab = std::make_shared<Table>();
        return v;
    }
    static Value array() {
        Value v;
        v.type = Type::Array;
        v.arr = std::make_shared<Array>();
        return v;
    }
    const char* type_name() const {
        switch (type) {
            case Type::String: return "string";
            case Type::Integer: return "integer";
            case Type::Float: return "float";
            case Type::Boolean: return "boolean";
            case Type::Array: return "array";
            default: return "table";
        }
    }
};

class Parser {
  public:
    explicit Parser(const std::string& text) : t_(text) {}

    Table parse() {
        Value root = Value::table();
        Value* cur = &root;
        while (true) {
            skip_ws_comments_newlines();
            if (eof()) break;
            if (peek() == '[') {
                const bool arr = peek(1) == '[';
                pos_ += arr ? 2 : 1;
                skip_ws();
                std::vector<std::string> path = key_path();
                skip_ws();
                expect(']');
                if (arr) expect(']');
                end_of_line();
                cur = arr ? open_table_array(root, path) : open_table(root, path);
            } else {
                std::vector<std::string> path = key_path();
                skip_ws();
                expect('=');
                skip_ws();
                Value v = value();
                assign(*cur, path, std::move(v));
                end_of_line();
            }
        }
        return *root.tab;
    }

  private:
    [[noreturn]] void error(const std::string& what) const {
        int line = 1;
        for (size_t k = 0; k < pos_ && k < t_.size(); ++k) line += t_[k] == '\n';
        throw Error(ErrorKind::InvalidArgument,
                    "TOML parse error at line " + std::to_string(line) + ": " + what);
    }
    bool eof() const { return pos_ >= t_.size(); }
    char peek(size_t k = 0) const { return pos_ + k < t_.size() ? t_[pos_ + k] : '\0'; }
    void expect(char c) {
        if (peek() != c) error(std::string("expected '") + c + "'");
        ++pos_;
    }
    void skip_ws() {
        while (peek() == ' ' || peek() == '\t') ++pos_;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') ++pos_;
    }
    void skip_ws_comments_newlines() {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\r' || peek() == '\n') {
                ++pos_;
                continue;
            }
            break;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (peek() == '\r') ++pos_;
        if (!eof() && peek() != '\n') error("expected the end of the line");
    }

    static bool bare(char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
               c == '_' || c == '-';
    }
    std::string key() {
        if (peek() == '"') return basic_string();
        if (peek() == '\'') return literal_string();
        const size_t b = pos_;
        while (bare(peek())) ++pos_;
        if (pos_ == b) error("expected a key");
        return t_.substr(b, pos_ - b);
    }
    std::vector<std::string> key_path() {
        std::vector<std::string> p{key()};
        for (;;) {
            skip_ws();
            if (peek() != '.') break;
            ++pos_;
            skip_ws();
            p.push_back(key());
        }
        return p;
    }

    std::string basic_string() {
        expect('"');
        if (peek() == '"' && peek(1) == '"') error("multi-line strings are not supported");
        std::string out;
        for (;;) {
            if (eof() || peek() == '\n') error("unterminated string");
            const char c = t_[pos_++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (eof()) error("unterminated string");
            const char e = t_[pos_++];
            switch (e) {
                case 'b': out += '\b'; break;
                case 't': out += '\t'; break;
                case 'n': out += '\n'; break;
                case 'f': out += '\f'; break;
                case 'r': out += '\r'; break;
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case 'u':
                case 'U': {
                    const int n = e == 'u' ? 4 : 8;
                    uint32_t cp = 0;
                    for (int k = 0; k < n; ++k) {
                        const char h = peek();
                        int d = h >= '0' && h <= '9'   ? h - '0'
                                : h >= 'a' && h <= 'f' ? h - 'a' + 10
                                : h >= 'A' && h <= 'F' ? h - 'A' + 10
                                                       : -1;
                        if (d < 0) error("bad unicode escape");
                        cp = cp * 16 + (uint32_t)d;
                        ++pos_;
                    }
                    utf8(cp, out);
                    break;
                }
                default: error(std::string("bad escape \\") + e);
            }
        }
        return out;
    }
    static void utf8(uint32_t cp, std::string& out) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string literal_string() {
        expect('\'');
        if (peek() == '\'' && peek(1) == '\'') error("multi-line strings are not supported");
        const size_t b = pos_;
        while (!eof() && peek() != '\'' && peek() != '\n') ++pos_;
        if (peek() != '\'') error("unterminated string");
        std::string out = t_.substr(b, pos_ - b);
        ++pos_;
        return out;
    }

    Value value() {
        Value v;
        const char c = peek();
        if (c == '"') {
            v.type = Value::Type::String;
            v.s = basic_string();
        } else if (c == '\'') {
            v.type = Value::Type::String;
            v.s = literal_string();
        } else if (c == '[') {
            ++pos_;
            v = Value::array();
            for (;;) {
                skip_ws_comments_newlines();
                if (peek() == ']') {
                    ++pos_;
                    break;
                }
                v.arr->push_back(value());
                skip_ws_comments_newlines();
                if (peek() == ',') {
                    ++pos_;
                    continue;
                }
                skip_ws_comments_newlines();
                expect(']');
                break;
            }
        } else if (c == '{') {
            ++pos_;
            v = Value::table();
            skip_ws();
            if (peek() == '}') {
                ++pos_;
                return v;
            }
            for (;;) {
                skip_ws();
                std::vector<std::string> path = key_path();
                skip_ws();
                expect('=');
                skip_ws();
                assign(v, path, value());
                skip_ws();
                if (peek() == ',') {
                    ++pos_;
                    continue;
                }
                expect('}');
                break;
            }
        } else if (t_.compare(pos_, 4, "true") == 0 && !bare(peek(4))) {
            pos_ += 4;
            v.type = Value::Type::Boolean;
            v.b = true;
        } else if (t_.compare(pos_, 5, "false") == 0 && !bare(peek(5))) {
            pos_ += 5;
            v.type = Value::Type::Boolean;
            v.b = false;
        } else {
            v = number();
        }
        return v;
    }

    Value number() {
        const size_t b = pos_;
        while (!eof()) {
            const char c = peek();
            if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == '_' ||
                c == 'e' || c == 'E' || (c >= 'a' && c <= 'z'))
                ++pos_;
            else
                break;
        }
        std::string tok = t_.substr(b, pos_ - b), clean;
        if (tok.empty()) error("expected a value");
        for (size_t k = 0; k < tok.size(); ++k) {
            if (tok[k] != '_') {
                clean += tok[k];
                continue;
            }
            const bool ok = k > 0 && k + 1 < tok.size() && isdigit((unsigned char)tok[k - 1]) &&
                            isdigit((unsigned char)tok[k + 1]);
            if (!ok) error("misplaced '_' in number " + tok);
        }
        Value v;
        const std::string body = (clean[0] == '+' || clean[0] == '-') ? clean.substr(1) : clean;
        if (body == "inf" || body == "nan") {
            v.type = Value::Type::Float;
            v.f = body == "inf" ? INFINITY : NAN;
            if (clean[0] == '-') v.f = -v.f;
            return v;
        }
        const bool is_float = clean.find_first_of(".eE") != std::string::npos;
        for (char c : body)
            if (!(isdigit((unsigned char)c) || c == '.' || c == 'e' || c == 'E' || c == '+' ||
                  c == '-'))
                error("bad value " + tok);
        if (body.empty() || !isdigit((unsigned char)body[0])) error("bad number " + tok);
        if (!is_float && body.size() > 1 && body[0] == '0') error("leading zero in " + tok);
        std::istringstream in(clean);
        in.imbue(std::locale::classic());
        if (is_float) {
            const size_t dot = clean.find('.');
            if (dot != std::string::npos &&
                (dot + 1 >= clean.size() || !isdigit((unsigned char)clean[dot + 1])))
                error("a float needs digits after '.' in " + tok);
            v.type = Value::Type::Float;
            in >> v.f;
        } else {
            v.type = Value::Type::Integer;
            in >> v.i;
        }
        if (in.fail() || !in.eof()) error("bad number " + tok);
        return v;
    }

    // Walks (creating) the tables of path[0..n-1] below t.
    Value* walk(Value& t, const std::vector<std::string>& path, size_t n, bool headers) {
        Value* cur = &t;
        for (size_t k = 0; k < n; ++k) {
            Table& tab = *cur->tab;
            auto it = tab.find(path[k]);
            if (it == tab.end()) {
                it = tab.emplace(path[k], Value::table()).first;
            } else if (it->second.type == Value::Type::Array && it->second.is_table_array &&
                       headers) {
                cur = &it->second.arr->back();
                continue;
            } else if (it->second.type != Value::Type::Table) {
                error("key `" + path[k] + "` is not a table");
            }
            cur = &it->second;
        }
        return cur;
    }
    void assign(Value& t, const std::vector<std::string>& path, Value v) {
        Value* parent = walk(t, path, path.size() - 1, false);
        Table& tab = *parent->tab;
        if (tab.count(path.back())) error("duplicate key `" + path.back() + "`");
        tab.emplace(path.back(), std::move(v));
    }
    Value* open_table(Value& root, const std::vector<std::string>& path) {
        Value* parent = walk(root, path, path.size() - 1, true);
        Table& tab = *parent->tab;
        auto it = tab.find(path.back());
        if (it == tab.end()) {
            it = tab.emplace(path.back(), Value::table()).first;
        } else if (it->second.type != Value::Type::Table || it->second.defined_by_header) {
            error("table [" + path.back() + "] defined twice");
        }
        it->second.defined_by_header = true;
        return &it->second;
    }
    Value* open_table_array(Value& root, const std::vector<std::string>& path) {
        Value* parent = walk(root, path, path.size() - 1, true);
        Table& tab = *parent->tab;
        auto it = tab.find(path.back());
        if (it == tab.end()) {
            Value a = Value::array();
            a.is_table_array = true;
            it = tab.emplace(path.back(), a).first;
        } else if (!(it->second.type == Value::Type::Array && it->second.is_table_array)) {
            error("key `" + path.back() + "` is not an array of tables");
        }
        it->second.arr->push_back(Value::table());
        return &it->second.arr->back();
    }

    const std::string& t_;
    size_t pos_ = 0;
};

inline Table parse(const std::string& text) { return Parser(text).parse(); }

// Writing: values as TOML literals.
inline std::string quote(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\t': out += "\\t"; break;
            case '\r': out += "\\r"; break;
            default: out += c;
        }
    }
    return out + "\"";
}
// Shortest decimal that reads back as the same double, with a '.0' on
// integral values so it stays a float.
inline std::string float_literal(double v) {
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s = buf;
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

}  // namespace toml
}  // namespace vegas
