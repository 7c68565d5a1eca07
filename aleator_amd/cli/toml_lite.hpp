// toml_lite.hpp -- the part of TOML v1.0 that run configurations use, read into a tree of nodes
// that remember where they came from (line, column), so that configuration errors name the
// exact key.
//
// Supported: comments; [table] and [[array.of.tables]] headers with dotted names; bare, quoted
// and dotted keys; basic strings with the standard escapes, literal strings; integers (decimal
// with '_' separators, 0x / 0o / 0b); floats (fraction and/or exponent, inf, nan); booleans;
// arrays of values on one or more lines.  Not supported, and reported as a parse error:
// multi-line strings, inline tables, dates and times.  Redefining a key is an error.
#pragma once

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace toml_lite {

struct ParseError : std::runtime_error {
    int line, column;
    ParseError(const std::string& msg, int l, int c) : std::runtime_error(msg), line(l), column(c) {}
};

struct Node {
    enum Kind { String, Integer, Float, Boolean, Array, Table } kind = Table;
    std::string str;
    std::int64_t integer = 0;
    double real = 0.0;
    bool boolean = false;
    std::vector<Node> items;                       // Array
    std::vector<std::pair<std::string, Node>> kv;  // Table, in file order
    int line = 1, column = 1;
    bool header_defined = false;  // a [table] header named this table
    bool array_of_tables = false;  // an Array built by [[...]] headers

    const Node* find(const std::string& key) const {
        for (const auto& [k, v] : kv)
            if (k == key) return &v;
        return nullptr;
    }
    Node* find(const std::string& key) {
        for (auto& [k, v] : kv)
            if (k == key) return &v;
        return nullptr;
    }
    const char* kind_name() const {
        switch (kind) {
            case String: return "string";
            case Integer: return "integer";
            case Float: return "floating-point";
            case Boolean: return "boolean";
            case Array: return "array";
            default: return "table";
        }
    }
    // Typed reads in the manner of toml++'s value<T>(): an integer also reads as a double.
    std::optional<double> as_double() const {
        if (kind == Float) return real;
        if (kind == Integer) return (double)integer;
        return std::nullopt;
    }
    std::optional<std::int64_t> as_int() const {
        return kind == Integer ? std::optional<std::int64_t>(integer) : std::nullopt;
    }
    std::optional<std::string> as_string() const {
        return kind == String ? std::optional<std::string>(str) : std::nullopt;
    }
};

namespace detail {

class Reader {
public:
    explicit Reader(std::string text) : t_(std::move(text)) {}

    Node parse() {
        Node root;
        root.kind = Node::Table;
        Node* cur = &root;
        for (;;) {
            skip_ws_comments_newlines();
            if (at_end()) break;
            if (peek() == '[') {
                cur = header(root);
            } else {
                key_value(*cur);
            }
            end_of_line();
        }
        return root;
    }

private:
    [[noreturn]] void error(const std::string& msg) const { throw ParseError(msg, line_, col_); }
    bool at_end() const { return p_ >= t_.size(); }
    char peek(size_t k = 0) const { return p_ + k < t_.size() ? t_[p_ + k] : '\0'; }
    char get() {
        const char c = t_[p_++];
        if (c == '\n') {
            ++line_;
            col_ = 1;
        } else {
            ++col_;
        }
        return c;
    }
    void skip_ws() {
        while (!at_end() && (peek() == ' ' || peek() == '\t')) get();
    }
    void skip_comment() {
        if (peek() == '#')
            while (!at_end() && peek() != '\n') get();
    }
    void skip_ws_comments_newlines() {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\r' && peek(1) == '\n') get();
            if (!at_end() && peek() == '\n') {
                get();
                continue;
            }
            return;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (peek() == '\r' && peek(1) == '\n') get();
        if (!at_end() && peek() != '\n') error(std::string("unexpected '") + peek() + "' after a value");
    }

    static bool bare(char c) {
        return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-';
    }

    std::vector<std::string> dotted_key() {
        std::vector<std::string> parts;
        for (;;) {
            skip_ws();
            if (peek() == '"') {
                parts.push_back(basic_string());
            } else if (peek() == '\'') {
                parts.push_back(literal_string());
            } else {
                std::string k;
                while (bare(peek())) k += get();
                if (k.empty()) error("expected a key");
                parts.push_back(k);
            }
            skip_ws();
            if (peek() != '.') return parts;
            get();
        }
    }

    // Walks / creates the tables of `path` below `root` (arrays of tables resolve to their
    // last element).
    Node* descend(Node& root, const std::vector<std::string>& path, size_t upto) {
        Node* n = &root;
        for (size_t i = 0; i < upto; ++i) {
            Node* next = n->find(path[i]);
            if (!next) {
                Node t;
                t.kind = Node::Table;
                t.line = line_;
                t.column = col_;
                n->kv.emplace_back(path[i], std::move(t));
                next = &n->kv.back().second;
            }
            if (next->kind == Node::Array && next->array_of_tables && !next->items.empty()) {
                next = &next->items.back();
            } else if (next->kind != Node::Table) {
                error("key '" + path[i] + "' is not a table");
            }
            n = next;
        }
        return n;
    }

    Node* header(Node& root) {
        const int l = line_, c = col_;
        get();
        const bool arr = peek() == '[';
        if (arr) get();
        const auto path = dotted_key();
        if (get() != ']' || (arr && get() != ']')) error("malformed table header");
        Node* parent = descend(root, path, path.size() - 1);
        const std::string& last = path.back();
        Node* existing = parent->find(last);
        if (arr) {
            if (!existing) {
                Node a;
                a.kind = Node::Array;
                a.array_of_tables = true;
                a.line = l;
                a.column = c;
                parent->kv.emplace_back(last, std::move(a));
                existing = &parent->kv.back().second;
            } else if (!(existing->kind == Node::Array && existing->array_of_tables)) {
                throw ParseError("redefinition of '" + last + "' as an array of tables", l, c);
            }
            Node t;
            t.kind = Node::Table;
            t.line = l;
            t.column = c;
            existing->items.push_back(std::move(t));
            return &existing->items.back();
        }
        if (existing) {
            if (existing->kind != Node::Table || existing->header_defined)
                throw ParseError("redefinition of table '" + last + "'", l, c);
            existing->header_defined = true;
            return existing;
        }
        Node t;
        t.kind = Node::Table;
        t.line = l;
        t.column = c;
        t.header_defined = true;
        parent->kv.emplace_back(last, std::move(t));
        return &parent->kv.back().second;
    }

    void key_value(Node& table) {
        const int l = line_, c = col_;
        const auto path = dotted_key();
        skip_ws();
        if (get() != '=') throw ParseError("expected '=' after a key", l, c);
        skip_ws();
        Node v = value();
        v.line = l;
        v.column = c;
        Node* parent = descend(table, path, path.size() - 1);
        if (parent->find(path.back())) throw ParseError("redefinition of key '" + path.back() + "'", l, c);
        parent->kv.emplace_back(path.back(), std::move(v));
    }

    Node value() {
        Node v;
        const char c = peek();
        if (c == '"') {
            if (peek(1) == '"' && peek(2) == '"') error("multi-line strings are not supported");
            v.kind = Node::String;
            v.str = basic_string();
        } else if (c == '\'') {
            if (peek(1) == '\'' && peek(2) == '\'') error("multi-line strings are not supported");
            v.kind = Node::String;
            v.str = literal_string();
        } else if (c == '[') {
            get();
            v.kind = Node::Array;
            for (;;) {
                skip_ws_comments_newlines();
                if (peek() == ']') {
                    get();
                    break;
                }
                const int l = line_, cc = col_;
                Node item = value();
                item.line = l;
                item.column = cc;
                v.items.push_back(std::move(item));
                skip_ws_comments_newlines();
                if (peek() == ',') {
                    get();
                } else if (peek() == ']') {
                    get();
                    break;
                } else {
                    error("expected ',' or ']' in an array");
                }
            }
        } else if (c == '{') {
            error("inline tables are not supported");
        } else if (t_.compare(p_, 4, "true") == 0 && !bare(peek(4))) {
            for (int i = 0; i < 4; ++i) get();
            v.kind = Node::Boolean;
            v.boolean = true;
        } else if (t_.compare(p_, 5, "false") == 0 && !bare(peek(5))) {
            for (int i = 0; i < 5; ++i) get();
            v.kind = Node::Boolean;
        } else {
            number(v);
        }
        return v;
    }

    std::string basic_string() {
        get();  // opening quote
        std::string s;
        for (;;) {
            if (at_end() || peek() == '\n') error("unterminated string");
            char ch = get();
            if (ch == '"') return s;
            if (ch != '\\') {
                s += ch;
                continue;
            }
            ch = at_end() ? '\0' : get();
            switch (ch) {
                case 'b': s += '\b'; break;
                case 't': s += '\t'; break;
                case 'n': s += '\n'; break;
                case 'f': s += '\f'; break;
                case 'r': s += '\r'; break;
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case 'u':
                case 'U': {
                    const int n = ch == 'u' ? 4 : 8;
                    uint32_t cp = 0;
                    for (int i = 0; i < n; ++i) {
                        const char h = at_end() ? '\0' : get();
                        const int d = (h >= '0' && h <= '9') ? h - '0'
                                      : (h >= 'a' && h <= 'f') ? h - 'a' + 10
                                      : (h >= 'A' && h <= 'F') ? h - 'A' + 10
                                                               : -1;
                        if (d < 0) error("malformed unicode escape");
                        cp = cp * 16 + (uint32_t)d;
                    }
                    utf8(s, cp);
                    break;
                }
                default: error("unknown escape sequence in a string");
            }
        }
    }

    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }

    std::string literal_string() {
        get();
        std::string s;
        for (;;) {
            if (at_end() || peek() == '\n') error("unterminated string");
            const char ch = get();
            if (ch == '\'') return s;
            s += ch;
        }
    }

    void number(Node& v) {
        std::string raw;
        while (!at_end() && (bare(peek()) || peek() == '.' || peek() == '+')) raw += get();
        if (raw.empty()) error("expected a value");
        std::string s;  // without the '_' separators, each of which must sit between digits
        for (size_t i = 0; i < raw.size(); ++i) {
            if (raw[i] == '_') {
                if (i == 0 || i + 1 == raw.size() || !std::isalnum((unsigned char)raw[i - 1]) ||
                    !std::isalnum((unsigned char)raw[i + 1]))
                    error("misplaced '_' in a number");
                continue;
            }
            s += raw[i];
        }
        const std::string body = (s[0] == '+' || s[0] == '-') ? s.substr(1) : s;
        const bool neg = s[0] == '-';
        if (body == "inf" || body == "nan") {
            v.kind = Node::Float;
            v.real = body == "inf" ? (neg ? -HUGE_VAL : HUGE_VAL) : std::nan("");
            return;
        }
        if (body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b')) {
            if (s[0] == '+' || s[0] == '-') error("a sign is not allowed on a hex, octal or binary integer");
            const int base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
            errno = 0;
            char* end = nullptr;
            const unsigned long long u = std::strtoull(body.c_str() + 2, &end, base);
            if (*end != '\0' || errno == ERANGE || u > (unsigned long long)INT64_MAX) error("malformed integer");
            v.kind = Node::Integer;
            v.integer = (std::int64_t)u;
            return;
        }
        const bool is_float = body.find_first_of(".eE") != std::string::npos;
        for (char ch : body)
            if (!(std::isdigit((unsigned char)ch) || ch == '.' || ch == 'e' || ch == 'E' || ch == '+' || ch == '-'))
                error("malformed value '" + raw + "'");
        if (body.empty() || !std::isdigit((unsigned char)body[0])) error("malformed number '" + raw + "'");
        if (body.size() > 1 && body[0] == '0' && std::isdigit((unsigned char)body[1]))
            error("leading zeros are not allowed in '" + raw + "'");
        errno = 0;
        char* end = nullptr;
        if (is_float) {
            const size_t dot = body.find('.');
            if (dot != std::string::npos && (dot + 1 >= body.size() || !std::isdigit((unsigned char)body[dot + 1])))
                error("a decimal point must be followed by a digit in '" + raw + "'");
            v.kind = Node::Float;
            v.real = std::strtod(s.c_str(), &end);
        } else {
            v.kind = Node::Integer;
            v.integer = std::strtoll(s.c_str(), &end, 10);
        }
        if (*end != '\0' || errno == ERANGE) error("malformed number '" + raw + "'");
    }

    std::string t_;
    size_t p_ = 0;
    int line_ = 1, col_ = 1;
};

}  // namespace detail

inline Node parse(const std::string& text) { return detail::Reader(text).parse(); }

inline Node parse_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ParseError("could not open file", 1, 1);
    std::ostringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

}  // namespace toml_lite
