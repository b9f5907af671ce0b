// JSON reader and writer for scene files (json.hpp). Behaviour follows the reference parser
// (src/scene/json.cpp:9-160): objects keep the last of duplicate keys, \u escapes (with surrogate
// pairs) become UTF-8, nesting is limited to 512 levels, and an error names the line it was found on.
// Numbers follow the RFC 8259 grammar and are converted by strtod in the C locale.
#include "json.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace phh::json {

const Value& Value::operator[](const std::string& key) const {
    static const Value none;
    if (kind != Kind::Object) return none;
    const auto it = members.find(key);
    return it == members.end() ? none : it->second;
}

const Value& Value::operator[](size_t i) const {
    static const Value none;
    return kind == Kind::Array && i < items.size() ? items[i] : none;
}

namespace {

class Reader {
public:
    explicit Reader(const std::string& s) : src_(s), n_(s.size()) {}

    Value document() {
        Value v = value(0);
        skip_space();
        if (pos_ != n_) error("trailing characters after the JSON value");
        return v;
    }

private:
    const std::string& src_;
    size_t n_;
    size_t pos_ = 0;

    [[noreturn]] void error(const std::string& what) const {
        size_t line = 1;
        for (size_t k = 0; k < pos_ && k < n_; k++) line += src_[k] == '\n' ? 1 : 0;
        throw ParseError("JSON error at line " + std::to_string(line) + ": " + what);
    }

    void skip_space() {
        while (pos_ < n_) {
            const char c = src_[pos_];
            if (c != ' ' && c != '\t' && c != '\n' && c != '\r') break;
            pos_++;
        }
    }

    // Skips white space, then takes `c` if it is next.
    bool take(char c) {
        skip_space();
        if (pos_ < n_ && src_[pos_] == c) {
            pos_++;
            return true;
        }
        return false;
    }

    void require(char c) {
        if (!take(c)) error(std::string("expected '") + c + "'");
    }

    bool literal(const char* word) {
        const size_t len = std::strlen(word);
        if (src_.compare(pos_, len, word) != 0) return false;
        pos_ += len;
        return true;
    }

    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80u) {
            out.push_back(char(cp));
            return;
        }
        if (cp < 0x800u) {
            out.push_back(char(0xC0u | (cp >> 6)));
        } else if (cp < 0x10000u) {
            out.push_back(char(0xE0u | (cp >> 12)));
            out.push_back(char(0x80u | ((cp >> 6) & 0x3Fu)));
        } else {
            out.push_back(char(0xF0u | (cp >> 18)));
            out.push_back(char(0x80u | ((cp >> 12) & 0x3Fu)));
            out.push_back(char(0x80u | ((cp >> 6) & 0x3Fu)));
        }
        out.push_back(char(0x80u | (cp & 0x3Fu)));
    }

    uint32_t hex_quad() {
        if (pos_ + 4 > n_) error("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = src_[pos_++];
            int d;
            if (c >= '0' && c <= '9')
                d = c - '0';
            else if (c >= 'a' && c <= 'f')
                d = c - 'a' + 10;
            else if (c >= 'A' && c <= 'F')
                d = c - 'A' + 10;
            else
                error("invalid \\u escape");
            v = (v << 4) | uint32_t(d);
        }
        return v;
    }

    std::string string_literal() {
        require('"');
        std::string out;
        for (;;) {
            // Runs without quotes or escapes (such as a base64 buffer) are copied in one piece.
            size_t run = pos_;
            while (run < n_ && src_[run] != '"' && src_[run] != '\\') run++;
            out.append(src_, pos_, run - pos_);
            pos_ = run;
            if (pos_ >= n_) error("unterminated string");
            const char c = src_[pos_++];
            if (c == '"') return out;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (pos_ >= n_) error("unterminated escape");
            const char e = src_[pos_++];
            switch (e) {
                case '"':
                case '\\':
                case '/': out.push_back(e); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    uint32_t cp = hex_quad();
                    // A high surrogate followed by an escaped low one is one code point.
                    if (cp >= 0xD800u && cp < 0xDC00u && pos_ + 1 < n_ && src_[pos_] == '\\' && src_[pos_ + 1] == 'u') {
                        pos_ += 2;
                        const uint32_t lo = hex_quad();
                        cp = 0x10000u + ((cp - 0xD800u) << 10) + (lo - 0xDC00u);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: error(std::string("invalid escape \\") + e);
            }
        }
    }

    static bool digit(char c) { return c >= '0' && c <= '9'; }

    double number() {
        // -? (0 | [1-9][0-9]*) (. [0-9]+)? ([eE] [+-]? [0-9]+)?
        const size_t start = pos_;
        if (src_[pos_] == '-') pos_++;
        if (pos_ >= n_ || !digit(src_[pos_])) error("invalid number");
        if (src_[pos_] == '0')
            pos_++;
        else
            while (pos_ < n_ && digit(src_[pos_])) pos_++;
        if (pos_ < n_ && src_[pos_] == '.') {
            pos_++;
            if (pos_ >= n_ || !digit(src_[pos_])) error("invalid number");
            while (pos_ < n_ && digit(src_[pos_])) pos_++;
        }
        if (pos_ < n_ && (src_[pos_] == 'e' || src_[pos_] == 'E')) {
            pos_++;
            if (pos_ < n_ && (src_[pos_] == '+' || src_[pos_] == '-')) pos_++;
            if (pos_ >= n_ || !digit(src_[pos_])) error("invalid number");
            while (pos_ < n_ && digit(src_[pos_])) pos_++;
        }
        const std::string lexeme = src_.substr(start, pos_ - start);
        return std::strtod(lexeme.c_str(), nullptr);
    }

    Value value(int depth) {
        if (depth > 512) error("nesting too deep");
        skip_space();
        if (pos_ >= n_) error("unexpected end of input");
        Value v;
        const char c = src_[pos_];
        if (c == '{') {
            pos_++;
            v.kind = Value::Kind::Object;
            if (take('}')) return v;
            do {
                skip_space();
                std::string key = string_literal();
                require(':');
                v.members[key] = value(depth + 1);
            } while (take(','));
            require('}');
        } else if (c == '[') {
            pos_++;
            v.kind = Value::Kind::Array;
            if (take(']')) return v;
            do v.items.push_back(value(depth + 1));
            while (take(','));
            require(']');
        } else if (c == '"') {
            v.kind = Value::Kind::String;
            v.text = string_literal();
        } else if (c == '-' || digit(c)) {
            v.kind = Value::Kind::Number;
            v.num = number();
        } else if (literal("true")) {
            v.kind = Value::Kind::Bool;
            v.flag = true;
        } else if (literal("false")) {
            v.kind = Value::Kind::Bool;
        } else if (!literal("null")) {
            error(std::string("unexpected character '") + c + "'");
        }
        return v;
    }
};

void write_to(const Value& v, std::string& out) {
    switch (v.kind) {
        case Value::Kind::Null: out += "null"; break;
        case Value::Kind::Bool: out += v.flag ? "true" : "false"; break;
        case Value::Kind::Number: {
            char buf[40];
            std::snprintf(buf, sizeof(buf), "%.17g", v.num);
            out += buf;
            break;
        }
        case Value::Kind::String: out += quote(v.text); break;
        case Value::Kind::Array: {
            out.push_back('[');
            for (size_t i = 0; i < v.items.size(); i++) {
                if (i) out.push_back(',');
                write_to(v.items[i], out);
            }
            out.push_back(']');
            break;
        }
        case Value::Kind::Object: {
            out.push_back('{');
            bool first = true;
            for (const auto& [k, m] : v.members) {
                if (!first) out.push_back(',');
                first = false;
                out += quote(k);
                out.push_back(':');
                write_to(m, out);
            }
            out.push_back('}');
            break;
        }
    }
}

}  // namespace

Value parse(const std::string& text) { return Reader(text).document(); }

std::string quote(const std::string& s) {
    std::string out = "\"";
    for (const char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if (static_cast<unsigned char>(c) < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", unsigned(static_cast<unsigned char>(c)));
                    out += buf;
                } else {
                    out.push_back(c);
                }
        }
    }
    out.push_back('"');
    return out;
}

std::string write(const Value& v) {
    std::string out;
    write_to(v, out);
    return out;
}

}  // namespace phh::json
