// The flat JSON document: reading (RFC 8259 grammar, checked character by character), lookups, the
// canonical text of a value, and the front-to-back writer the glTF exporter uses.
#include "json.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace phh::json {

// ---- reading -------------------------------------------------------------------------------------

class Reader {
public:
    Reader(std::string_view text, Document& doc) : in_(text), doc_(doc) {}

    void run() {
        blank();
        const uint32_t top = value(0);
        blank();
        if (at_ < in_.size()) stop("text continues after the value");
        (void)top;
    }

private:
    std::string_view in_;
    Document& doc_;
    size_t at_ = 0;
    std::vector<uint32_t> open_;  // children of the containers being read, innermost last

    [[noreturn]] void stop(const std::string& why) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < at_ && k < in_.size(); k++) {
            if (in_[k] == '\n') {
                line++;
                col = 1;
            } else {
                col++;
            }
        }
        throw SyntaxError("json: " + why + " (line " + std::to_string(line) + ", column " + std::to_string(col) + ")");
    }

    bool more() const { return at_ < in_.size(); }
    char peek() const { return in_[at_]; }

    void blank() {
        while (more()) {
            const char c = peek();
            if (c != ' ' && c != '\t' && c != '\n' && c != '\r') break;
            at_++;
        }
    }

    void expect(char c, const char* what) {
        blank();
        if (!more() || peek() != c) stop(std::string("expected ") + what);
        at_++;
    }

    uint32_t add(Tag t) {
        doc_.entries_.push_back(Document::Entry{t, 0, 0, 0.0});
        return uint32_t(doc_.entries_.size() - 1);
    }

    void word(std::string_view w) {
        if (in_.substr(at_, w.size()) != w) stop("unknown literal");
        at_ += w.size();
    }

    static int hex_digit(char c) {
        if (c >= '0' && c <= '9') return c - '0';
        if (c >= 'a' && c <= 'f') return c - 'a' + 10;
        if (c >= 'A' && c <= 'F') return c - 'A' + 10;
        return -1;
    }

    uint32_t code_unit() {
        if (in_.size() - at_ < 4) stop("short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const int d = hex_digit(in_[at_++]);
            if (d < 0) stop("bad hex digit in \\u escape");
            v = v * 16 + uint32_t(d);
        }
        return v;
    }

    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out.push_back(char(cp));
            return;
        }
        const int tail = cp < 0x800 ? 1 : cp < 0x10000 ? 2 : 3;
        static const unsigned char lead[4] = {0, 0xC0, 0xE0, 0xF0};
        out.push_back(char(lead[tail] | (cp >> (6 * tail))));
        for (int k = tail - 1; k >= 0; k--) out.push_back(char(0x80 | ((cp >> (6 * k)) & 0x3F)));
    }

    // A string literal, appended to the document's character store; returns its [lo, hi) range.
    std::pair<uint32_t, uint32_t> string_range() {
        expect('"', "a string");
        std::string& out = doc_.chars_;
        const size_t lo = out.size();
        for (;;) {
            // Plain runs (such as base64 buffers) are copied whole: the next quote, and the first
            // backslash before it, found with memchr.
            const char* base = in_.data();
            const void* q = std::memchr(base + at_, '"', in_.size() - at_);
            if (!q) stop("string not closed");
            size_t stop_at = size_t(static_cast<const char*>(q) - base);
            if (const void* bs = std::memchr(base + at_, '\\', stop_at - at_)) stop_at = size_t(static_cast<const char*>(bs) - base);
            out.append(in_.data() + at_, stop_at - at_);
            at_ = stop_at + 1;
            if (in_[stop_at] == '"') break;
            if (!more()) stop("string not closed");
            const char e = in_[at_++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    uint32_t cp = code_unit();
                    if (cp >= 0xD800 && cp <= 0xDBFF && in_.substr(at_, 2) == "\\u") {
                        at_ += 2;
                        const uint32_t low = code_unit();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (low - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: stop("unknown escape");
            }
        }
        if (out.size() > UINT32_MAX) stop("text too large");
        return {uint32_t(lo), uint32_t(out.size())};
    }

    // -? int frac? exp?, with int = 0 | [1-9][0-9]*.
    double number_literal() {
        const size_t lo = at_;
        auto digits = [&] {
            const size_t d0 = at_;
            while (more() && peek() >= '0' && peek() <= '9') at_++;
            return at_ - d0;
        };
        if (peek() == '-') at_++;
        if (!more()) stop("number without digits");
        if (peek() == '0') {
            at_++;
        } else if (digits() == 0) {
            stop("number without digits");
        }
        if (more() && peek() == '.') {
            at_++;
            if (digits() == 0) stop("no digits after the decimal point");
        }
        if (more() && (peek() == 'e' || peek() == 'E')) {
            at_++;
            if (more() && (peek() == '+' || peek() == '-')) at_++;
            if (digits() == 0) stop("no digits in the exponent");
        }
        return std::strtod(std::string(in_.substr(lo, at_ - lo)).c_str(), nullptr);
    }

    // Closes the container `id` whose children were pushed to open_ from `mark` on.
    void close(uint32_t id, size_t mark) {
        Document::Entry& e = doc_.entries_[id];
        e.lo = uint32_t(doc_.links_.size());
        doc_.links_.insert(doc_.links_.end(), open_.begin() + long(mark), open_.end());
        e.hi = uint32_t(doc_.links_.size());
        open_.resize(mark);
    }

    uint32_t value(int depth) {
        if (depth > 512) stop("nesting too deep");
        blank();
        if (!more()) stop("text ends where a value should start");
        const char c = peek();
        if (c == '{') {
            at_++;
            const uint32_t id = add(Tag::Object);
            const size_t mark = open_.size();
            blank();
            if (more() && peek() == '}') {
                at_++;
            } else {
                for (;;) {
                    blank();
                    const auto [lo, hi] = string_range();
                    const uint32_t k = add(Tag::String);
                    doc_.entries_[k].lo = lo;
                    doc_.entries_[k].hi = hi;
                    expect(':', "':' after a member name");
                    const uint32_t v = value(depth + 1);
                    open_.push_back(k);
                    open_.push_back(v);
                    blank();
                    if (more() && peek() == ',') {
                        at_++;
                        continue;
                    }
                    expect('}', "',' or '}' in an object");
                    break;
                }
            }
            close(id, mark);
            return id;
        }
        if (c == '[') {
            at_++;
            const uint32_t id = add(Tag::Array);
            const size_t mark = open_.size();
            blank();
            if (more() && peek() == ']') {
                at_++;
            } else {
                for (;;) {
                    open_.push_back(value(depth + 1));
                    blank();
                    if (more() && peek() == ',') {
                        at_++;
                        continue;
                    }
                    expect(']', "',' or ']' in an array");
                    break;
                }
            }
            close(id, mark);
            return id;
        }
        if (c == '"') {
            const auto [lo, hi] = string_range();
            const uint32_t id = add(Tag::String);
            doc_.entries_[id].lo = lo;
            doc_.entries_[id].hi = hi;
            return id;
        }
        if (c == '-' || (c >= '0' && c <= '9')) {
            const double v = number_literal();
            const uint32_t id = add(Tag::Number);
            doc_.entries_[id].number = v;
            return id;
        }
        if (c == 't') {
            word("true");
            return add(Tag::True);
        }
        if (c == 'f') {
            word("false");
            return add(Tag::False);
        }
        if (c == 'n') {
            word("null");
            return add(Tag::Null);
        }
        stop(std::string("unexpected character '") + c + "'");
    }
};

Document Document::read(std::string_view text) {
    Document d;
    d.chars_.reserve(text.size() / 2);
    Reader(text, d).run();
    return d;
}

// ---- lookups -------------------------------------------------------------------------------------

Tag Node::tag() const { return doc_ ? doc_->entries_[id_].tag : Tag::Null; }

size_t Node::length() const {
    const Tag t = tag();
    if (t != Tag::Array && t != Tag::Object) return 0;
    const Document::Entry& e = doc_->entries_[id_];
    return t == Tag::Array ? e.hi - e.lo : (e.hi - e.lo) / 2;
}

uint32_t Node::find(std::string_view key) const {
    if (tag() != Tag::Object) return 0;
    const Document::Entry& e = doc_->entries_[id_];
    for (uint32_t k = e.hi; k > e.lo; k -= 2) {  // from the end: the last of duplicates counts
        const Document::Entry& name = doc_->entries_[doc_->links_[k - 2]];
        if (std::string_view(doc_->chars_).substr(name.lo, name.hi - name.lo) == key) return k - 1;
    }
    return 0;
}

bool Node::holds(std::string_view key) const { return find(key) != 0; }

Node Node::member(std::string_view key) const {
    const uint32_t at = find(key);
    return at ? Node(doc_, doc_->links_[at]) : Node();
}

Node Node::element(size_t i) const {
    if (tag() != Tag::Array) return Node();
    const Document::Entry& e = doc_->entries_[id_];
    return i < size_t(e.hi - e.lo) ? Node(doc_, doc_->links_[e.lo + i]) : Node();
}

std::string_view Node::key_at(size_t i) const {
    if (tag() != Tag::Object || i >= length()) return {};
    const Document::Entry& name = doc_->entries_[doc_->links_[doc_->entries_[id_].lo + 2 * i]];
    return std::string_view(doc_->chars_).substr(name.lo, name.hi - name.lo);
}

Node Node::value_at(size_t i) const {
    if (tag() != Tag::Object || i >= length()) return Node();
    return Node(doc_, doc_->links_[doc_->entries_[id_].lo + 2 * i + 1]);
}

double Node::real(double otherwise) const { return tag() == Tag::Number ? doc_->entries_[id_].number : otherwise; }

int Node::whole(int otherwise) const { return tag() == Tag::Number ? int(doc_->entries_[id_].number) : otherwise; }

bool Node::truth(bool otherwise) const {
    const Tag t = tag();
    return t == Tag::True ? true : t == Tag::False ? false : otherwise;
}

std::string_view Node::text() const {
    if (tag() != Tag::String) return {};
    const Document::Entry& e = doc_->entries_[id_];
    return std::string_view(doc_->chars_).substr(e.lo, e.hi - e.lo);
}

// ---- text out -------------------------------------------------------------------------------------

std::string escaped(std::string_view s) {
    std::string out;
    out.reserve(s.size() + 2);
    out.push_back('"');
    for (const char c : s) {
        const unsigned char u = static_cast<unsigned char>(c);
        if (c == '"' || c == '\\') {
            out.push_back('\\');
            out.push_back(c);
        } else if (u >= 0x20) {
            out.push_back(c);
        } else {
            static const char* named = "btn\0fr";  // \b \t \n (\v has no short form) \f \r
            if (u >= 8 && u <= 13 && u != 11) {
                out.push_back('\\');
                out.push_back(named[u - 8]);
            } else {
                char buf[8];
                std::snprintf(buf, sizeof(buf), "\\u%04x", unsigned(u));
                out += buf;
            }
        }
    }
    out.push_back('"');
    return out;
}

namespace {

void number_text(double v, std::string& out) {
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    out += buf;
}

void canonical_into(Node v, std::string& out) {
    switch (v.tag()) {
        case Tag::Null: out += "null"; return;
        case Tag::True: out += "true"; return;
        case Tag::False: out += "false"; return;
        case Tag::Number: number_text(v.real(), out); return;
        case Tag::String: out += escaped(v.text()); return;
        case Tag::Array:
            out.push_back('[');
            for (size_t i = 0; i < v.length(); i++) {
                if (i) out.push_back(',');
                canonical_into(v.element(i), out);
            }
            out.push_back(']');
            return;
        case Tag::Object: {
            std::map<std::string_view, Node> sorted;  // a later duplicate replaces an earlier one
            for (size_t i = 0; i < v.length(); i++) sorted[v.key_at(i)] = v.value_at(i);
            out.push_back('{');
            bool any = false;
            for (const auto& [k, m] : sorted) {
                if (any) out.push_back(',');
                any = true;
                out += escaped(k);
                out.push_back(':');
                canonical_into(m, out);
            }
            out.push_back('}');
            return;
        }
    }
}

}  // namespace

std::string canonical(Node v) {
    std::string out;
    canonical_into(v, out);
    return out;
}

void Writer::separate() {
    if (after_name_) {
        after_name_ = false;
        return;
    }
    if (!first_.empty()) {
        if (!first_.back()) out_.push_back(',');
        first_.back() = false;
    }
}

Writer& Writer::open_object() {
    separate();
    out_.push_back('{');
    first_.push_back(true);
    return *this;
}

Writer& Writer::close_object() {
    out_.push_back('}');
    first_.pop_back();
    return *this;
}

Writer& Writer::open_array() {
    separate();
    out_.push_back('[');
    first_.push_back(true);
    return *this;
}

Writer& Writer::close_array() {
    out_.push_back(']');
    first_.pop_back();
    return *this;
}

Writer& Writer::name(std::string_view key) {
    separate();
    out_ += escaped(key);
    out_.push_back(':');
    after_name_ = true;
    return *this;
}

Writer& Writer::number(double v) {
    separate();
    number_text(v, out_);
    return *this;
}

Writer& Writer::integer(int64_t v) {
    separate();
    out_ += std::to_string(v);
    return *this;
}

Writer& Writer::boolean(bool v) {
    separate();
    out_ += v ? "true" : "false";
    return *this;
}

Writer& Writer::string(std::string_view s) {
    separate();
    out_ += escaped(s);
    return *this;
}

Writer& Writer::numbers(const float* v, int n) {
    open_array();
    for (int i = 0; i < n; i++) number(double(v[i]));
    return close_array();
}

}  // namespace phh::json
