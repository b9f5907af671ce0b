// JSON (RFC 8259) for scene files, as a flat document: every value of a text is one entry of an array,
// and arrays and objects name their children by position in a second array, so that a parsed glTF
// document is three allocations however many values it holds. Lookups follow the conventions glTF
// readers rely on (the reference's loader, src/scene/gltf.cpp, reads its documents this way): a missing
// member or element reads as null, a typed read of a value of another type returns the caller's
// default, and of duplicate member names the last one counts.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace phh::json {

// A malformed text: the message names the reason, then the line and column where reading stopped.
struct SyntaxError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

enum class Tag : uint8_t { Null, False, True, Number, String, Array, Object };

class Document;

// A read-only handle to one value of a Document. Default-constructed, or obtained by looking up what is
// not there, it is a null value; every lookup on it is null again.
class Node {
public:
    Node() = default;

    Tag tag() const;
    bool null() const { return tag() == Tag::Null; }
    bool is_array() const { return tag() == Tag::Array; }
    bool is_object() const { return tag() == Tag::Object; }
    bool is_number() const { return tag() == Tag::Number; }
    bool is_string() const { return tag() == Tag::String; }

    // Object member by name (the last of duplicates); array element by index.
    Node member(std::string_view key) const;
    Node element(size_t i) const;
    bool holds(std::string_view key) const;
    // Elements of an array or members of an object; 0 for any other value.
    size_t length() const;
    // The i-th member of an object in document order (duplicates included).
    std::string_view key_at(size_t i) const;
    Node value_at(size_t i) const;

    double real(double otherwise = 0.0) const;
    // The number truncated toward zero, as a C cast does.
    int whole(int otherwise = -1) const;
    bool truth(bool otherwise = false) const;
    // The string's bytes (UTF-8), or an empty view for a value of another type.
    std::string_view text() const;

private:
    friend class Document;
    Node(const Document* d, uint32_t id) : doc_(d), id_(id) {}
    // Position in the links array of the value of the last member named `key`, or 0 when there is none.
    uint32_t find(std::string_view key) const;
    const Document* doc_ = nullptr;
    uint32_t id_ = 0;
};

class Document {
public:
    // Reads one JSON text; white space may surround it. Nesting deeper than 512 levels is refused.
    static Document read(std::string_view text);
    Node root() const { return entries_.empty() ? Node() : Node(this, 0); }

private:
    friend class Node;
    friend class Reader;
    struct Entry {
        Tag tag = Tag::Null;
        uint32_t lo = 0, hi = 0;  // strings: byte range in chars_; arrays/objects: range in links_
        double number = 0.0;
    };
    std::vector<Entry> entries_;
    // Children of containers, contiguous per container; an object's run alternates key entry, value entry.
    std::vector<uint32_t> links_;
    std::string chars_;
};

// The text of a JSON string literal for `s`, quotes included.
std::string escaped(std::string_view s);

// Canonical text of a value: members sorted by name (duplicates resolved to the last), numbers with 17
// significant digits, no white space.
std::string canonical(Node v);

// Builds JSON text front to back; commas and colons are placed by the writer.
class Writer {
public:
    Writer& open_object();
    Writer& close_object();
    Writer& open_array();
    Writer& close_array();
    Writer& name(std::string_view key);
    Writer& number(double v);
    Writer& integer(int64_t v);
    Writer& boolean(bool v);
    Writer& string(std::string_view s);
    // Shorthands for an object member.
    template <class T>
    Writer& field(std::string_view key, const T& v) {
        name(key);
        return put(v);
    }
    Writer& numbers(const float* v, int n);
    const std::string& str() const { return out_; }

private:
    Writer& put(double v) { return number(v); }
    Writer& put(float v) { return number(double(v)); }
    Writer& put(int v) { return integer(v); }
    Writer& put(int64_t v) { return integer(v); }
    Writer& put(size_t v) { return integer(int64_t(v)); }
    Writer& put(bool v) { return boolean(v); }
    Writer& put(const std::string& v) { return string(v); }
    Writer& put(const char* v) { return string(v); }
    void separate();
    std::string out_;
    std::vector<bool> first_;  // per open container: nothing written in it yet
    bool after_name_ = false;
};

}  // namespace phh::json
