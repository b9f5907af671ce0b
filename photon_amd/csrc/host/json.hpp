// JSON (RFC 8259) for scene files: a value tree, a parser with line-numbered errors, and a canonical
// writer. Replaces the reference's json::Value / json::parse (src/scene/json.h:11-58) with the same
// lookup semantics: a missing member or an out-of-range element reads as null, and the typed getters
// return a fallback for a value of another type.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace phh::json {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Value {
public:
    enum class Kind : uint8_t { Null, Bool, Number, String, Array, Object };

    Kind kind = Kind::Null;
    bool flag = false;
    double num = 0.0;
    std::string text;
    std::vector<Value> items;
    std::map<std::string, Value> members;

    bool is_null() const { return kind == Kind::Null; }
    bool is_object() const { return kind == Kind::Object; }
    bool is_array() const { return kind == Kind::Array; }
    bool is_number() const { return kind == Kind::Number; }
    bool is_string() const { return kind == Kind::String; }

    bool has(const std::string& key) const { return kind == Kind::Object && members.count(key) != 0; }
    const Value& operator[](const std::string& key) const;
    const Value& operator[](size_t i) const;
    // Elements of an array, members of an object, 0 otherwise.
    size_t size() const { return kind == Kind::Array ? items.size() : kind == Kind::Object ? members.size() : 0; }

    double number(double fallback = 0.0) const { return kind == Kind::Number ? num : fallback; }
    // Truncates toward zero, as a C cast of the number does.
    int integer(int fallback = -1) const { return kind == Kind::Number ? int(num) : fallback; }
    bool boolean(bool fallback = false) const { return kind == Kind::Bool ? flag : fallback; }
    // The string, or "" for a value of another kind.
    const std::string& string() const { return text; }
};

// Parses one JSON text (surrounding white space allowed). Throws ParseError naming the line.
Value parse(const std::string& text);

// A JSON string literal for `s` (quotes included).
std::string quote(const std::string& s);

// Canonical text: members in key order, numbers with 17 significant digits, no white space.
std::string write(const Value& v);

}  // namespace phh::json
