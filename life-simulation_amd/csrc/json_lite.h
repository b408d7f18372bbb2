// Minimal JSON reader for scene and sweep files: the value kinds of RFC 8259,
// parsed into a tree.  Lookups mirror how the reference reads its scene JSON
// (nlohmann::json::value(key, default)): a missing key gives the default, a
// present key of the wrong kind is an error.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace lenia {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool is_number() const { return kind == Number; }
    bool is_array() const { return kind == Array; }
    bool is_object() const { return kind == Object; }
    bool is_string() const { return kind == String; }
    bool is_bool() const { return kind == Bool; }
    // member of an object (nullptr if absent or not an object)
    const Value *get(const std::string &key) const;
    bool contains(const std::string &key) const { return get(key) != nullptr; }
};

// Parse `len` bytes of text.  On failure returns false and describes the
// first error (with its byte offset) in `err`.
bool parse(const char *text, size_t len, Value &out, std::string &err);

// Typed member lookups with a default.  Each returns false (and sets `err`)
// when the member exists with a kind the target cannot take.
bool get_number(const Value &o, const std::string &key, double dflt, double &out, std::string &err);
bool get_string(const Value &o, const std::string &key, const std::string &dflt, std::string &out,
                std::string &err);
bool get_bool(const Value &o, const std::string &key, bool dflt, bool &out, std::string &err);

// compact JSON text of a value (numbers with up to 17 significant digits)
std::string dump(const Value &v);

}  // namespace json
}  // namespace lenia
