// A TOML reader for scene files: the subset of TOML v0.5 that the toml 0.5 crate reads for
// examples/test_mods/config.rs (tables, dotted keys, arrays of tables, inline tables,
// basic/literal strings, integers, floats, booleans).  Tables keep their keys sorted by
// bytes, the iteration order of toml 0.5's BTreeMap-backed tables, which fixes the
// material numbering of a scene.  Dates are rejected.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace arp_toml {

enum class Kind { String, Integer, Float, Boolean, Array, Table };

struct Value;
using VPtr = std::shared_ptr<Value>;

struct Value {
    Kind kind = Kind::Table;
    std::string str;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    std::vector<VPtr> arr;
    std::map<std::string, VPtr> tab;
    // parser bookkeeping
    bool defined = false;  // a [header] or a key = value defined this table
    bool closed = false;   // inline table or value array: no later additions
    bool aot = false;      // array created by [[header]]
    bool dotted = false;   // table created by a dotted key

    const Value* get(const std::string& k) const {
        if (kind != Kind::Table) return nullptr;
        auto it = tab.find(k);
        return it == tab.end() ? nullptr : it->second.get();
    }
    bool is_table() const { return kind == Kind::Table; }
    bool is_array() const { return kind == Kind::Array; }
    bool is_float() const { return kind == Kind::Float; }
    bool is_int() const { return kind == Kind::Integer; }
    bool is_str() const { return kind == Kind::String; }
    bool is_bool() const { return kind == Kind::Boolean; }
};

// Parses a document; returns the root table, or nullptr with a message (line number
// included) in err.
VPtr parse(const char* text, size_t len, std::string& err);

// JSON text of a value (floats always carry a '.', an exponent or Infinity/NaN, so a JSON
// reader tells them from integers).
std::string to_json(const Value& v);

const char* kind_name(Kind k);

}  // namespace arp_toml
