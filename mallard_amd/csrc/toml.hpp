// Minimal TOML reader covering the subset Mallard inputs use: tables,
// arrays of tables, dotted keys, basic / literal / multi-line strings
// (with line-ending backslash), integers, floats, booleans, arrays and
// inline tables. It replaces toml11 for the keys the reference reads with
// toml::find / toml::find_or (solver.cpp:51 and the init routines).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mlr::toml {

struct Value {
    enum Kind { Null, Bool, Int, Float, String, Array, Table };
    Kind kind = Null;
    bool b = false;
    int64_t i = 0;
    double f = 0.0;
    std::string s;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> tab;  // insertion order

    bool is_table() const { return kind == Table; }
    const Value* get(const std::string& key) const;
    Value* get_mut(const std::string& key);
    bool contains(const std::string& key) const { return get(key) != nullptr; }
    Value& set(const std::string& key, Value v);

    double as_float() const;   // accepts integers too
    int64_t as_int() const;
    bool as_bool() const;
    const std::string& as_string() const;
    std::string dump(int indent = 0) const;
};

Value parse(const std::string& text, const std::string& origin = "<string>");
Value parse_file(const std::string& path);

// find(v, "a", "b") walks nested tables; throws with the key path when absent.
const Value& find(const Value& v, const std::string& k1);
const Value& find(const Value& v, const std::string& k1, const std::string& k2);
const Value& find(const Value& v, const std::string& k1, const std::string& k2,
                  const std::string& k3);
const Value* lookup(const Value& v, std::initializer_list<const char*> path);

double find_or(const Value& v, std::initializer_list<const char*> path, double def);
int64_t find_or_int(const Value& v, std::initializer_list<const char*> path, int64_t def);
bool find_or_bool(const Value& v, std::initializer_list<const char*> path, bool def);
std::string find_or_str(const Value& v, std::initializer_list<const char*> path,
                        const std::string& def);

}  // namespace mlr::toml
