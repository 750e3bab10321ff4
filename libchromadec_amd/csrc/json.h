// Minimal JSON document reader for ld-decode `.tbc.json` sidecars.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace chd::json {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Value> a;
    std::vector<std::pair<std::string, Value>> o;

    const Value *get(const char *key) const;  // object member or nullptr
    bool is_number() const { return type == Number; }
    int as_int(int fallback) const { return type == Number ? static_cast<int>(n) : fallback; }
    double as_double(double fallback) const { return type == Number ? n : fallback; }
    bool as_bool(bool fallback) const { return type == Bool ? b : fallback; }
};

// Parse `text`; on failure returns false and describes the error in `err`.
bool parse(const std::string &text, Value &out, std::string &err);

}  // namespace chd::json
