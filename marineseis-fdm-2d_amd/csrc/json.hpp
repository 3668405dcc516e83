// Minimal JSON document model and dotted-path lookup for the solver's config.
//
// Mirrors the behaviour the reference relies on (include/json_config.hpp):
//   * paths "a.b.c", "a.b[0]", "a.b[0].c" (json_config.hpp:241-356);
//   * numbers are doubles; float = (float)value (json_config.hpp:392-406);
//     int = cJSON's valueint, i.e. the double saturated to [INT_MIN, INT_MAX]
//     and truncated toward zero (json_config.hpp:376-389);
//   * a missing path, a non-number where a number is expected, a non-string
//     where a path string is expected, and a parse failure are errors.
// The parser is a plain recursive-descent RFC 8259 reader written for this
// project; no third-party JSON code is used or linked.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace msfdm {

struct JsonError : std::runtime_error {
    explicit JsonError(const std::string &m) : std::runtime_error(m) {}
};

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<std::unique_ptr<JsonValue>> items;                       // Array
    std::vector<std::pair<std::string, std::unique_ptr<JsonValue>>> members;  // Object (ordered)
};

class JsonDoc {
public:
    static JsonDoc parse(const std::string &text);      // throws JsonError
    static JsonDoc load_file(const std::string &path);  // throws JsonError

    const JsonValue *find(const std::string &path) const;  // nullptr if absent
    bool has(const std::string &path) const { return find(path) != nullptr; }

    int get_int(const std::string &path) const;
    int get_int(const std::string &path, size_t index) const;
    float get_float(const std::string &path) const;
    std::string get_string(const std::string &path) const;
    size_t array_size(const std::string &path) const;

private:
    std::unique_ptr<JsonValue> root_;
};

// cJSON-compatible double -> int conversion (saturating, truncating).
int json_valueint(double v);

}  // namespace msfdm
