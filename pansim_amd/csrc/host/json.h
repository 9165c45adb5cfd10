// Minimal JSON reader for PanSim's input formats (agents, locations, agent
// types, location types, parameters, progression config and matrices,
// closure rules, configRandom).  It replaces the reference's rapidjson-based
// DECODE_JSON_FILE (include/IO/JSONDecoder.h:436-462).  Numbers are parsed
// with strtod (correctly rounded); the reference's rapidjson default parse is
// not full-precision, so a value may differ from the reference by one ulp of
// double before it is narrowed to float (see ps::json::Value::num).
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ps {
namespace json {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

class Value {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double n = 0.0;
    std::string raw;  // number text (for string-typed members given as numbers) or string value
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool has(const std::string& key) const;
    const Value& at(const std::string& key) const;  // throws Error when missing
    const Value& operator[](size_t i) const { return arr.at(i); }
    size_t size() const { return kind == Array ? arr.size() : obj.size(); }

    double num() const;              // Number (or numeric string)
    int64_t integer() const;         // Number truncated toward zero
    bool boolean() const;            // Bool (or 0/1 number)
    std::string str() const;         // String, or the text of a Number
    std::vector<double> numbers() const;
    std::vector<std::string> strings() const;
};

Value parse(const std::string& text, const std::string& what = "<memory>");
Value parse_file(const std::string& path);

}  // namespace json
}  // namespace ps
