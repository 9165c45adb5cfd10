#include "json.h"

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace ps {
namespace json {

namespace {

struct Parser {
    const char* begin;
    const char* p;
    const char* end;
    const std::string& what;

    [[noreturn]] void fail(const std::string& msg) {
        throw Error("JSON parse error in " + what + " at offset " + std::to_string(p - begin) +
                    ": " + msg);
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* s) {
        size_t n = std::strlen(s);
        if (size_t(end - p) >= n && std::memcmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    std::string string_body() {
        // p points after the opening quote
        std::string out;
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= end) fail("bad escape");
                char e = *p++;
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
                    if (end - p < 4) fail("bad \\u escape");
                    unsigned cp = std::strtoul(std::string(p, p + 4).c_str(), nullptr, 16);
                    p += 4;
                    if (cp < 0x80) out.push_back(char(cp));
                    else if (cp < 0x800) {
                        out.push_back(char(0xC0 | (cp >> 6)));
                        out.push_back(char(0x80 | (cp & 0x3F)));
                    } else {
                        out.push_back(char(0xE0 | (cp >> 12)));
                        out.push_back(char(0x80 | ((cp >> 6) & 0x3F)));
                        out.push_back(char(0x80 | (cp & 0x3F)));
                    }
                    break;
                }
                default: fail("bad escape");
                }
            } else {
                out.push_back(c);
            }
        }
        if (p >= end) fail("unterminated string");
        ++p;
        return out;
    }
    Value value() {
        ws();
        if (p >= end) fail("unexpected end");
        Value v;
        char c = *p;
        if (c == '{') {
            ++p;
            v.kind = Value::Object;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return v;
            }
            for (;;) {
                ws();
                if (p >= end || *p != '"') fail("expected key");
                ++p;
                std::string key = string_body();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                ++p;
                v.obj.emplace_back(std::move(key), value());
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            ++p;
            v.kind = Value::Array;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            ++p;
            v.kind = Value::String;
            v.raw = string_body();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
            v.b = false;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else {
            const char* s = p;
            if (p < end && (*p == '-' || *p == '+')) ++p;
            while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' || *p == '+'))
                ++p;
            if (p == s) fail(std::string("unexpected character '") + c + "'");
            v.kind = Value::Number;
            v.raw.assign(s, p);
            errno = 0;
            v.n = std::strtod(v.raw.c_str(), nullptr);
        }
        return v;
    }
};

}  // namespace

bool Value::has(const std::string& key) const {
    if (kind != Object) return false;
    for (auto& kv : obj)
        if (kv.first == key) return true;
    return false;
}

const Value& Value::at(const std::string& key) const {
    if (kind == Object)
        for (auto& kv : obj)
            if (kv.first == key) return kv.second;
    throw Error("missing JSON member '" + key + "'");
}

double Value::num() const {
    if (kind == Number) return n;
    if (kind == String) {
        char* e = nullptr;
        double d = std::strtod(raw.c_str(), &e);
        if (e != raw.c_str()) return d;
    }
    if (kind == Bool) return b ? 1.0 : 0.0;
    throw Error("JSON value is not a number");
}

int64_t Value::integer() const { return int64_t(num()); }

bool Value::boolean() const {
    if (kind == Bool) return b;
    if (kind == Number) return n != 0.0;
    throw Error("JSON value is not a boolean");
}

std::string Value::str() const {
    if (kind == String || kind == Number) return raw;
    if (kind == Bool) return b ? "true" : "false";
    throw Error("JSON value is not a string");
}

std::vector<double> Value::numbers() const {
    std::vector<double> out;
    if (kind != Array) throw Error("JSON value is not an array");
    for (auto& v : arr) out.push_back(v.num());
    return out;
}

std::vector<std::string> Value::strings() const {
    std::vector<std::string> out;
    if (kind != Array) throw Error("JSON value is not an array");
    for (auto& v : arr) out.push_back(v.str());
    return out;
}

Value parse(const std::string& text, const std::string& what) {
    Parser ps{text.data(), text.data(), text.data() + text.size(), what};
    Value v = ps.value();
    ps.ws();
    if (ps.p != ps.end) ps.fail("trailing characters");
    return v;
}

Value parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open input file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), path);
}

}  // namespace json
}  // namespace ps
