// json.hpp -- a small JSON reader for the scene files.
//
// The reference parses scenes with nlohmann/json 3.11.3 (scene/src/loader.cpp,
// fetched at configure time by third_party/CMakeLists.txt:16-18).  This reader
// covers RFC 8259 values (objects, arrays, strings with escapes, numbers,
// true/false/null).  As with nlohmann's stream extraction, parsing stops after
// the first complete value; duplicate object keys keep the last value.
#pragma once

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace wsim {
namespace json {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct TypeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Value {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    bool is_integer = false;
    long long integer = 0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool contains(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
    const Value& at(const std::string& k) const {
        if (kind != Object) throw TypeError("json: not an object");
        auto it = obj.find(k);
        if (it == obj.end()) throw TypeError("json: missing key " + k);
        return it->second;
    }
    const Value& at(size_t i) const {
        if (kind != Array) throw TypeError("json: not an array");
        if (i >= arr.size()) throw TypeError("json: array index out of range");
        return arr[i];
    }
    float as_float() const {
        if (kind != Number) throw TypeError("json: value is not a number");
        return is_integer ? static_cast<float>(integer) : static_cast<float>(num);
    }
    int32_t as_int() const {
        if (kind != Number) throw TypeError("json: value is not a number");
        return is_integer ? static_cast<int32_t>(integer) : static_cast<int32_t>(num);
    }
    const std::string& as_string() const {
        if (kind != String) throw TypeError("json: value is not a string");
        return str;
    }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        ws();
        Value v = value(0);
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void err(const std::string& m) const {
        throw ParseError("json parse error at offset " + std::to_string(i_) + ": " + m);
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r'))
            ++i_;
    }
    char peek() const { return i_ < s_.size() ? s_[i_] : '\0'; }
    void expect(char c) {
        if (peek() != c) err(std::string("expected '") + c + "'");
        ++i_;
    }
    void literal(const char* w) {
        for (const char* p = w; *p; ++p) {
            if (peek() != *p) err(std::string("bad literal, expected ") + w);
            ++i_;
        }
    }
    Value value(int depth) {
        if (depth > 256) err("nesting too deep");
        if (i_ >= s_.size()) err("unexpected end of input");
        Value v;
        const char c = peek();
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            ws();
            if (peek() == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                if (peek() != '"') err("expected object key");
                std::string k = string_lit();
                ws();
                expect(':');
                ws();
                v.obj[k] = value(depth + 1);
                ws();
                if (peek() == ',') {
                    ++i_;
                    continue;
                }
                expect('}');
                break;
            }
        } else if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            ws();
            if (peek() == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                v.arr.push_back(value(depth + 1));
                ws();
                if (peek() == ',') {
                    ++i_;
                    continue;
                }
                expect(']');
                break;
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
        } else if (c == 't') {
            literal("true");
            v.kind = Value::Bool;
            v.b = true;
        } else if (c == 'f') {
            literal("false");
            v.kind = Value::Bool;
            v.b = false;
        } else if (c == 'n') {
            literal("null");
        } else {
            number(v);
        }
        return v;
    }
    void number(Value& v) {
        const size_t st = i_;
        bool integral = true;
        if (peek() == '-') ++i_;
        if (peek() == '0') {
            ++i_;
        } else if (peek() >= '1' && peek() <= '9') {
            while (peek() >= '0' && peek() <= '9') ++i_;
        } else {
            err("invalid value");
        }
        if (peek() == '.') {
            integral = false;
            ++i_;
            if (!(peek() >= '0' && peek() <= '9')) err("digit expected after '.'");
            while (peek() >= '0' && peek() <= '9') ++i_;
        }
        if (peek() == 'e' || peek() == 'E') {
            integral = false;
            ++i_;
            if (peek() == '+' || peek() == '-') ++i_;
            if (!(peek() >= '0' && peek() <= '9')) err("digit expected in exponent");
            while (peek() >= '0' && peek() <= '9') ++i_;
        }
        const std::string tok = s_.substr(st, i_ - st);
        v.kind = Value::Number;
        v.num = std::strtod(tok.c_str(), nullptr);
        if (integral) {
            errno = 0;
            const long long ll = std::strtoll(tok.c_str(), nullptr, 10);
            if (errno == 0) {
                v.is_integer = true;
                v.integer = ll;
            }
        }
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = peek();
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
            else err("bad \\u escape");
            ++i_;
        }
        return v;
    }
    std::string string_lit() {
        expect('"');
        std::string out;
        for (;;) {
            if (i_ >= s_.size()) err("unterminated string");
            const char c = s_[i_++];
            if (c == '"') break;
            if (static_cast<unsigned char>(c) < 0x20) err("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) err("unterminated escape");
            const char e = s_[i_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (peek() != '\\') err("lone high surrogate");
                        ++i_;
                        if (peek() != 'u') err("lone high surrogate");
                        ++i_;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) err("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: err("bad escape");
            }
        }
        return out;
    }
};

}  // namespace json
}  // namespace wsim
