// Minimal JSON reader for psim model files (the schema of psim/src/inputManager.cpp:14-115).
// Numbers are parsed with strtod, which rounds correctly like the reference's nlohmann_json.
#pragma once

#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace psim_json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value& at(const std::string& key) const {
        if (kind != Object) { throw std::runtime_error("JSON: expected an object looking up '" + key + "'"); }
        for (const auto& kv : obj) {
            if (kv.first == key) { return kv.second; }
        }
        throw std::runtime_error("JSON: missing key '" + key + "'");
    }
    bool contains(const std::string& key) const {
        if (kind != Object) { return false; }
        for (const auto& kv : obj) {
            if (kv.first == key) { return true; }
        }
        return false;
    }
    double number() const {
        if (kind != Number) { throw std::runtime_error("JSON: expected a number"); }
        return num;
    }
    const std::string& string() const {
        if (kind != String) { throw std::runtime_error("JSON: expected a string"); }
        return str;
    }
    const std::vector<Value>& array() const {
        if (kind != Array) { throw std::runtime_error("JSON: expected an array"); }
        return arr;
    }
};

class Parser {
public:
    explicit Parser(const char* text) : p_(text) {}

    Value parse() {
        Value v = value();
        ws();
        if (*p_ != '\0') { throw std::runtime_error("JSON: trailing characters"); }
        return v;
    }

private:
    const char* p_;

    void ws() {
        while (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t') { ++p_; }
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if (std::strncmp(p_, s, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Value value() {
        ws();
        Value v;
        if (*p_ == '{') {
            ++p_;
            v.kind = Value::Object;
            ws();
            if (*p_ == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (*p_ != '"') { throw std::runtime_error("JSON: expected a key"); }
                std::string k = str();
                ws();
                if (*p_++ != ':') { throw std::runtime_error("JSON: expected ':'"); }
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (*p_ == ',') {
                    ++p_;
                    continue;
                }
                if (*p_ == '}') {
                    ++p_;
                    break;
                }
                throw std::runtime_error("JSON: expected ',' or '}'");
            }
        } else if (*p_ == '[') {
            ++p_;
            v.kind = Value::Array;
            ws();
            if (*p_ == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (*p_ == ',') {
                    ++p_;
                    continue;
                }
                if (*p_ == ']') {
                    ++p_;
                    break;
                }
                throw std::runtime_error("JSON: expected ',' or ']'");
            }
        } else if (*p_ == '"') {
            v.kind = Value::String;
            v.str = str();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else {
            char* end = nullptr;
            v.num = std::strtod(p_, &end);
            if (end == p_) { throw std::runtime_error("JSON: unexpected character"); }
            v.kind = Value::Number;
            p_ = end;
        }
        return v;
    }
    std::string str() {
        ++p_;  // opening quote
        std::string out;
        while (*p_ && *p_ != '"') {
            if (*p_ == '\\') {
                ++p_;
                switch (*p_) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    unsigned code = 0;
                    for (int i = 0; i < 4; ++i) {
                        ++p_;
                        const char ch = *p_;
                        code = code * 16 + (unsigned)((ch >= '0' && ch <= '9') ? ch - '0'
                                                      : (ch >= 'a' && ch <= 'f') ? ch - 'a' + 10
                                                      : (ch >= 'A' && ch <= 'F') ? ch - 'A' + 10 : 0);
                    }
                    if (code < 0x80) {
                        out += (char)code;
                    } else if (code < 0x800) {
                        out += (char)(0xC0 | (code >> 6));
                        out += (char)(0x80 | (code & 0x3F));
                    } else {
                        out += (char)(0xE0 | (code >> 12));
                        out += (char)(0x80 | ((code >> 6) & 0x3F));
                        out += (char)(0x80 | (code & 0x3F));
                    }
                    break;
                }
                default: out += *p_;
                }
                ++p_;
            } else {
                out += *p_++;
            }
        }
        if (*p_ != '"') { throw std::runtime_error("JSON: unterminated string"); }
        ++p_;
        return out;
    }
};

inline Value parse(const char* text) { return Parser(text).parse(); }

}  // namespace psim_json
