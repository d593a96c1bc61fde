// Minimal JSON reader for MRay scene files ("jsonc": JSON plus // and /* */
// comments, no trailing commas).  Numbers are kept as double; objects keep
// their key order.  Errors throw MRayError with the byte offset.
#pragma once

#include "mray_tracer.hpp"

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <string_view>
#include <vector>

namespace mray_amd::json
{

struct Value
{
    enum class T { Null, Bool, Num, Str, Arr, Obj } t = T::Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Value> a;
    std::vector<std::pair<std::string, Value>> o;

    bool IsNull() const { return t == T::Null; }
    bool IsArr() const { return t == T::Arr; }
    bool IsObj() const { return t == T::Obj; }
    bool IsStr() const { return t == T::Str; }
    bool IsNum() const { return t == T::Num; }
    const Value* Find(std::string_view key) const
    {
        if(t != T::Obj) return nullptr;
        for(const auto& kv : o)
            if(kv.first == key) return &kv.second;
        return nullptr;
    }
    const Value& At(std::string_view key) const
    {
        const Value* v = Find(key);
        if(!v) throw MRayError("scene: missing field \"" + std::string(key) + "\"");
        return *v;
    }
    double Num() const
    {
        if(t == T::Bool) return b ? 1.0 : 0.0;
        if(t != T::Num) throw MRayError("scene: expected a number");
        return n;
    }
    // A non-negative integer that fits in 32 bits (IDs, indices, counts); a
    // fraction, a negative, huge or non-finite value throws instead of being
    // cast (the cast would be undefined behaviour).
    uint32_t U32() const
    {
        double d = Num();
        if(!(d >= 0.0 && d <= 4294967295.0) || d != double(uint64_t(d)))
            throw MRayError("scene: expected an unsigned 32-bit integer");
        return uint32_t(d);
    }
    // A finite number within float range (positions, colours, angles).
    float F32() const
    {
        double d = Num();
        if(!(d >= -3.4028234663852886e38 && d <= 3.4028234663852886e38))
            throw MRayError("scene: expected a finite 32-bit float");
        return float(d);
    }
    bool Bool() const
    {
        if(t == T::Bool) return b;
        if(t == T::Num) return n != 0.0;
        throw MRayError("scene: expected a boolean");
    }
    const std::string& Str() const
    {
        if(t != T::Str) throw MRayError("scene: expected a string");
        return s;
    }
    const std::vector<Value>& Arr() const
    {
        if(t != T::Arr) throw MRayError("scene: expected an array");
        return a;
    }
};

class Parser
{
    std::string_view src;
    size_t p = 0;

    [[noreturn]] void Fail(const char* what) const
    {
        throw MRayError(std::string("scene json: ") + what + " at byte " + std::to_string(p));
    }
    void Skip()
    {
        for(;;)
        {
            while(p < src.size() && (src[p] == ' ' || src[p] == '\t' || src[p] == '\n' || src[p] == '\r'))
                p++;
            if(p + 1 < src.size() && src[p] == '/' && src[p + 1] == '/')
            {
                while(p < src.size() && src[p] != '\n') p++;
                continue;
            }
            if(p + 1 < src.size() && src[p] == '/' && src[p + 1] == '*')
            {
                size_t e = src.find("*/", p + 2);
                if(e == std::string_view::npos) Fail("unterminated comment");
                p = e + 2;
                continue;
            }
            return;
        }
    }
    std::string ParseString()
    {
        if(p >= src.size() || src[p] != '"') Fail("expected a string");
        p++;
        std::string out;
        while(p < src.size() && src[p] != '"')
        {
            char c = src[p++];
            if(c == '\\')
            {
                if(p >= src.size()) Fail("bad escape");
                char e = src[p++];
                switch(e)
                {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                    {
                        if(p + 4 > src.size()) Fail("bad \\u escape");
                        unsigned cp = unsigned(std::strtoul(std::string(src.substr(p, 4)).c_str(), nullptr, 16));
                        p += 4;
                        if(cp < 0x80) out += char(cp);
                        else if(cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                        else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: out += e; break;
                }
            }
            else out += c;
        }
        if(p >= src.size()) Fail("unterminated string");
        p++;
        return out;
    }
    Value ParseValue(int depth)
    {
        if(depth > 256) Fail("nesting too deep");
        Skip();
        if(p >= src.size()) Fail("unexpected end");
        Value v;
        char c = src[p];
        if(c == '{')
        {
            v.t = Value::T::Obj;
            p++;
            Skip();
            if(p < src.size() && src[p] == '}') { p++; return v; }
            for(;;)
            {
                Skip();
                std::string key = ParseString();
                Skip();
                if(p >= src.size() || src[p] != ':') Fail("expected ':'");
                p++;
                v.o.emplace_back(std::move(key), ParseValue(depth + 1));
                Skip();
                if(p < src.size() && src[p] == ',') { p++; continue; }
                if(p < src.size() && src[p] == '}') { p++; return v; }
                Fail("expected ',' or '}'");
            }
        }
        if(c == '[')
        {
            v.t = Value::T::Arr;
            p++;
            Skip();
            if(p < src.size() && src[p] == ']') { p++; return v; }
            for(;;)
            {
                v.a.push_back(ParseValue(depth + 1));
                Skip();
                if(p < src.size() && src[p] == ',') { p++; continue; }
                if(p < src.size() && src[p] == ']') { p++; return v; }
                Fail("expected ',' or ']'");
            }
        }
        if(c == '"') { v.t = Value::T::Str; v.s = ParseString(); return v; }
        if(src.compare(p, 4, "true") == 0) { p += 4; v.t = Value::T::Bool; v.b = true; return v; }
        if(src.compare(p, 5, "false") == 0) { p += 5; v.t = Value::T::Bool; v.b = false; return v; }
        if(src.compare(p, 4, "null") == 0) { p += 4; return v; }
        // number
        size_t s0 = p;
        while(p < src.size() && (std::isdigit(static_cast<unsigned char>(src[p])) || src[p] == '-' ||
                                 src[p] == '+' || src[p] == '.' || src[p] == 'e' || src[p] == 'E'))
            p++;
        if(p == s0) Fail("unexpected character");
        std::string num(src.substr(s0, p - s0));
        char* end = nullptr;
        v.t = Value::T::Num;
        v.n = std::strtod(num.c_str(), &end);
        if(!end || *end != '\0') Fail("bad number");
        return v;
    }

    public:
    explicit Parser(std::string_view text) : src(text) {}
    Value Parse()
    {
        Value v = ParseValue(0);
        Skip();
        if(p != src.size()) Fail("trailing characters");
        return v;
    }
};

inline Value Parse(std::string_view text) { return Parser(text).Parse(); }

} // namespace mray_amd::json
