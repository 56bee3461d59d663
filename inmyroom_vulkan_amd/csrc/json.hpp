// A small JSON reader (RFC 8259) for glTF documents: a DOM of values with
// objects kept as key/value vectors in document order.  Host code only.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace cdk {
namespace json {

struct Value {
    enum Kind : uint8_t { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value* find(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    // an integer that int64_t holds exactly (|x| < 2^53 keeps the cast defined)
    bool is_int() const {
        return kind == Number && num > -9007199254740992.0 && num < 9007199254740992.0 && num == double(int64_t(num));
    }
};

class Parser {
  public:
    Parser(const char* p, size_t n) : s_(p), end_(p + n), beg_(p) {}

    bool parse(Value& out, std::string& err) {
        ws();
        if (!value(out, 0)) {
            err = "JSON: " + err_ + " at byte " + std::to_string(size_t(s_ - beg_));
            return false;
        }
        ws();
        if (s_ != end_) {
            err = "JSON: trailing characters at byte " + std::to_string(size_t(s_ - beg_));
            return false;
        }
        return true;
    }

  private:
    const char* s_;
    const char* end_;
    const char* beg_;
    std::string err_;

    bool fail(const char* m) {
        err_ = m;
        return false;
    }
    void ws() {
        while (s_ < end_ && (*s_ == ' ' || *s_ == '\t' || *s_ == '\n' || *s_ == '\r')) ++s_;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (size_t(end_ - s_) < n || std::memcmp(s_, w, n) != 0) return fail("invalid literal");
        s_ += n;
        return true;
    }
    bool value(Value& v, int depth) {
        if (depth > 256) return fail("nesting too deep");
        if (s_ >= end_) return fail("unexpected end");
        switch (*s_) {
            case '{': return object(v, depth);
            case '[': return array(v, depth);
            case '"': v.kind = Value::String; return string(v.str);
            case 't': v.kind = Value::Bool; v.b = true; return lit("true");
            case 'f': v.kind = Value::Bool; v.b = false; return lit("false");
            case 'n': v.kind = Value::Null; return lit("null");
            default: return number(v);
        }
    }
    bool object(Value& v, int depth) {
        v.kind = Value::Object;
        ++s_;
        ws();
        if (s_ < end_ && *s_ == '}') { ++s_; return true; }
        for (;;) {
            ws();
            if (s_ >= end_ || *s_ != '"') return fail("expected a key");
            std::string key;
            if (!string(key)) return false;
            ws();
            if (s_ >= end_ || *s_ != ':') return fail("expected ':'");
            ++s_;
            ws();
            v.obj.emplace_back(std::move(key), Value());
            if (!value(v.obj.back().second, depth + 1)) return false;
            ws();
            if (s_ < end_ && *s_ == ',') { ++s_; continue; }
            if (s_ < end_ && *s_ == '}') { ++s_; return true; }
            return fail("expected ',' or '}'");
        }
    }
    bool array(Value& v, int depth) {
        v.kind = Value::Array;
        ++s_;
        ws();
        if (s_ < end_ && *s_ == ']') { ++s_; return true; }
        for (;;) {
            ws();
            v.arr.emplace_back();
            if (!value(v.arr.back(), depth + 1)) return false;
            ws();
            if (s_ < end_ && *s_ == ',') { ++s_; continue; }
            if (s_ < end_ && *s_ == ']') { ++s_; return true; }
            return fail("expected ',' or ']'");
        }
    }
    static void put_utf8(std::string& o, uint32_t c) {
        if (c < 0x80) o += char(c);
        else if (c < 0x800) { o += char(0xC0 | (c >> 6)); o += char(0x80 | (c & 63)); }
        else if (c < 0x10000) { o += char(0xE0 | (c >> 12)); o += char(0x80 | ((c >> 6) & 63)); o += char(0x80 | (c & 63)); }
        else {
            o += char(0xF0 | (c >> 18));
            o += char(0x80 | ((c >> 12) & 63));
            o += char(0x80 | ((c >> 6) & 63));
            o += char(0x80 | (c & 63));
        }
    }
    bool hex4(uint32_t& c) {
        if (end_ - s_ < 4) return fail("short \\u escape");
        c = 0;
        for (int i = 0; i < 4; ++i) {
            const char h = *s_++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') c |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= uint32_t(h - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& o) {
        ++s_;  // opening quote
        for (;;) {
            if (s_ >= end_) return fail("unterminated string");
            const char c = *s_++;
            if (c == '"') return true;
            if (static_cast<unsigned char>(c) < 0x20) return fail("control character in string");
            if (c != '\\') { o += c; continue; }
            if (s_ >= end_) return fail("unterminated escape");
            const char e = *s_++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                        uint32_t lo;
                        if (end_ - s_ < 2 || s_[0] != '\\' || s_[1] != 'u') return fail("lone high surrogate");
                        s_ += 2;
                        if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: return fail("bad escape");
            }
        }
    }
    bool number(Value& v) {
        const char* p = s_;
        if (p < end_ && *p == '-') ++p;
        if (p >= end_ || !(*p >= '0' && *p <= '9')) return fail("invalid value");
        if (*p == '0') ++p;
        else while (p < end_ && *p >= '0' && *p <= '9') ++p;
        if (p < end_ && *p == '.') {
            ++p;
            if (p >= end_ || !(*p >= '0' && *p <= '9')) return fail("bad fraction");
            while (p < end_ && *p >= '0' && *p <= '9') ++p;
        }
        if (p < end_ && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p < end_ && (*p == '+' || *p == '-')) ++p;
            if (p >= end_ || !(*p >= '0' && *p <= '9')) return fail("bad exponent");
            while (p < end_ && *p >= '0' && *p <= '9') ++p;
        }
        const std::string tok(s_, p);  // strtod needs a terminated copy
        v.kind = Value::Number;
        v.num = std::strtod(tok.c_str(), nullptr);
        // A number written as an integer is an integer, as the JSON libraries
        // glTF loaders use keep it: "-0" is the integer 0, read as +0.0.
        if (tok.find_first_of(".eE") == std::string::npos && v.num == 0.0) v.num = 0.0;
        s_ = p;
        return true;
    }
};

inline bool parse(const char* p, size_t n, Value& out, std::string& err) { return Parser(p, n).parse(out, err); }

}  // namespace json
}  // namespace cdk
