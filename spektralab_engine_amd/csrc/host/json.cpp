// Strict JSON reader and compact writer; behaviour stated in json.h.
#include "json.h"

#include <clocale>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <locale.h>

namespace spk::json {

const Value* Value::find(std::string_view key) const {
    if (!is_object()) return nullptr;
    const Object& o = as_object();
    auto it = o.find(key);
    return it == o.end() ? nullptr : &it->second;
}

Value& Value::operator[](std::string_view key) {
    if (is_null()) v_ = Object{};
    Object& o = as_object();
    auto it = o.find(key);
    if (it == o.end()) it = o.emplace(std::string(key), Value()).first;
    return it->second;
}

void Value::push_back(Value v) {
    if (is_null()) v_ = Array{};
    as_array().push_back(std::move(v));
}

namespace {

locale_t c_locale() {
    static locale_t loc = newlocale(LC_ALL_MASK, "C", nullptr);
    return loc;
}

class Parser {
public:
    explicit Parser(std::string_view s) : s_(s) {}

    Value run() {
        check_utf8();
        ws();
        Value v = value(0);
        ws();
        if (i_ != s_.size()) err("unexpected data after the JSON value");
        return v;
    }

private:
    std::string_view s_;
    size_t i_ = 0;

    [[noreturn]] void err(const std::string& m) const { throw ParseError(i_, m); }
    [[noreturn]] void err_at(size_t at, const std::string& m) const { throw ParseError(at, m); }
    bool eof() const { return i_ >= s_.size(); }

    void check_utf8() const {
        size_t i = 0;
        const auto* p = reinterpret_cast<const unsigned char*>(s_.data());
        while (i < s_.size()) {
            const unsigned c = p[i];
            size_t n = 0;
            unsigned cp = 0;
            if (c < 0x80) { ++i; continue; }
            if ((c & 0xE0) == 0xC0) { n = 1; cp = c & 0x1F; }
            else if ((c & 0xF0) == 0xE0) { n = 2; cp = c & 0x0F; }
            else if ((c & 0xF8) == 0xF0) { n = 3; cp = c & 0x07; }
            else throw ParseError(i, "invalid UTF-8");
            for (size_t k = 1; k <= n; ++k) {
                if (i + k >= s_.size() || (p[i + k] & 0xC0) != 0x80) throw ParseError(i, "invalid UTF-8");
                cp = (cp << 6) | (p[i + k] & 0x3F);
            }
            if ((n == 1 && cp < 0x80) || (n == 2 && cp < 0x800) || (n == 3 && cp < 0x10000) || cp > 0x10FFFF ||
                (cp >= 0xD800 && cp <= 0xDFFF))
                throw ParseError(i, "invalid UTF-8");
            i += n + 1;
        }
    }

    void ws() {
        while (!eof() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }

    void expect_word(const char* w) {
        for (const char* p = w; *p; ++p) {
            if (eof() || s_[i_] != *p) err(std::string("invalid literal, expected '") + w + "'");
            ++i_;
        }
    }

    Value value(int depth) {
        if (eof()) err("unexpected end of input");
        const char c = s_[i_];
        if (c == '{') return object(depth + 1);
        if (c == '[') return array(depth + 1);
        if (c == '"') return Value(string());
        if (c == 't') { expect_word("true"); return Value(true); }
        if (c == 'f') { expect_word("false"); return Value(false); }
        if (c == 'n') { expect_word("null"); return Value(); }
        if (c == '-' || (c >= '0' && c <= '9')) return Value(number());
        err("unexpected character");
    }

    Value object(int depth) {
        if (depth > 64) err("nesting deeper than 64");
        ++i_;
        Object o;
        ws();
        if (!eof() && s_[i_] == '}') { ++i_; return Value(std::move(o)); }
        for (;;) {
            ws();
            if (eof() || s_[i_] != '"') err("expected a string key");
            std::string k = string();
            ws();
            if (eof() || s_[i_] != ':') err("expected ':'");
            ++i_;
            ws();
            Value v = value(depth);
            o[k] = std::move(v);
            ws();
            if (eof()) err("unexpected end of input in an object");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == '}') { ++i_; break; }
            err("expected ',' or '}'");
        }
        return Value(std::move(o));
    }

    Value array(int depth) {
        if (depth > 64) err("nesting deeper than 64");
        ++i_;
        Array a;
        ws();
        if (!eof() && s_[i_] == ']') { ++i_; return Value(std::move(a)); }
        for (;;) {
            ws();
            a.push_back(value(depth));
            ws();
            if (eof()) err("unexpected end of input in an array");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == ']') { ++i_; break; }
            err("expected ',' or ']'");
        }
        return Value(std::move(a));
    }

    static void put_utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) o += char(cp);
        else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        } else {
            o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F));
            o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        }
    }

    unsigned hex4() {
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            if (eof()) err("truncated \\u escape");
            const char c = s_[i_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= unsigned(c - '0');
            else if (c >= 'a' && c <= 'f') v |= unsigned(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= unsigned(c - 'A' + 10);
            else err("invalid \\u escape");
        }
        return v;
    }

    std::string string() {
        ++i_;
        std::string o;
        for (;;) {
            if (eof()) err("unterminated string");
            const unsigned char c = static_cast<unsigned char>(s_[i_]);
            if (c == '"') { ++i_; break; }
            if (c < 0x20) err("control character in a string");
            if (c != '\\') { o += char(c); ++i_; continue; }
            ++i_;
            if (eof()) err("unterminated escape");
            const char e = s_[i_++];
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (i_ + 1 >= s_.size() || s_[i_] != '\\' || s_[i_ + 1] != 'u') err("unpaired surrogate");
                        i_ += 2;
                        const unsigned lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) err("unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                        err("unpaired surrogate");
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: err("invalid escape");
            }
        }
        return o;
    }

    double number() {
        const size_t start = i_;
        if (s_[i_] == '-') ++i_;
        if (eof()) err("invalid number");
        if (s_[i_] == '0') {
            ++i_;
        } else if (s_[i_] >= '1' && s_[i_] <= '9') {
            while (!eof() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        } else {
            err("invalid number");
        }
        if (!eof() && s_[i_] == '.') {
            ++i_;
            if (eof() || s_[i_] < '0' || s_[i_] > '9') err("invalid number: digit expected after '.'");
            while (!eof() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        }
        if (!eof() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            ++i_;
            if (!eof() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (eof() || s_[i_] < '0' || s_[i_] > '9') err("invalid number: digit expected in exponent");
            while (!eof() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        }
        const std::string tmp(s_.substr(start, i_ - start));
        char* end = nullptr;
        const double v = strtod_l(tmp.c_str(), &end, c_locale());
        if (end != tmp.c_str() + tmp.size()) err_at(start, "invalid number");
        if (std::isinf(v)) err_at(start, "number out of range");
        return v;
    }
};

void append_string(std::string& out, const std::string& s) {
    out += '"';
    for (const char ch : s) {
        const unsigned char c = static_cast<unsigned char>(ch);
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", c);
                    out += buf;
                } else {
                    out += ch;
                }
        }
    }
    out += '"';
}

void append_value(std::string& out, const Value& v) {
    switch (v.type()) {
        case Type::Null: out += "null"; break;
        case Type::Bool: out += v.as_bool() ? "true" : "false"; break;
        case Type::Number: out += format_number(v.as_number()); break;
        case Type::String: append_string(out, v.as_string()); break;
        case Type::Array: {
            out += '[';
            bool first = true;
            for (const Value& e : v.as_array()) {
                if (!first) out += ',';
                first = false;
                append_value(out, e);
            }
            out += ']';
            break;
        }
        case Type::Object: {
            out += '{';
            bool first = true;
            for (const auto& [k, e] : v.as_object()) {  // std::map: ascending bytewise order
                if (!first) out += ',';
                first = false;
                append_string(out, k);
                out += ':';
                append_value(out, e);
            }
            out += '}';
            break;
        }
    }
}

}  // namespace

Value parse(std::string_view text) { return Parser(text).run(); }

std::string format_number(double d) {
    if (!std::isfinite(d)) return "null";
    char buf[40];
    const locale_t old = uselocale(c_locale());
    std::snprintf(buf, sizeof buf, "%.17g", d);
    uselocale(old);
    return buf;
}

std::string write(const Value& v) {
    std::string out;
    append_value(out, v);
    return out;
}

}  // namespace spk::json
