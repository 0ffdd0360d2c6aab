// The JSON document model; rules stated in json.h.
#include "json.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace spk::json {

// ---- the model ------------------------------------------------------------------------------
Value Value::list(std::vector<Value> items) {
    Value v;
    v.kind_ = Kind::List;
    v.list_ = std::move(items);
    return v;
}

Value Value::map() {
    Value v;
    v.kind_ = Kind::Map;
    return v;
}

namespace {
[[noreturn]] void wrong_kind(const char* wanted) { throw std::logic_error(std::string("JSON value is not ") + wanted); }

bool key_less(const Member& m, std::string_view k) { return std::string_view(m.first) < k; }
}  // namespace

bool Value::boolean() const {
    if (kind_ != Kind::Boolean) wrong_kind("a boolean");
    return flag_;
}
double Value::number() const {
    if (kind_ != Kind::Number) wrong_kind("a number");
    return num_;
}
const std::string& Value::text() const {
    if (kind_ != Kind::Text) wrong_kind("a string");
    return text_;
}
const std::vector<Value>& Value::items() const {
    if (kind_ != Kind::List) wrong_kind("an array");
    return list_;
}
const std::vector<Member>& Value::members() const {
    if (kind_ != Kind::Map) wrong_kind("an object");
    return map_;
}

const Value* Value::get(std::string_view key) const {
    if (kind_ != Kind::Map) return nullptr;
    auto it = std::lower_bound(map_.begin(), map_.end(), key, key_less);
    return (it != map_.end() && it->first == key) ? &it->second : nullptr;
}

Value& Value::set(std::string_view key, Value v) {
    if (kind_ == Kind::Null) kind_ = Kind::Map;
    if (kind_ != Kind::Map) wrong_kind("an object");
    auto it = std::lower_bound(map_.begin(), map_.end(), key, key_less);
    if (it != map_.end() && it->first == key) {
        it->second = std::move(v);
        return it->second;
    }
    return map_.insert(it, Member(std::string(key), std::move(v)))->second;
}

Value& Value::append(Value v) {
    if (kind_ == Kind::Null) kind_ = Kind::List;
    if (kind_ != Kind::List) wrong_kind("an array");
    list_.push_back(std::move(v));
    return list_.back();
}

size_t Value::size() const { return kind_ == Kind::List ? list_.size() : kind_ == Kind::Map ? map_.size() : 0; }

bool Value::operator==(const Value& o) const {
    if (kind_ != o.kind_) return false;
    switch (kind_) {
        case Kind::Null: return true;
        case Kind::Boolean: return flag_ == o.flag_;
        case Kind::Number: return num_ == o.num_;
        case Kind::Text: return text_ == o.text_;
        case Kind::List: return list_ == o.list_;
        case Kind::Map: return map_ == o.map_;
    }
    return false;
}

SyntaxError::SyntaxError(size_t at, const std::string& why)
    : std::runtime_error("invalid JSON at offset " + std::to_string(at) + ": " + why), at_(at) {}

// ---- reading ---------------------------------------------------------------------------------
namespace {

struct Reader {
    std::string_view src;
    size_t pos = 0;

    [[noreturn]] void stop(const std::string& why) const { throw SyntaxError(pos, why); }
    bool done() const { return pos >= src.size(); }
    char peek() const { return done() ? '\0' : src[pos]; }

    void blanks() {
        while (!done()) {
            const char c = src[pos];
            if (c != ' ' && c != '\t' && c != '\n' && c != '\r') break;
            ++pos;
        }
    }

    void word(std::string_view w) {
        if (src.substr(pos, w.size()) != w) stop("unrecognised token");
        pos += w.size();
    }

    Value value(int depth) {
        blanks();
        switch (peek()) {
            case '{': return map(depth + 1);
            case '[': return list(depth + 1);
            case '"': return Value(text());
            case 't': word("true"); return Value(true);
            case 'f': word("false"); return Value(false);
            case 'n': word("null"); return Value();
            case '\0':
                if (done()) stop("the text ends where a value should begin");
                stop("unexpected character");
            default: return Value(number());
        }
    }

    Value map(int depth) {
        if (depth > kMaxDepth) stop("containers nested more than " + std::to_string(kMaxDepth) + " deep");
        ++pos;  // '{'
        Value out = Value::map();
        blanks();
        if (peek() == '}') {
            ++pos;
            return out;
        }
        for (;;) {
            blanks();
            if (peek() != '"') stop("an object key must be a string");
            std::string key = text();
            blanks();
            if (peek() != ':') stop("':' expected after an object key");
            ++pos;
            out.set(key, value(depth));
            blanks();
            const char c = peek();
            ++pos;
            if (c == '}') return out;
            if (c != ',') {
                --pos;
                stop("',' or '}' expected in an object");
            }
        }
    }

    Value list(int depth) {
        if (depth > kMaxDepth) stop("containers nested more than " + std::to_string(kMaxDepth) + " deep");
        ++pos;  // '['
        Value out = Value::list();
        blanks();
        if (peek() == ']') {
            ++pos;
            return out;
        }
        for (;;) {
            out.append(value(depth));
            blanks();
            const char c = peek();
            ++pos;
            if (c == ']') return out;
            if (c != ',') {
                --pos;
                stop("',' or ']' expected in an array");
            }
        }
    }

    // Numbers: -? (0 | [1-9][0-9]*) (. [0-9]+)? ([eE] [+-]? [0-9]+)?
    double number() {
        const size_t start = pos;
        auto digits = [&] {
            const size_t d0 = pos;
            while (!done() && src[pos] >= '0' && src[pos] <= '9') ++pos;
            return pos - d0;
        };
        if (peek() == '-') ++pos;
        if (peek() == '0') {
            ++pos;
        } else if (digits() == 0) {
            pos = start;
            stop("unexpected character");
        }
        if (peek() == '.') {
            ++pos;
            if (digits() == 0) stop("a digit must follow the decimal point");
        }
        if (peek() == 'e' || peek() == 'E') {
            ++pos;
            if (peek() == '+' || peek() == '-') ++pos;
            if (digits() == 0) stop("an exponent needs digits");
        }
        double d = 0.0;
        const std::string_view lexeme = src.substr(start, pos - start);
        const auto r = std::from_chars(lexeme.data(), lexeme.data() + lexeme.size(), d);
        if (r.ec == std::errc::result_out_of_range) {
            // beyond the doubles: the nearest one, +-inf for a large magnitude, +-0 for a tiny one
            const size_t e = lexeme.find_first_of("eE");
            const bool tiny = e != std::string_view::npos && e + 1 < lexeme.size() && lexeme[e + 1] == '-';
            const bool neg = lexeme.front() == '-';
            d = tiny ? (neg ? -0.0 : 0.0) : (neg ? -HUGE_VAL : HUGE_VAL);
        } else if (r.ec != std::errc()) {
            stop("malformed number");
        }
        return d;
    }

    uint32_t hex4() {
        if (src.size() - pos < 4) stop("\\u needs four hex digits");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = src[pos++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else stop("\\u needs four hex digits");
        }
        return v;
    }

    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }

    // One raw UTF-8 sequence, validated (no overlongs, no surrogates, at most U+10FFFF).
    void raw_utf8(std::string& out) {
        const unsigned char c = static_cast<unsigned char>(src[pos]);
        int extra = 0;
        uint32_t cp = 0, lo = 0;
        if (c >= 0xC2 && c <= 0xDF) {
            extra = 1;
            cp = c & 0x1F;
            lo = 0x80;
        } else if (c >= 0xE0 && c <= 0xEF) {
            extra = 2;
            cp = c & 0x0F;
            lo = 0x800;
        } else if (c >= 0xF0 && c <= 0xF4) {
            extra = 3;
            cp = c & 0x07;
            lo = 0x10000;
        } else {
            stop("invalid UTF-8 in a string");
        }
        if (src.size() - pos <= size_t(extra)) stop("invalid UTF-8 in a string");
        for (int i = 1; i <= extra; ++i) {
            const unsigned char d = static_cast<unsigned char>(src[pos + size_t(i)]);
            if ((d & 0xC0) != 0x80) stop("invalid UTF-8 in a string");
            cp = (cp << 6) | (d & 0x3F);
        }
        if (cp < lo || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) stop("invalid UTF-8 in a string");
        out.append(src.data() + pos, size_t(extra) + 1);
        pos += size_t(extra) + 1;
    }

    std::string text() {
        ++pos;  // '"'
        std::string out;
        for (;;) {
            if (done()) stop("a string is not closed");
            const unsigned char c = static_cast<unsigned char>(src[pos]);
            if (c == '"') {
                ++pos;
                return out;
            }
            if (c < 0x20) stop("a control character inside a string");
            if (c >= 0x80) {
                raw_utf8(out);
                continue;
            }
            if (c != '\\') {
                out += char(c);
                ++pos;
                continue;
            }
            ++pos;
            if (done()) stop("a string is not closed");
            const char e = src[pos++];
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
                    if (cp >= 0xDC00 && cp <= 0xDFFF) stop("a low surrogate without a high one");
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (src.substr(pos, 2) != "\\u") stop("a high surrogate without a low one");
                        pos += 2;
                        const uint32_t low = hex4();
                        if (low < 0xDC00 || low > 0xDFFF) stop("a high surrogate without a low one");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (low - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: --pos; stop("unknown escape in a string");
            }
        }
    }
};

}  // namespace

Value parse(std::string_view text) {
    Reader r{text};
    Value v = r.value(0);
    r.blanks();
    if (!r.done()) r.stop("text continues after the value");
    return v;
}

// ---- writing ---------------------------------------------------------------------------------
std::string number_text(double d) {
    if (!std::isfinite(d)) return "null";
    if (d == 0.0) return std::signbit(d) ? "-0" : "0";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, d);
    return std::string(buf, r.ptr);
}

namespace {

void quote(std::string& out, const std::string& s) {
    static const char* hex = "0123456789abcdef";
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
                    out += "\\u00";
                    out += hex[c >> 4];
                    out += hex[c & 15];
                } else {
                    out += ch;
                }
        }
    }
    out += '"';
}

void emit(std::string& out, const Value& v) {
    switch (v.kind()) {
        case Value::Kind::Null: out += "null"; break;
        case Value::Kind::Boolean: out += v.boolean() ? "true" : "false"; break;
        case Value::Kind::Number: out += number_text(v.number()); break;
        case Value::Kind::Text: quote(out, v.text()); break;
        case Value::Kind::List: {
            out += '[';
            bool first = true;
            for (const Value& e : v.items()) {
                if (!first) out += ',';
                first = false;
                emit(out, e);
            }
            out += ']';
            break;
        }
        case Value::Kind::Map: {
            out += '{';
            bool first = true;
            for (const Member& m : v.members()) {
                if (!first) out += ',';
                first = false;
                quote(out, m.first);
                out += ':';
                emit(out, m.second);
            }
            out += '}';
            break;
        }
    }
}

}  // namespace

std::string dump(const Value& v) {
    std::string out;
    emit(out, v);
    return out;
}

}  // namespace spk::json
