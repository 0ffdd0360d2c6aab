// Strict JSON (RFC 8259) reader and compact writer.
//
// Wire behaviour follows the reference's src/util/json.hpp:1-30: strict reader (no NaN/Infinity,
// depth <= 64, duplicate keys: last wins, whitespace only after the value, byte offset in
// errors); writer with object keys in ascending bytewise order, numbers as "%.17g", non-finite
// numbers as null.
#pragma once

#include <cstddef>
#include <map>
#include <stdexcept>
#include <string>
#include <string_view>
#include <variant>
#include <vector>

namespace spk::json {

class Value;
using Array = std::vector<Value>;
using Object = std::map<std::string, Value, std::less<>>;

enum class Type : unsigned char { Null, Bool, Number, String, Array, Object };

class Value {
public:
    Value() = default;
    Value(std::nullptr_t) {}
    Value(bool b) : v_(b) {}
    Value(double d) : v_(d) {}
    Value(int i) : v_(double(i)) {}
    Value(unsigned i) : v_(double(i)) {}
    Value(long i) : v_(double(i)) {}
    Value(unsigned long i) : v_(double(i)) {}
    Value(long long i) : v_(double(i)) {}
    Value(unsigned long long i) : v_(double(i)) {}
    Value(const char* s) : v_(std::string(s)) {}
    Value(std::string s) : v_(std::move(s)) {}
    Value(std::string_view s) : v_(std::string(s)) {}
    Value(Array a) : v_(std::move(a)) {}
    Value(Object o) : v_(std::move(o)) {}

    Type type() const { return Type(v_.index()); }
    bool is_null() const { return type() == Type::Null; }
    bool is_bool() const { return type() == Type::Bool; }
    bool is_number() const { return type() == Type::Number; }
    bool is_string() const { return type() == Type::String; }
    bool is_array() const { return type() == Type::Array; }
    bool is_object() const { return type() == Type::Object; }

    bool as_bool() const { return std::get<bool>(v_); }
    double as_number() const { return std::get<double>(v_); }
    const std::string& as_string() const { return std::get<std::string>(v_); }
    const Array& as_array() const { return std::get<Array>(v_); }
    Array& as_array() { return std::get<Array>(v_); }
    const Object& as_object() const { return std::get<Object>(v_); }
    Object& as_object() { return std::get<Object>(v_); }

    // nullptr when this is not an object or the key is absent.
    const Value* find(std::string_view key) const;
    // A null value becomes an empty object first.
    Value& operator[](std::string_view key);
    void push_back(Value v);
    bool operator==(const Value& o) const { return v_ == o.v_; }

private:
    std::variant<std::monostate, bool, double, std::string, Array, Object> v_;
};

class ParseError : public std::runtime_error {
public:
    ParseError(size_t offset, const std::string& what)
        : std::runtime_error("JSON parse error at byte " + std::to_string(offset) + ": " + what), offset_(offset) {}
    size_t offset() const { return offset_; }

private:
    size_t offset_;
};

Value parse(std::string_view text);
std::string write(const Value& v);
std::string format_number(double d);

}  // namespace spk::json
