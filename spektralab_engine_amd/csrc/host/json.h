// A small JSON document model for the engine's option strings, parameter deltas, resource files
// and replies.
//
// Reading is strict RFC 8259: one value, whitespace around it, no comments, no NaN or Infinity
// literals, strings valid UTF-8 after escapes, at most kMaxDepth nested containers.  A repeated
// object key keeps its last value.  Writing is compact, objects in ascending bytewise key order,
// numbers in the shortest form that reads back to the same double (non-finite numbers as null).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace spk::json {

inline constexpr int kMaxDepth = 64;

class Value;
using Member = std::pair<std::string, Value>;

class Value {
public:
    enum class Kind : unsigned char { Null, Boolean, Number, Text, List, Map };

    Value() = default;
    Value(std::nullptr_t) {}
    Value(bool b) : kind_(Kind::Boolean), flag_(b) {}
    Value(double d) : kind_(Kind::Number), num_(d) {}
    Value(int i) : Value(double(i)) {}
    Value(unsigned i) : Value(double(i)) {}
    Value(long i) : Value(double(i)) {}
    Value(unsigned long i) : Value(double(i)) {}
    Value(long long i) : Value(double(i)) {}
    Value(unsigned long long i) : Value(double(i)) {}
    Value(const char* s) : kind_(Kind::Text), text_(s) {}
    Value(std::string s) : kind_(Kind::Text), text_(std::move(s)) {}
    Value(std::string_view s) : kind_(Kind::Text), text_(s) {}

    static Value list(std::vector<Value> items = {});
    static Value map();

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Kind::Null; }
    bool is_bool() const { return kind_ == Kind::Boolean; }
    bool is_number() const { return kind_ == Kind::Number; }
    bool is_text() const { return kind_ == Kind::Text; }
    bool is_list() const { return kind_ == Kind::List; }
    bool is_map() const { return kind_ == Kind::Map; }

    // Typed reads; a read of the wrong kind throws std::logic_error.
    bool boolean() const;
    double number() const;
    const std::string& text() const;
    const std::vector<Value>& items() const;     // list elements
    const std::vector<Member>& members() const;  // map members, ascending by key

    // Maps: the member's value or null when absent (or when this is not a map).
    const Value* get(std::string_view key) const;
    // Maps (a null value becomes an empty map): inserts or replaces, keeping keys sorted.
    Value& set(std::string_view key, Value v);
    // Lists (a null value becomes an empty list).
    Value& append(Value v);
    size_t size() const;  // elements of a list, members of a map, else 0

    bool operator==(const Value& o) const;

private:
    Kind kind_ = Kind::Null;
    bool flag_ = false;
    double num_ = 0.0;
    std::string text_;
    std::vector<Value> list_;
    std::vector<Member> map_;
};

// A syntax error with the byte offset where reading stopped.
class SyntaxError : public std::runtime_error {
public:
    SyntaxError(size_t at, const std::string& why);
    size_t at() const { return at_; }

private:
    size_t at_;
};

Value parse(std::string_view text);
std::string dump(const Value& v);
// The shortest decimal form that reads back to d; "null" when d is not finite.
std::string number_text(double d);

}  // namespace spk::json
