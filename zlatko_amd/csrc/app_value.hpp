// app_value.hpp -- the JSON value of the Spec/Value front door: what the analyses read
// (specifications) and return (results), the role of ccr::app::Value
// (include/ccr/app/value.hpp:18-146) with JSON text as its wire form. Host code only.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ccrhip {
namespace app {

class Value {
 public:
  enum class Kind { Undefined, Null, Bool, Number, String, Array, Object };

  Value() = default;
  static Value null() { return Value(Kind::Null); }
  static Value boolean(bool b) {
    Value v(Kind::Bool);
    v.num_ = b ? 1.0 : 0.0;
    return v;
  }
  static Value number(double x) {
    Value v(Kind::Number);
    v.num_ = x;
    return v;
  }
  static Value string(std::string s) {
    Value v(Kind::String);
    v.str_ = std::move(s);
    return v;
  }
  static Value array() { return Value(Kind::Array); }
  static Value object() { return Value(Kind::Object); }
  static Value numbers(const std::vector<double>& xs) {
    Value v(Kind::Array);
    v.items_.reserve(xs.size());
    for (double x : xs) v.items_.push_back(number(x));
    return v;
  }

  Kind kind() const { return kind_; }
  bool present() const { return kind_ != Kind::Undefined && kind_ != Kind::Null; }
  bool isString() const { return kind_ == Kind::String; }
  bool isArray() const { return kind_ == Kind::Array; }
  bool isObject() const { return kind_ == Kind::Object; }

  // Objects keep their keys in insertion order; setting an existing key replaces its value.
  Value& set(const std::string& key, Value v) {
    if (kind_ != Kind::Object) throw std::logic_error("Value::set on a non-object");
    for (auto& kv : members_)
      if (kv.first == key) {
        kv.second = std::move(v);
        return kv.second;
      }
    members_.emplace_back(key, std::move(v));
    return members_.back().second;
  }
  Value& set(const std::string& key, double x) { return set(key, number(x)); }
  Value& set(const std::string& key, const char* s) { return set(key, string(s)); }
  Value& set(const std::string& key, const std::string& s) { return set(key, string(s)); }
  Value& setBool(const std::string& key, bool b) { return set(key, boolean(b)); }
  Value& setNumbers(const std::string& key, const std::vector<double>& xs) { return set(key, numbers(xs)); }
  const Value& operator[](const std::string& key) const {
    if (kind_ == Kind::Object)
      for (const auto& kv : members_)
        if (kv.first == key) return kv.second;
    return undefined();
  }
  const std::vector<std::pair<std::string, Value>>& members() const { return members_; }

  void push(Value v) {
    if (kind_ != Kind::Array) throw std::logic_error("Value::push on a non-array");
    items_.push_back(std::move(v));
  }
  const Value& at(std::size_t i) const { return kind_ == Kind::Array && i < items_.size() ? items_[i] : undefined(); }
  std::size_t size() const {
    return kind_ == Kind::Array ? items_.size() : kind_ == Kind::Object ? members_.size() : 0;
  }

  double asNumber(const char* what = "value") const {
    if (kind_ != Kind::Number && kind_ != Kind::Bool) throw std::invalid_argument(std::string(what) + ": expected a number");
    return num_;
  }
  bool asBool(const char* what = "value") const {
    if (kind_ != Kind::Number && kind_ != Kind::Bool) throw std::invalid_argument(std::string(what) + ": expected a boolean");
    return num_ != 0.0;
  }
  const std::string& asString(const char* what = "value") const {
    if (kind_ != Kind::String) throw std::invalid_argument(std::string(what) + ": expected a string");
    return str_;
  }
  std::vector<double> asNumbers(const char* what = "value") const {
    if (kind_ != Kind::Array) throw std::invalid_argument(std::string(what) + ": expected an array of numbers");
    std::vector<double> out;
    out.reserve(items_.size());
    for (const auto& v : items_) out.push_back(v.asNumber(what));
    return out;
  }

  // JSON text <-> Value. Non-finite numbers are written as null (JSON has no NaN or infinity).
  static Value parse(const std::string& text);
  std::string dump() const {
    std::string out;
    write(out);
    return out;
  }

 private:
  explicit Value(Kind k) : kind_(k) {}
  static const Value& undefined() {
    static const Value u;
    return u;
  }
  void write(std::string& out) const;

  Kind kind_ = Kind::Undefined;
  double num_ = 0.0;
  std::string str_;
  std::vector<Value> items_;
  std::vector<std::pair<std::string, Value>> members_;
};

namespace detail {

class JsonReader {
 public:
  explicit JsonReader(const std::string& s) : s_(s) {}
  Value document() {
    Value v = value();
    space();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  [[noreturn]] void fail(const std::string& why) const {
    throw std::invalid_argument("spec JSON: " + why + " at offset " + std::to_string(i_));
  }
  void space() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
  }
  bool literal(const char* word) {
    std::size_t n = 0;
    while (word[n]) ++n;
    if (s_.compare(i_, n, word) != 0) return false;
    i_ += n;
    return true;
  }
  Value value() {
    space();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') return Value::string(text());
    if (literal("true")) return Value::boolean(true);
    if (literal("false")) return Value::boolean(false);
    if (literal("null")) return Value::null();
    return number();
  }
  Value object() {
    Value o = Value::object();
    ++i_;
    space();
    if (i_ < s_.size() && s_[i_] == '}') {
      ++i_;
      return o;
    }
    for (;;) {
      space();
      if (i_ >= s_.size() || s_[i_] != '"') fail("expected a key");
      std::string key = text();
      space();
      if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
      ++i_;
      o.set(key, value());
      space();
      if (i_ < s_.size() && s_[i_] == ',') {
        ++i_;
        continue;
      }
      if (i_ < s_.size() && s_[i_] == '}') {
        ++i_;
        return o;
      }
      fail("expected ',' or '}'");
    }
  }
  Value array() {
    Value a = Value::array();
    ++i_;
    space();
    if (i_ < s_.size() && s_[i_] == ']') {
      ++i_;
      return a;
    }
    for (;;) {
      a.push(value());
      space();
      if (i_ < s_.size() && s_[i_] == ',') {
        ++i_;
        continue;
      }
      if (i_ < s_.size() && s_[i_] == ']') {
        ++i_;
        return a;
      }
      fail("expected ',' or ']'");
    }
  }
  static void utf8(std::string& out, unsigned cp) {
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
  unsigned hex4() {
    if (i_ + 4 > s_.size()) fail("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = s_[i_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= static_cast<unsigned>(c - '0');
      else if (c >= 'a' && c <= 'f') v |= static_cast<unsigned>(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= static_cast<unsigned>(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string text() {
    std::string out;
    ++i_;  // opening quote
    while (i_ < s_.size()) {
      const char c = s_[i_++];
      if (c == '"') return out;
      if (c != '\\') {
        out += c;
        continue;
      }
      if (i_ >= s_.size()) break;
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
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && i_ + 1 < s_.size() && s_[i_] == '\\' && s_[i_ + 1] == 'u') {
            i_ += 2;
            const unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    fail("unterminated string");
  }
  Value number() {
    const char* begin = s_.c_str() + i_;
    char* end = nullptr;
    const double x = std::strtod(begin, &end);
    if (end == begin) fail("unexpected character");
    i_ += static_cast<std::size_t>(end - begin);
    return Value::number(x);
  }

  const std::string& s_;
  std::size_t i_ = 0;
};

inline void writeString(std::string& out, const std::string& s) {
  out += '"';
  for (const char ch : s) {
    const unsigned char c = static_cast<unsigned char>(ch);
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
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

}  // namespace detail

inline Value Value::parse(const std::string& text) { return detail::JsonReader(text).document(); }

inline void Value::write(std::string& out) const {
  switch (kind_) {
    case Kind::Undefined:
    case Kind::Null: out += "null"; break;
    case Kind::Bool: out += num_ != 0.0 ? "true" : "false"; break;
    case Kind::Number: {
      if (!std::isfinite(num_)) {
        out += "null";
        break;
      }
      char buf[32];
      std::snprintf(buf, sizeof buf, "%.17g", num_);
      out += buf;
      break;
    }
    case Kind::String: detail::writeString(out, str_); break;
    case Kind::Array: {
      out += '[';
      for (std::size_t i = 0; i < items_.size(); ++i) {
        if (i) out += ',';
        items_[i].write(out);
      }
      out += ']';
      break;
    }
    case Kind::Object: {
      out += '{';
      bool first = true;
      for (const auto& kv : members_) {
        if (kv.second.kind_ == Kind::Undefined) continue;
        if (!first) out += ',';
        first = false;
        detail::writeString(out, kv.first);
        out += ':';
        kv.second.write(out);
      }
      out += '}';
      break;
    }
  }
}

}  // namespace app
}  // namespace ccrhip
