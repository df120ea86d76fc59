// weldform/Json.h -- a small JSON reader for the scene files of the
// reference's command-line driver (Source/WeldForm.cpp:147-944; the
// reference reads them with nlohmann::json).  Values are null, booleans,
// numbers (double), strings, arrays and objects; object members keep their
// order and the first of repeated keys wins.  Looking up a missing member
// or index gives null, as the reference's non-const lookups do.
#pragma once
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "matvec.h"

namespace SPH {
namespace json {

struct Value {
  enum Type { Null, Bool, Number, String, Array, Object };
  Type type = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  bool is_null() const { return type == Null; }
  size_t size() const { return type == Array ? arr.size() : (type == Object ? obj.size() : 0); }
  const Value &operator[](const std::string &k) const {
    if (type == Object)
      for (const auto &m : obj)
        if (m.first == k) return m.second;
    return null_value();
  }
  const Value &operator[](const char *k) const { return (*this)[std::string(k)]; }
  const Value &operator[](size_t i) const {
    if (type == Array && i < arr.size()) return arr[i];
    return null_value();
  }
  const Value &operator[](int i) const { return i < 0 ? null_value() : (*this)[(size_t)i]; }
  static const Value &null_value() {
    static const Value v;
    return v;
  }
};

class Parser {
 public:
  explicit Parser(const std::string &s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char *what) {
    std::ostringstream o;
    o << "JSON: " << what << " at offset " << p_;
    throw new Fatal(o.str());
  }
  void ws() {
    while (p_ < s_.size()) {
      const char c = s_[p_];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r') ++p_;
      else break;
    }
  }
  bool lit(const char *w) {
    size_t k = 0;
    while (w[k] && p_ + k < s_.size() && s_[p_ + k] == w[k]) ++k;
    if (w[k]) return false;
    p_ += k;
    return true;
  }
  Value value() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    const char c = s_[p_];
    Value v;
    if (c == '{') {
      v.type = Value::Object;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("expected a member name");
        std::string k = string();
        ws();
        if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
        ++p_;
        v.obj.emplace_back(std::move(k), value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == '}') {
          ++p_;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.type = Value::Array;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ']') {
          ++p_;
          return v;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = Value::String;
      v.str = string();
      return v;
    }
    if (lit("true")) {
      v.type = Value::Bool;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.type = Value::Bool;
      return v;
    }
    if (lit("null")) return v;
    // number
    const char *b = s_.c_str() + p_;
    char *e = nullptr;
    v.num = std::strtod(b, &e);
    if (e == b) fail("unexpected character");
    p_ += (size_t)(e - b);
    v.type = Value::Number;
    return v;
  }
  std::string string() {
    std::string out;
    ++p_;  // opening quote
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\') {
        if (p_ >= s_.size()) fail("bad escape");
        const char e = s_[p_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u':
            if (p_ + 4 > s_.size()) fail("bad \\u escape");
            p_ += 4;
            c = '?';
            break;
          default: c = e;
        }
      }
      out.push_back(c);
    }
    if (p_ >= s_.size()) fail("unterminated string");
    ++p_;
    return out;
  }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

inline Value parse_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw new Fatal("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace json

// The reference's input helpers (Source/Input.h:10-60): a present value is
// converted and stored, an absent (null) one leaves the target untouched.
// Numbers convert to integers by truncation; a boolean target needs a JSON
// boolean.
inline bool readValue(const json::Value &j, double &v) {
  if (j.is_null()) return false;
  if (j.type != json::Value::Number) throw new Fatal("JSON: a number was expected");
  v = j.num;
  return true;
}
inline bool readValue(const json::Value &j, int &v) {
  double d;
  if (!readValue(j, d)) return false;
  v = (int)d;
  return true;
}
inline bool readValue(const json::Value &j, size_t &v) {
  double d;
  if (!readValue(j, d)) return false;
  v = (size_t)d;
  return true;
}
inline bool readValue(const json::Value &j, bool &v) {
  if (j.is_null()) return false;
  if (j.type != json::Value::Bool) throw new Fatal("JSON: a boolean was expected");
  v = j.b;
  return true;
}
inline bool readValue(const json::Value &j, std::string &v) {
  if (j.is_null()) return false;
  if (j.type != json::Value::String) throw new Fatal("JSON: a string was expected");
  v = j.str;
  return true;
}
inline bool readArray(const json::Value &j, std::vector<double> &vec) {
  if (j.is_null()) return false;
  if (j.type != json::Value::Array) throw new Fatal("JSON: an array was expected");
  vec.resize(j.arr.size());
  for (size_t i = 0; i < j.arr.size(); ++i) {
    if (j.arr[i].type != json::Value::Number) throw new Fatal("JSON: an array of numbers was expected");
    vec[i] = j.arr[i].num;
  }
  return true;
}
inline bool readValue(const json::Value &j, std::vector<double> &v) { return readArray(j, v); }
inline bool readVector(const json::Value &j, Vec3_t &vec) {
  std::vector<double> v;
  if (!readArray(j, v)) return false;
  for (size_t i = 0; i < v.size() && i < 3; ++i) vec(i) = v[i];
  return true;
}
inline bool readBoolVector(const json::Value &j, bool vec[]) {
  if (j.is_null()) return false;
  if (j.type != json::Value::Array) throw new Fatal("JSON: an array was expected");
  for (size_t i = 0; i < j.arr.size(); ++i) {
    if (j.arr[i].type != json::Value::Bool) throw new Fatal("JSON: an array of booleans was expected");
    vec[i] = j.arr[i].b;
  }
  return true;
}

}  // namespace SPH
