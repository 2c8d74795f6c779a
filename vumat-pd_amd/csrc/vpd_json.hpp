// A small strict JSON reader for the case and material files of the case runner, with the
// typed accessors and error messages of the reference's reader (src/config/SimpleJson.cpp:
// 226-318): optional fields fall back to a default, present fields must have the right type,
// integers must be exact and in range, booleans also accept numbers.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace vpdjson {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool boolean = false;
  double number = 0.0;
  std::string string;
  std::vector<Value> array;
  std::map<std::string, Value> object;
};

// Recursive descent; where input is malformed it stops at the same point, with the same
// message, as the reference's reader (SimpleJson.cpp:17-215).
class Reader {
public:
  explicit Reader(const std::string &text) : s_(text) {}

  Value document() {
    Value v = value();
    space();
    if (p_ != s_.size())
      throw std::runtime_error("Unexpected trailing content in JSON input.");
    return v;
  }

private:
  const std::string &s_;
  size_t p_ = 0;

  void space() {
    while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_])))
      ++p_;
  }
  bool at(char c) const { return p_ < s_.size() && s_[p_] == c; }
  void expect(char c) {
    space();
    if (!at(c))
      throw std::runtime_error(std::string("Expected JSON token: ") + c);
    ++p_;
  }
  bool literal(const char *word) {
    const size_t len = std::char_traits<char>::length(word);
    if (s_.compare(p_, len, word) != 0)
      return false;
    p_ += len;
    return true;
  }

  Value value() {
    space();
    if (p_ >= s_.size())
      throw std::runtime_error("Unexpected end of JSON input.");
    const char c = s_[p_];
    Value v;
    if (c == '"') {
      v.kind = Value::String;
      v.string = text();
    } else if (c == '[') {
      v.kind = Value::Array;
      expect('[');
      space();
      if (at(']')) {
        ++p_;
        return v;
      }
      for (;;) {
        v.array.push_back(value());
        space();
        if (at(']')) {
          ++p_;
          return v;
        }
        expect(',');
      }
    } else if (c == '{') {
      v.kind = Value::Object;
      expect('{');
      space();
      if (at('}')) {
        ++p_;
        return v;
      }
      for (;;) {
        space();
        const std::string key = text();
        expect(':');
        v.object[key] = value();
        space();
        if (at('}')) {
          ++p_;
          return v;
        }
        expect(',');
      }
    } else if (std::isdigit(static_cast<unsigned char>(c)) || c == '-') {
      v.kind = Value::Number;
      v.number = number();
    } else if (literal("true")) {
      v.kind = Value::Bool;
      v.boolean = true;
    } else if (literal("false")) {
      v.kind = Value::Bool;
    } else if (!literal("null")) {
      throw std::runtime_error("Unsupported JSON token.");
    }
    return v;
  }

  std::string text() {
    expect('"');
    std::string out;
    while (p_ < s_.size()) {
      const char c = s_[p_++];
      if (c == '"')
        return out;
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= s_.size())
        throw std::runtime_error("Unterminated JSON escape sequence.");
      const char e = s_[p_++];
      const char *from = "\"\\/bfnrt";
      const char *to = "\"\\/\b\f\n\r\t";
      const char *hit = std::strchr(from, e);
      if (e == '\0' || !hit)
        throw std::runtime_error("Unsupported JSON string escape.");
      out += to[hit - from];
    }
    throw std::runtime_error("Unterminated JSON string.");
  }

  // JSON number grammar (a leading 0 takes no further integer digits); the conversion is
  // strtod's correctly rounded one and must consume the whole token.
  double number() {
    const size_t start = p_;
    if (at('-'))
      ++p_;
    auto digits = [&] {
      const size_t d = p_;
      while (p_ < s_.size() && std::isdigit(static_cast<unsigned char>(s_[p_])))
        ++p_;
      if (p_ == d)
        throw std::runtime_error("Invalid JSON number: expected digits.");
    };
    if (at('0'))
      ++p_;
    else
      digits();
    if (at('.')) {
      ++p_;
      digits();
    }
    if (at('e') || at('E')) {
      ++p_;
      if (at('+') || at('-'))
        ++p_;
      digits();
    }
    const std::string token = s_.substr(start, p_ - start);
    char *stop = nullptr;
    const double x = std::strtod(token.c_str(), &stop);
    if (stop != token.c_str() + token.size() || !std::isfinite(x))
      throw std::runtime_error("JSON number must be finite and representable as a double.");
    return x;
  }
};

inline Value parse(const std::string &text) { return Reader(text).document(); }

inline Value load_file(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in)
    throw std::runtime_error("Failed to open JSON file: " + path);
  std::ostringstream buf;
  buf << in.rdbuf();
  return parse(buf.str());
}

inline const Value *find(const Value &object, const std::string &name) {
  if (object.kind != Value::Object)
    throw std::runtime_error("JSON root must be an object.");
  const auto it = object.object.find(name);
  return it == object.object.end() ? nullptr : &it->second;
}

inline std::string get_string(const Value &o, const std::string &name, bool required = true,
                              const std::string &fallback = "") {
  const Value *v = find(o, name);
  if (!v) {
    if (required)
      throw std::runtime_error("Missing string field: " + name);
    return fallback;
  }
  if (v->kind != Value::String)
    throw std::runtime_error("Expected string field: " + name);
  return v->string;
}

inline int get_int(const Value &o, const std::string &name, bool required = true,
                   int fallback = 0) {
  const Value *v = find(o, name);
  if (!v) {
    if (required)
      throw std::runtime_error("Missing integer field: " + name);
    return fallback;
  }
  if (v->kind != Value::Number)
    throw std::runtime_error("Expected numeric field: " + name);
  if (!std::isfinite(v->number) || std::trunc(v->number) != v->number ||
      v->number < std::numeric_limits<int>::min() || v->number > std::numeric_limits<int>::max())
    throw std::runtime_error("Expected representable integer field: " + name);
  return static_cast<int>(v->number);
}

inline double get_double(const Value &o, const std::string &name, bool required = true,
                         double fallback = 0.0) {
  const Value *v = find(o, name);
  if (!v) {
    if (required)
      throw std::runtime_error("Missing numeric field: " + name);
    return fallback;
  }
  if (v->kind != Value::Number)
    throw std::runtime_error("Expected numeric field: " + name);
  return v->number;
}

inline bool get_bool(const Value &o, const std::string &name, bool required = true,
                     bool fallback = false) {
  const Value *v = find(o, name);
  if (!v) {
    if (required)
      throw std::runtime_error("Missing boolean field: " + name);
    return fallback;
  }
  if (v->kind == Value::Bool)
    return v->boolean;
  if (v->kind == Value::Number)
    return v->number != 0.0;
  throw std::runtime_error("Expected boolean field: " + name);
}

inline std::vector<double> get_double_array(const Value &o, const std::string &name) {
  const Value *v = find(o, name);
  if (!v || v->kind != Value::Array)
    throw std::runtime_error("Expected numeric array field: " + name);
  std::vector<double> out;
  out.reserve(v->array.size());
  for (const Value &e : v->array) {
    if (e.kind != Value::Number)
      throw std::runtime_error("Expected numeric entry in array field: " + name);
    out.push_back(e.number);
  }
  return out;
}

} // namespace vpdjson
