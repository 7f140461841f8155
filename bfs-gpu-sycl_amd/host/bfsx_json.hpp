// Minimal JSON value for the CLI's output file: objects (sorted keys),
// arrays, strings, booleans, integers and doubles; parse and pretty-print.
// The reference writes its results with nlohmann::json (src/host.cpp:228-271,
// `combined.dump(4)`); this type supports the subset that output needs with
// the same shape: sorted object keys, 4-space indentation, integers printed
// as integers and doubles in shortest round-trip form.
#pragma once

#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <istream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace bfsx {

class Json {
 public:
  enum class Type { Null, Bool, Int, UInt, Double, String, Array, Object };

  Json() = default;
  Json(std::nullptr_t) {}
  Json(bool b) : t_(Type::Bool), b_(b) {}
  Json(double d) : t_(Type::Double), d_(d) {}
  Json(float d) : t_(Type::Double), d_(d) {}
  Json(const char *s) : t_(Type::String), s_(s) {}
  Json(const std::string &s) : t_(Type::String), s_(s) {}
  template <class I, std::enable_if_t<std::is_integral_v<I> && !std::is_same_v<I, bool>, int> = 0>
  Json(I v) {
    if constexpr (std::is_signed_v<I>) {
      t_ = Type::Int;
      i_ = (int64_t)v;
    } else {
      t_ = Type::UInt;
      u_ = (uint64_t)v;
    }
  }
  template <class T>
  Json(const std::vector<T> &v) : t_(Type::Array) {
    for (const T &x : v) a_.emplace_back(x);
  }

  static Json object() {
    Json j;
    j.t_ = Type::Object;
    return j;
  }
  static Json array() {
    Json j;
    j.t_ = Type::Array;
    return j;
  }

  Type type() const { return t_; }
  bool is_object() const { return t_ == Type::Object; }
  bool is_array() const { return t_ == Type::Array; }
  bool is_null() const { return t_ == Type::Null; }
  bool contains(const std::string &k) const { return t_ == Type::Object && o_.count(k); }
  size_t size() const { return t_ == Type::Array ? a_.size() : t_ == Type::Object ? o_.size() : 0; }

  Json &operator[](const std::string &k) {
    if (t_ == Type::Null) t_ = Type::Object;
    if (t_ != Type::Object) throw std::runtime_error("json: not an object");
    return o_[k];
  }
  Json &operator[](const char *k) { return (*this)[std::string(k)]; }
  const Json &at(const std::string &k) const {
    auto it = o_.find(k);
    if (t_ != Type::Object || it == o_.end()) throw std::runtime_error("json: no key " + k);
    return it->second;
  }
  Json &operator[](size_t i) {
    if (t_ != Type::Array || i >= a_.size()) throw std::runtime_error("json: bad index");
    return a_[i];
  }
  void push_back(const Json &v) {
    if (t_ == Type::Null) t_ = Type::Array;
    if (t_ != Type::Array) throw std::runtime_error("json: not an array");
    a_.push_back(v);
  }
  const std::map<std::string, Json> &items() const { return o_; }
  const std::vector<Json> &elements() const { return a_; }

  double as_double() const {
    switch (t_) {
      case Type::Double: return d_;
      case Type::Int: return (double)i_;
      case Type::UInt: return (double)u_;
      case Type::Bool: return b_ ? 1.0 : 0.0;
      default: throw std::runtime_error("json: not a number");
    }
  }
  uint64_t as_uint() const {
    switch (t_) {
      case Type::UInt: return u_;
      case Type::Int: return (uint64_t)i_;
      case Type::Double: return (uint64_t)d_;
      case Type::Bool: return b_;
      default: throw std::runtime_error("json: not a number");
    }
  }
  int64_t as_int() const { return t_ == Type::Int ? i_ : (int64_t)as_uint(); }
  bool as_bool() const {
    if (t_ != Type::Bool) throw std::runtime_error("json: not a boolean");
    return b_;
  }
  const std::string &as_string() const {
    if (t_ != Type::String) throw std::runtime_error("json: not a string");
    return s_;
  }
  // conversions used by reference-shaped code: static_cast<double>(obj["k"])
  explicit operator double() const { return as_double(); }
  explicit operator unsigned int() const { return (unsigned int)as_uint(); }
  explicit operator uint64_t() const { return as_uint(); }
  explicit operator int() const { return (int)as_int(); }
  explicit operator bool() const { return as_bool(); }
  explicit operator std::string() const { return as_string(); }

  std::string dump(int indent = -1) const {
    std::string out;
    write(out, indent, 0);
    return out;
  }

  static Json parse(const std::string &text) {
    size_t pos = 0;
    Json j = parse_value(text, pos);
    skip_ws(text, pos);
    if (pos != text.size()) throw std::runtime_error("json: trailing characters");
    return j;
  }
  static Json parse(std::istream &in) {
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return parse(text);
  }

 private:
  Type t_ = Type::Null;
  bool b_ = false;
  int64_t i_ = 0;
  uint64_t u_ = 0;
  double d_ = 0;
  std::string s_;
  std::vector<Json> a_;
  std::map<std::string, Json> o_;

  static void write_string(std::string &out, const std::string &s) {
    out += '"';
    for (unsigned char c : s) {
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
            std::snprintf(buf, sizeof(buf), "\\u%04x", c);
            out += buf;
          } else {
            out += (char)c;
          }
      }
    }
    out += '"';
  }

  static void write_double(std::string &out, double d) {
    if (!std::isfinite(d)) {
      out += "null";
      return;
    }
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), d);
    std::string s(buf, r.ptr);
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    out += s;
  }

  void write(std::string &out, int indent, int depth) const {
    const std::string nl = indent >= 0 ? "\n" : "";
    auto pad = [&](int d) { return indent >= 0 ? std::string((size_t)(indent * d), ' ') : ""; };
    switch (t_) {
      case Type::Null: out += "null"; break;
      case Type::Bool: out += b_ ? "true" : "false"; break;
      case Type::Int: out += std::to_string(i_); break;
      case Type::UInt: out += std::to_string(u_); break;
      case Type::Double: write_double(out, d_); break;
      case Type::String: write_string(out, s_); break;
      case Type::Array:
        if (a_.empty()) {
          out += "[]";
          break;
        }
        out += "[" + nl;
        for (size_t k = 0; k < a_.size(); k++) {
          out += pad(depth + 1);
          a_[k].write(out, indent, depth + 1);
          out += (k + 1 < a_.size() ? "," : "") + nl;
        }
        out += pad(depth) + "]";
        break;
      case Type::Object: {
        if (o_.empty()) {
          out += "{}";
          break;
        }
        out += "{" + nl;
        size_t k = 0;
        for (const auto &kv : o_) {
          out += pad(depth + 1);
          write_string(out, kv.first);
          out += indent >= 0 ? ": " : ":";
          kv.second.write(out, indent, depth + 1);
          out += (++k < o_.size() ? "," : "") + nl;
        }
        out += pad(depth) + "}";
        break;
      }
    }
  }

  static void skip_ws(const std::string &t, size_t &p) {
    while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) p++;
  }

  static void expect(const std::string &t, size_t &p, const char *lit) {
    for (const char *c = lit; *c; c++, p++)
      if (p >= t.size() || t[p] != *c) throw std::runtime_error("json: unexpected token");
  }

  static void put_utf8(std::string &out, uint32_t cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }

  static std::string parse_string(const std::string &t, size_t &p) {
    if (p >= t.size() || t[p] != '"') throw std::runtime_error("json: expected string");
    p++;
    std::string s;
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c != '\\') {
        s += c;
        continue;
      }
      if (p >= t.size()) break;
      char e = t[p++];
      switch (e) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          if (p + 4 > t.size()) throw std::runtime_error("json: bad escape");
          uint32_t cp = (uint32_t)std::stoul(t.substr(p, 4), nullptr, 16);
          p += 4;
          if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= t.size() && t[p] == '\\' && t[p + 1] == 'u') {
            const uint32_t lo = (uint32_t)std::stoul(t.substr(p + 2, 4), nullptr, 16);
            p += 6;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(s, cp);
          break;
        }
        default: throw std::runtime_error("json: bad escape");
      }
    }
    if (p >= t.size()) throw std::runtime_error("json: unterminated string");
    p++;
    return s;
  }

  static Json parse_number(const std::string &t, size_t &p) {
    const size_t start = p;
    if (t[p] == '-') p++;
    bool frac = false;
    while (p < t.size() && (std::isdigit((unsigned char)t[p]) || t[p] == '.' || t[p] == 'e' ||
                            t[p] == 'E' || t[p] == '+' || t[p] == '-')) {
      if (t[p] == '.' || t[p] == 'e' || t[p] == 'E') frac = true;
      p++;
    }
    const std::string num = t.substr(start, p - start);
    if (num.empty() || num == "-") throw std::runtime_error("json: bad number");
    if (!frac) {
      if (num[0] == '-') {
        int64_t v = 0;
        auto r = std::from_chars(num.data(), num.data() + num.size(), v);
        if (r.ec == std::errc()) return Json(v);
      } else {
        uint64_t v = 0;
        auto r = std::from_chars(num.data(), num.data() + num.size(), v);
        if (r.ec == std::errc()) return Json(v);
      }
    }
    return Json(std::stod(num));
  }

  static Json parse_value(const std::string &t, size_t &p) {
    skip_ws(t, p);
    if (p >= t.size()) throw std::runtime_error("json: unexpected end");
    const char c = t[p];
    if (c == '{') {
      p++;
      Json o = object();
      skip_ws(t, p);
      if (p < t.size() && t[p] == '}') {
        p++;
        return o;
      }
      for (;;) {
        skip_ws(t, p);
        std::string k = parse_string(t, p);
        skip_ws(t, p);
        expect(t, p, ":");
        o.o_[k] = parse_value(t, p);
        skip_ws(t, p);
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        expect(t, p, "}");
        return o;
      }
    }
    if (c == '[') {
      p++;
      Json a = array();
      skip_ws(t, p);
      if (p < t.size() && t[p] == ']') {
        p++;
        return a;
      }
      for (;;) {
        a.a_.push_back(parse_value(t, p));
        skip_ws(t, p);
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        expect(t, p, "]");
        return a;
      }
    }
    if (c == '"') return Json(parse_string(t, p));
    if (c == 't') {
      expect(t, p, "true");
      return Json(true);
    }
    if (c == 'f') {
      expect(t, p, "false");
      return Json(false);
    }
    if (c == 'n') {
      expect(t, p, "null");
      return Json();
    }
    return parse_number(t, p);
  }
};

inline std::istream &operator>>(std::istream &in, Json &j) {
  j = Json::parse(in);
  return in;
}

}  // namespace bfsx
