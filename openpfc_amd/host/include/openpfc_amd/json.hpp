// json.hpp -- a small JSON document model and parser for the settings files.
//
// The reference reads its settings with nlohmann::json
// (include/openpfc/frontend/ui/settings_loader.hpp); this image has no JSON
// library for C++, so the App layer uses this one: objects keep their key
// order, numbers remember whether they were written as integers (the
// reference checks is_number_integer() for grid sizes), parse errors throw
// std::runtime_error with the line and column.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstdio>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pfc::json {

class Value {
public:
  enum class Kind { Null, Bool, Integer, Float, String, Array, Object };

  Value() = default;
  static Value boolean(bool b) { Value v; v.m_kind = Kind::Bool; v.m_bool = b; return v; }
  static Value integer(long long i) { Value v; v.m_kind = Kind::Integer; v.m_int = i; return v; }
  static Value number(double d) { Value v; v.m_kind = Kind::Float; v.m_num = d; return v; }
  static Value string(std::string s) { Value v; v.m_kind = Kind::String; v.m_str = std::move(s); return v; }
  static Value array() { Value v; v.m_kind = Kind::Array; return v; }
  static Value object() { Value v; v.m_kind = Kind::Object; return v; }

  Kind kind() const noexcept { return m_kind; }
  bool is_null() const noexcept { return m_kind == Kind::Null; }
  bool is_bool() const noexcept { return m_kind == Kind::Bool; }
  bool is_number() const noexcept { return m_kind == Kind::Integer || m_kind == Kind::Float; }
  bool is_number_integer() const noexcept { return m_kind == Kind::Integer; }
  bool is_string() const noexcept { return m_kind == Kind::String; }
  bool is_array() const noexcept { return m_kind == Kind::Array; }
  bool is_object() const noexcept { return m_kind == Kind::Object; }

  double as_double() const {
    if (m_kind == Kind::Integer) return (double)m_int;
    if (m_kind == Kind::Float) return m_num;
    throw std::invalid_argument("json: value is not a number: " + dump());
  }
  long long as_int() const {
    if (m_kind == Kind::Integer) return m_int;
    throw std::invalid_argument("json: value is not an integer: " + dump());
  }
  bool as_bool() const {
    if (m_kind == Kind::Bool) return m_bool;
    throw std::invalid_argument("json: value is not a boolean: " + dump());
  }
  const std::string &as_string() const {
    if (m_kind == Kind::String) return m_str;
    throw std::invalid_argument("json: value is not a string: " + dump());
  }

  // arrays
  size_t size() const {
    return m_kind == Kind::Array ? m_arr.size() : m_kind == Kind::Object ? m_obj.size() : 0;
  }
  const Value &operator[](size_t i) const { return m_arr.at(i); }
  void push_back(Value v) {
    if (m_kind == Kind::Null) m_kind = Kind::Array;
    if (m_kind != Kind::Array) throw std::invalid_argument("json: push_back on a non-array");
    m_arr.push_back(std::move(v));
  }
  const std::vector<Value> &items() const { return m_arr; }
  std::vector<Value> &items() { return m_arr; }

  // objects (key order kept)
  bool contains(const std::string &k) const { return find(k) != nullptr; }
  const Value *find(const std::string &k) const {
    if (m_kind != Kind::Object) return nullptr;
    for (const auto &kv : m_obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  Value *find(const std::string &k) {
    if (m_kind != Kind::Object) return nullptr;
    for (auto &kv : m_obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  const Value &at(const std::string &k) const {
    const Value *v = find(k);
    if (!v) throw std::out_of_range("json: key '" + k + "' not found");
    return *v;
  }
  // insert-or-get
  Value &operator[](const std::string &k) {
    if (m_kind == Kind::Null) m_kind = Kind::Object;
    if (m_kind != Kind::Object) throw std::invalid_argument("json: key access on a non-object");
    if (Value *v = find(k)) return *v;
    m_obj.emplace_back(k, Value());
    return m_obj.back().second;
  }
  const std::vector<std::pair<std::string, Value>> &members() const { return m_obj; }

  bool operator==(const Value &o) const {
    if (is_number() && o.is_number()) {
      if (m_kind == Kind::Integer && o.m_kind == Kind::Integer) return m_int == o.m_int;
      return as_double() == o.as_double();
    }
    if (m_kind != o.m_kind) return false;
    switch (m_kind) {
    case Kind::Null: return true;
    case Kind::Bool: return m_bool == o.m_bool;
    case Kind::String: return m_str == o.m_str;
    case Kind::Array: return m_arr == o.m_arr;
    case Kind::Object: {
      if (m_obj.size() != o.m_obj.size()) return false;
      for (const auto &kv : m_obj) {
        const Value *w = o.find(kv.first);
        if (!w || !(kv.second == *w)) return false;
      }
      return true;
    }
    default: return false;
    }
  }
  bool operator!=(const Value &o) const { return !(*this == o); }

  std::string dump(int indent = -1) const {
    std::string out;
    dump_to(out, indent, 0);
    return out;
  }

private:
  static void escape(std::string &out, const std::string &s) {
    out += '"';
    for (unsigned char c : s) {
      switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\t': out += "\\t"; break;
      case '\r': out += "\\r"; break;
      default:
        if (c < 0x20) {
          char b[8];
          std::snprintf(b, sizeof b, "\\u%04x", c);
          out += b;
        } else {
          out += (char)c;
        }
      }
    }
    out += '"';
  }
  void dump_to(std::string &out, int indent, int depth) const {
    auto nl = [&](int d) {
      if (indent >= 0) {
        out += '\n';
        out.append((size_t)(indent * d), ' ');
      }
    };
    switch (m_kind) {
    case Kind::Null: out += "null"; break;
    case Kind::Bool: out += m_bool ? "true" : "false"; break;
    case Kind::Integer: out += std::to_string(m_int); break;
    case Kind::Float: {
      if (!std::isfinite(m_num)) { // the tokens Python's json module reads
        out += std::isnan(m_num) ? "NaN" : (m_num > 0 ? "Infinity" : "-Infinity");
        break;
      }
      char b[32];
      std::snprintf(b, sizeof b, "%.17g", m_num);
      std::string s = b;
      if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
      out += s;
      break;
    }
    case Kind::String: escape(out, m_str); break;
    case Kind::Array:
      out += '[';
      for (size_t i = 0; i < m_arr.size(); ++i) {
        if (i) out += ',';
        nl(depth + 1);
        m_arr[i].dump_to(out, indent, depth + 1);
      }
      if (!m_arr.empty()) nl(depth);
      out += ']';
      break;
    case Kind::Object:
      out += '{';
      for (size_t i = 0; i < m_obj.size(); ++i) {
        if (i) out += ',';
        nl(depth + 1);
        escape(out, m_obj[i].first);
        out += indent >= 0 ? ": " : ":";
        m_obj[i].second.dump_to(out, indent, depth + 1);
      }
      if (!m_obj.empty()) nl(depth);
      out += '}';
      break;
    }
  }

  Kind m_kind = Kind::Null;
  bool m_bool = false;
  long long m_int = 0;
  double m_num = 0;
  std::string m_str;
  std::vector<Value> m_arr;
  std::vector<std::pair<std::string, Value>> m_obj;
};

namespace detail {

class Parser {
public:
  explicit Parser(const std::string &s) : m_s(s) {}

  Value parse_document() {
    Value v = parse_value();
    ws();
    if (m_i != m_s.size()) fail("unexpected trailing characters");
    return v;
  }

private:
  [[noreturn]] void fail(const std::string &what) const {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < m_i && k < m_s.size(); ++k) {
      if (m_s[k] == '\n') {
        ++line;
        col = 1;
      } else {
        ++col;
      }
    }
    std::ostringstream o;
    o << "JSON parse error at line " << line << ", column " << col << ": " << what;
    throw std::runtime_error(o.str());
  }
  void ws() {
    while (m_i < m_s.size() && (m_s[m_i] == ' ' || m_s[m_i] == '\t' || m_s[m_i] == '\n' ||
                                m_s[m_i] == '\r'))
      ++m_i;
  }
  bool eat(char c) {
    ws();
    if (m_i < m_s.size() && m_s[m_i] == c) {
      ++m_i;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "'");
  }
  bool literal(const char *w) {
    const size_t n = std::char_traits<char>::length(w);
    if (m_s.compare(m_i, n, w) == 0) {
      m_i += n;
      return true;
    }
    return false;
  }

  Value parse_value() {
    ws();
    if (m_i >= m_s.size()) fail("unexpected end of input");
    const char c = m_s[m_i];
    if (c == '{') return parse_object();
    if (c == '[') return parse_array();
    if (c == '"') return Value::string(parse_string());
    if (literal("true")) return Value::boolean(true);
    if (literal("false")) return Value::boolean(false);
    if (literal("null")) return Value();
    if (c == '-' || (c >= '0' && c <= '9')) return parse_number();
    fail(std::string("unexpected character '") + c + "'");
  }

  Value parse_object() {
    expect('{');
    Value v = Value::object();
    if (eat('}')) return v;
    do {
      ws();
      if (m_i >= m_s.size() || m_s[m_i] != '"') fail("expected a string key");
      std::string k = parse_string();
      expect(':');
      if (v.contains(k)) fail("duplicate key '" + k + "'");
      v[k] = parse_value();
    } while (eat(','));
    expect('}');
    return v;
  }

  Value parse_array() {
    expect('[');
    Value v = Value::array();
    if (eat(']')) return v;
    do {
      v.push_back(parse_value());
    } while (eat(','));
    expect(']');
    return v;
  }

  static void put_utf8(std::string &o, unsigned cp) {
    if (cp < 0x80) {
      o += (char)cp;
    } else if (cp < 0x800) {
      o += (char)(0xC0 | (cp >> 6));
      o += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += (char)(0xE0 | (cp >> 12));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xF0 | (cp >> 18));
      o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (m_i + 4 > m_s.size()) fail("truncated \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      const char h = m_s[m_i++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
      else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string parse_string() {
    ++m_i; // opening quote
    std::string o;
    while (true) {
      if (m_i >= m_s.size()) fail("unterminated string");
      const char c = m_s[m_i++];
      if (c == '"') break;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') {
        o += c;
        continue;
      }
      if (m_i >= m_s.size()) fail("unterminated escape");
      const char e = m_s[m_i++];
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
        if (cp >= 0xD800 && cp < 0xDC00 && m_i + 1 < m_s.size() && m_s[m_i] == '\\' &&
            m_s[m_i + 1] == 'u') {
          m_i += 2;
          const unsigned lo = hex4();
          cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
        }
        put_utf8(o, cp);
        break;
      }
      default: fail(std::string("bad escape '\\") + e + "'");
      }
    }
    return o;
  }
  Value parse_number() {
    const size_t start = m_i;
    bool is_float = false;
    if (m_s[m_i] == '-') ++m_i;
    auto digits = [&] {
      const size_t b = m_i;
      while (m_i < m_s.size() && m_s[m_i] >= '0' && m_s[m_i] <= '9') ++m_i;
      if (m_i == b) fail("expected digits");
    };
    digits();
    if (m_i < m_s.size() && m_s[m_i] == '.') {
      is_float = true;
      ++m_i;
      digits();
    }
    if (m_i < m_s.size() && (m_s[m_i] == 'e' || m_s[m_i] == 'E')) {
      is_float = true;
      ++m_i;
      if (m_i < m_s.size() && (m_s[m_i] == '+' || m_s[m_i] == '-')) ++m_i;
      digits();
    }
    const std::string tok = m_s.substr(start, m_i - start);
    if (!is_float) {
      errno = 0;
      char *end = nullptr;
      const long long v = std::strtoll(tok.c_str(), &end, 10);
      if (errno == 0 && end && *end == '\0') return Value::integer(v);
    }
    return Value::number(std::strtod(tok.c_str(), nullptr));
  }

  const std::string &m_s;
  size_t m_i = 0;
};

} // namespace detail

inline Value parse(const std::string &text) { return detail::Parser(text).parse_document(); }

} // namespace pfc::json
