// toml.hpp -- TOML settings files read into the JSON document model.
//
// The reference converts TOML to JSON before wiring a run
// (include/openpfc/frontend/ui/settings_loader.hpp, tests/unit/frontend/ui/
// test_toml_to_json.cpp): tables become objects, arrays of tables arrays of
// objects, integers stay integers.  This reader covers the TOML the
// settings files use: comments, [table] and [dotted.table] headers, [[array
// of tables]], bare / quoted / dotted keys, basic and literal strings (with
// the basic escapes), integers (with _ separators, 0x/0o/0b), floats (with
// exponents, inf, nan), booleans, arrays (nested, multi-line, trailing
// comma) and inline tables.  Dates and multi-line strings are rejected with
// a parse error.  Errors throw std::runtime_error with the line number.
#pragma once

#include "json.hpp"

#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace pfc::toml {

namespace detail {

class Reader {
public:
  explicit Reader(const std::string &s) : m_s(s) {}

  json::Value parse() {
    json::Value root = json::Value::object();
    json::Value *table = &root;
    while (true) {
      skip_blank_lines();
      if (m_i >= m_s.size()) break;
      if (m_s[m_i] == '[') {
        const bool array = m_i + 1 < m_s.size() && m_s[m_i + 1] == '[';
        m_i += array ? 2 : 1;
        ws();
        std::vector<std::string> path = key_path();
        ws();
        if (!consume(']') || (array && !consume(']'))) fail("bad table header");
        end_of_line();
        table = array ? open_array_table(root, path) : open_table(root, path);
        continue;
      }
      std::vector<std::string> path = key_path();
      ws();
      if (!consume('=')) fail("expected '=' after key");
      ws();
      json::Value v = value();
      assign(*table, path, std::move(v));
      end_of_line();
    }
    return root;
  }

private:
  [[noreturn]] void fail(const std::string &what) const {
    size_t line = 1;
    for (size_t k = 0; k < m_i && k < m_s.size(); ++k)
      if (m_s[k] == '\n') ++line;
    throw std::runtime_error("TOML parse error at line " + std::to_string(line) + ": " + what);
  }
  void ws() {
    while (m_i < m_s.size() && (m_s[m_i] == ' ' || m_s[m_i] == '\t')) ++m_i;
  }
  bool consume(char c) {
    if (m_i < m_s.size() && m_s[m_i] == c) {
      ++m_i;
      return true;
    }
    return false;
  }
  void skip_comment() {
    if (m_i < m_s.size() && m_s[m_i] == '#')
      while (m_i < m_s.size() && m_s[m_i] != '\n') ++m_i;
  }
  void skip_blank_lines() {
    while (m_i < m_s.size()) {
      ws();
      skip_comment();
      if (m_i < m_s.size() && (m_s[m_i] == '\n' || m_s[m_i] == '\r')) {
        ++m_i;
        continue;
      }
      break;
    }
  }
  // whitespace, comments and newlines inside arrays
  void skip_array_space() {
    while (m_i < m_s.size()) {
      const char c = m_s[m_i];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
        ++m_i;
      } else if (c == '#') {
        skip_comment();
      } else {
        break;
      }
    }
  }
  void end_of_line() {
    ws();
    skip_comment();
    if (m_i >= m_s.size()) return;
    if (m_s[m_i] == '\r') ++m_i;
    if (m_i < m_s.size() && m_s[m_i] == '\n') {
      ++m_i;
      return;
    }
    fail("unexpected characters after value");
  }

  static bool bare_char(char c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') ||
           c == '_' || c == '-';
  }
  std::string key_part() {
    if (m_i < m_s.size() && (m_s[m_i] == '"' || m_s[m_i] == '\'')) return string_value();
    const size_t b = m_i;
    while (m_i < m_s.size() && bare_char(m_s[m_i])) ++m_i;
    if (m_i == b) fail("expected a key");
    return m_s.substr(b, m_i - b);
  }
  std::vector<std::string> key_path() {
    std::vector<std::string> p{key_part()};
    while (true) {
      ws();
      if (!consume('.')) break;
      ws();
      p.push_back(key_part());
    }
    return p;
  }

  std::string string_value() {
    const char q = m_s[m_i++];
    if (m_i + 1 < m_s.size() && m_s[m_i] == q && m_s[m_i + 1] == q)
      fail("multi-line strings are not supported");
    std::string o;
    while (true) {
      if (m_i >= m_s.size() || m_s[m_i] == '\n') fail("unterminated string");
      const char c = m_s[m_i++];
      if (c == q) break;
      if (q == '\'' || c != '\\') {
        o += c;
        continue;
      }
      const char e = m_i < m_s.size() ? m_s[m_i++] : '\0';
      switch (e) {
      case '"': o += '"'; break;
      case '\\': o += '\\'; break;
      case 'b': o += '\b'; break;
      case 'f': o += '\f'; break;
      case 'n': o += '\n'; break;
      case 'r': o += '\r'; break;
      case 't': o += '\t'; break;
      default: fail(std::string("unsupported escape '\\") + e + "'");
      }
    }
    return o;
  }

  json::Value number_value() {
    const size_t b = m_i;
    while (m_i < m_s.size() && (bare_char(m_s[m_i]) || m_s[m_i] == '+' || m_s[m_i] == '.'))
      ++m_i;
    std::string tok = m_s.substr(b, m_i - b), t;
    for (char c : tok)
      if (c != '_') t += c;
    if (t == "inf" || t == "+inf") return json::Value::number(std::numeric_limits<double>::infinity());
    if (t == "-inf") return json::Value::number(-std::numeric_limits<double>::infinity());
    if (t == "nan" || t == "+nan" || t == "-nan")
      return json::Value::number(std::numeric_limits<double>::quiet_NaN());
    int base = 10;
    std::string digits = t;
    if (t.size() > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'o' || t[1] == 'b')) {
      base = t[1] == 'x' ? 16 : t[1] == 'o' ? 8 : 2;
      digits = t.substr(2);
    }
    const bool is_float = base == 10 && t.find_first_of(".eE") != std::string::npos;
    char *end = nullptr;
    errno = 0;
    if (is_float) {
      const double d = std::strtod(digits.c_str(), &end);
      if (!end || *end != '\0' || digits.empty()) fail("bad number '" + tok + "'");
      return json::Value::number(d);
    }
    const long long v = std::strtoll(digits.c_str(), &end, base);
    if (!end || *end != '\0' || digits.empty() || errno) fail("bad number '" + tok + "'");
    return json::Value::integer(v);
  }

  json::Value value() {
    if (m_i >= m_s.size()) fail("expected a value");
    const char c = m_s[m_i];
    if (c == '"' || c == '\'') return json::Value::string(string_value());
    if (c == '[') {
      ++m_i;
      json::Value a = json::Value::array();
      while (true) {
        skip_array_space();
        if (consume(']')) break;
        a.push_back(value());
        skip_array_space();
        if (consume(',')) continue;
        skip_array_space();
        if (consume(']')) break;
        fail("expected ',' or ']' in array");
      }
      return a;
    }
    if (c == '{') {
      ++m_i;
      json::Value t = json::Value::object();
      ws();
      if (consume('}')) return t;
      while (true) {
        ws();
        std::vector<std::string> path = key_path();
        ws();
        if (!consume('=')) fail("expected '=' in inline table");
        ws();
        assign(t, path, value());
        ws();
        if (consume(',')) continue;
        if (consume('}')) break;
        fail("expected ',' or '}' in inline table");
      }
      return t;
    }
    if (m_s.compare(m_i, 4, "true") == 0 && !(m_i + 4 < m_s.size() && bare_char(m_s[m_i + 4]))) {
      m_i += 4;
      return json::Value::boolean(true);
    }
    if (m_s.compare(m_i, 5, "false") == 0 && !(m_i + 5 < m_s.size() && bare_char(m_s[m_i + 5]))) {
      m_i += 5;
      return json::Value::boolean(false);
    }
    if (c == '+' || c == '-' || (c >= '0' && c <= '9') || c == 'i' || c == 'n') {
      // reject dates (1979-05-27...) before reading a number
      size_t k = m_i;
      while (k < m_s.size() && m_s[k] >= '0' && m_s[k] <= '9') ++k;
      if (k - m_i == 4 && k < m_s.size() && m_s[k] == '-') fail("dates are not supported");
      return number_value();
    }
    fail(std::string("unexpected character '") + c + "'");
  }

  void assign(json::Value &table, const std::vector<std::string> &path, json::Value v) {
    json::Value *t = &table;
    for (size_t k = 0; k + 1 < path.size(); ++k) {
      json::Value &next = (*t)[path[k]];
      if (next.is_null()) next = json::Value::object();
      if (!next.is_object()) fail("key '" + path[k] + "' is not a table");
      t = &next;
    }
    if (t->contains(path.back())) fail("duplicate key '" + path.back() + "'");
    (*t)[path.back()] = std::move(v);
  }
  json::Value *descend(json::Value &root, const std::vector<std::string> &path, size_t n) {
    json::Value *t = &root;
    for (size_t k = 0; k < n; ++k) {
      json::Value &next = (*t)[path[k]];
      if (next.is_null()) next = json::Value::object();
      if (next.is_array()) {
        if (next.size() == 0) fail("empty array '" + path[k] + "'");
        t = &next.items().back(); // the last table of an array of tables
        continue;
      }
      if (!next.is_object()) fail("key '" + path[k] + "' is not a table");
      t = &next;
    }
    return t;
  }
  json::Value *open_table(json::Value &root, const std::vector<std::string> &path) {
    json::Value *t = descend(root, path, path.size());
    if (!t->is_object()) fail("table header names a non-table");
    return t;
  }
  json::Value *open_array_table(json::Value &root, const std::vector<std::string> &path) {
    json::Value *parent = descend(root, path, path.size() - 1);
    json::Value &arr = (*parent)[path.back()];
    if (arr.is_null()) arr = json::Value::array();
    if (!arr.is_array()) fail("'" + path.back() + "' is not an array of tables");
    arr.push_back(json::Value::object());
    return &arr.items().back();
  }

  const std::string &m_s;
  size_t m_i = 0;
};

} // namespace detail

inline json::Value parse(const std::string &text) { return detail::Reader(text).parse(); }

} // namespace pfc::toml
