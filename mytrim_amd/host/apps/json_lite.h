// Minimal JSON reader for the runmytrim input files.
//
// Accepts standard JSON plus `//` line comments and `/* */` block comments,
// which the MyTRIM input decks use (tests/json/*.json).  Values are held in a
// small tree; lookups of missing keys return a shared null value, so
// `root["a"]["b"].isNumber()` is safe, as with jsoncpp's Json::Value.
#ifndef MYTRIM_AMD_JSON_LITE_H
#define MYTRIM_AMD_JSON_LITE_H

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace json_lite {

class Value {
 public:
  enum Kind { Null, Bool, Number, String, Array, Object };

  Value() = default;
  Kind kind() const { return _kind; }
  bool isNull() const { return _kind == Null; }
  bool isBool() const { return _kind == Bool; }
  bool isNumber() const { return _kind == Number; }
  bool isString() const { return _kind == String; }
  bool isArray() const { return _kind == Array; }
  bool isObject() const { return _kind == Object; }

  bool asBool() const { return _b; }
  double asDouble() const { return _n; }
  // jsoncpp's asInt truncates toward zero
  int asInt() const { return int(_n); }
  const std::string& asString() const { return _s; }

  size_t size() const { return _kind == Array ? _a.size() : _kind == Object ? _o.size() : 0; }
  const Value& operator[](size_t i) const { return (_kind == Array && i < _a.size()) ? _a[i] : null(); }
  const Value& operator[](const std::string& k) const {
    if (_kind != Object) return null();
    auto it = _o.find(k);
    return it == _o.end() ? null() : it->second;
  }
  const Value& operator[](const char* k) const { return (*this)[std::string(k)]; }

  static const Value& null() {
    static const Value v;
    return v;
  }

 private:
  friend class Parser;
  Kind _kind = Null;
  bool _b = false;
  double _n = 0.0;
  std::string _s;
  std::vector<Value> _a;
  std::map<std::string, Value> _o;
};

class Parser {
 public:
  explicit Parser(const std::string& text) : _t(text) {}

  Value parse() {
    Value v = value();
    skip();
    if (_p != _t.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& _t;
  size_t _p = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1;
    for (size_t i = 0; i < _p && i < _t.size(); ++i)
      if (_t[i] == '\n') ++line;
    throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ": " + what);
  }

  void skip() {
    for (;;) {
      while (_p < _t.size() && std::isspace((unsigned char)_t[_p])) ++_p;
      if (_p + 1 < _t.size() && _t[_p] == '/' && _t[_p + 1] == '/') {
        while (_p < _t.size() && _t[_p] != '\n') ++_p;
      } else if (_p + 1 < _t.size() && _t[_p] == '/' && _t[_p + 1] == '*') {
        const size_t e = _t.find("*/", _p + 2);
        if (e == std::string::npos) fail("unterminated comment");
        _p = e + 2;
      } else {
        return;
      }
    }
  }

  char peek() {
    skip();
    if (_p >= _t.size()) fail("unexpected end of input");
    return _t[_p];
  }

  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++_p;
  }

  bool literal(const char* w) {
    const std::string s(w);
    if (_t.compare(_p, s.size(), s) == 0) {
      _p += s.size();
      return true;
    }
    return false;
  }

  Value value() {
    const char c = peek();
    Value v;
    if (c == '{') {
      ++_p;
      v._kind = Value::Object;
      if (peek() == '}') {
        ++_p;
        return v;
      }
      for (;;) {
        if (peek() != '"') fail("expected a key string");
        const std::string k = string();
        expect(':');
        v._o[k] = value();
        const char d = peek();
        ++_p;
        if (d == '}') break;
        if (d != ',') fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      ++_p;
      v._kind = Value::Array;
      if (peek() == ']') {
        ++_p;
        return v;
      }
      for (;;) {
        v._a.push_back(value());
        const char d = peek();
        ++_p;
        if (d == ']') break;
        if (d != ',') fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v._kind = Value::String;
      v._s = string();
    } else if (literal("true")) {
      v._kind = Value::Bool;
      v._b = true;
    } else if (literal("false")) {
      v._kind = Value::Bool;
      v._b = false;
    } else if (literal("null")) {
      v._kind = Value::Null;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const char* s = _t.c_str() + _p;
      char* end = nullptr;
      v._n = std::strtod(s, &end);
      if (end == s) fail("bad number");
      _p += size_t(end - s);
      v._kind = Value::Number;
    } else {
      fail(std::string("unexpected character '") + c + "'");
    }
    return v;
  }

  std::string string() {
    expect('"');
    std::string s;
    while (_p < _t.size() && _t[_p] != '"') {
      char c = _t[_p++];
      if (c == '\\') {
        if (_p >= _t.size()) break;
        const char e = _t[_p++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u':  // keep the escape as text (no unicode names occur in input decks)
            s += "\\u";
            continue;
          default: c = e;
        }
      }
      s += c;
    }
    if (_p >= _t.size()) fail("unterminated string");
    ++_p;
    return s;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json_lite

#endif
