#include "json.hpp"

#include <cctype>
#include <cstdlib>

namespace fwiflow {
namespace json {
namespace {

struct Reader {
  const std::string& s;
  size_t p = 0;
  std::string err;
  explicit Reader(const std::string& t) : s(t) {}

  void ws() {
    while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
  }
  bool fail(const char* m) {
    if (err.empty()) err = std::string(m) + " at offset " + std::to_string(p);
    return false;
  }
  bool lit(const char* w) {
    size_t n = 0;
    while (w[n]) n++;
    if (s.compare(p, n, w) != 0) return fail("bad literal");
    p += n;
    return true;
  }
  bool string(std::string& out) {
    if (p >= s.size() || s[p] != '"') return fail("expected string");
    p++;
    out.clear();
    while (p < s.size() && s[p] != '"') {
      char c = s[p++];
      if (c == '\\') {
        if (p >= s.size()) return fail("bad escape");
        char e = s[p++];
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
            if (p + 4 > s.size()) return fail("bad \\u escape");
            unsigned cp = (unsigned)std::strtoul(s.substr(p, 4).c_str(), nullptr, 16);
            p += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: return fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    if (p >= s.size()) return fail("unterminated string");
    p++;
    return true;
  }
  bool number(Value& v) {
    size_t q = p;
    if (q < s.size() && (s[q] == '-' || s[q] == '+')) q++;
    bool frac = false;
    while (q < s.size()) {
      char c = s[q];
      if (std::isdigit((unsigned char)c)) {
        q++;
      } else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
        frac = true;
        q++;
      } else {
        break;
      }
    }
    if (q == p) return fail("expected number");
    std::string tok = s.substr(p, q - p);
    char* end = nullptr;
    v.num = std::strtod(tok.c_str(), &end);
    if (!end || *end != '\0') return fail("bad number");
    v.kind = Value::Number;
    v.is_int = !frac;
    p = q;
    return true;
  }
  bool value(Value& v, int depth) {
    if (depth > 64) return fail("nesting too deep");
    ws();
    if (p >= s.size()) return fail("unexpected end");
    char c = s[p];
    if (c == '{') {
      p++;
      v.kind = Value::Object;
      ws();
      if (p < s.size() && s[p] == '}') {
        p++;
        return true;
      }
      for (;;) {
        ws();
        std::string key;
        if (!string(key)) return false;
        ws();
        if (p >= s.size() || s[p] != ':') return fail("expected ':'");
        p++;
        Value item;
        if (!value(item, depth + 1)) return false;
        v.obj.emplace_back(std::move(key), std::move(item));
        ws();
        if (p < s.size() && s[p] == ',') {
          p++;
          continue;
        }
        if (p < s.size() && s[p] == '}') {
          p++;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      p++;
      v.kind = Value::Array;
      ws();
      if (p < s.size() && s[p] == ']') {
        p++;
        return true;
      }
      for (;;) {
        Value item;
        if (!value(item, depth + 1)) return false;
        v.arr.push_back(std::move(item));
        ws();
        if (p < s.size() && s[p] == ',') {
          p++;
          continue;
        }
        if (p < s.size() && s[p] == ']') {
          p++;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Value::String;
      return string(v.str);
    }
    if (c == 't') {
      v.kind = Value::Bool;
      v.b = true;
      return lit("true");
    }
    if (c == 'f') {
      v.kind = Value::Bool;
      v.b = false;
      return lit("false");
    }
    if (c == 'n') {
      v.kind = Value::Null;
      return lit("null");
    }
    return number(v);
  }
};

}  // namespace

bool parse(const std::string& text, Value& out, std::string& err) {
  Reader r(text);
  out = Value();
  if (!r.value(out, 0)) {
    err = r.err;
    return false;
  }
  r.ws();
  if (r.p != text.size()) {
    err = "trailing characters at offset " + std::to_string(r.p);
    return false;
  }
  return true;
}

}  // namespace json
}  // namespace fwiflow
