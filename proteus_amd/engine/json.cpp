// JSON parsing and serialisation for plans.
#include "json.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace pj {
namespace engine {

Json Json::boolean(bool b) {
  Json j;
  j.kind_ = Bool;
  j.b_ = b;
  return j;
}
Json Json::integer(int64_t v) {
  Json j;
  j.kind_ = Number;
  j.is_int_ = true;
  j.i_ = v;
  j.d_ = (double)v;
  return j;
}
Json Json::real(double v) {
  Json j;
  j.kind_ = Number;
  j.d_ = v;
  return j;
}
Json Json::string(std::string s) {
  Json j;
  j.kind_ = String;
  j.s_ = std::move(s);
  return j;
}
Json Json::array() {
  Json j;
  j.kind_ = Array;
  return j;
}
Json Json::object() {
  Json j;
  j.kind_ = Object;
  return j;
}

Json &Json::push(Json v) {
  arr_.push_back(std::move(v));
  return arr_.back();
}

bool Json::has(const char *k) const { return get(k) != nullptr; }

const Json *Json::get(const char *k) const {
  if (kind_ != Object) return nullptr;
  for (auto &m : obj_)
    if (m.first == k) return &m.second;
  return nullptr;
}

const Json &Json::operator[](const char *k) const {
  static const Json null_json;
  const Json *p = get(k);
  return p ? *p : null_json;
}

Json &Json::set(const std::string &k, Json v) {
  for (auto &m : obj_)
    if (m.first == k) {
      m.second = std::move(v);
      return m.second;
    }
  obj_.emplace_back(k, std::move(v));
  return obj_.back().second;
}

static void dump_string(const std::string &s, std::string &o) {
  o += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      default:
        if (c < 0x20) {
          char b[8];
          snprintf(b, sizeof b, "\\u%04x", c);
          o += b;
        } else {
          o += (char)c;
        }
    }
  }
  o += '"';
}

static void dump_rec(const Json &j, std::string &o) {
  switch (j.kind()) {
    case Json::Null: o += "null"; break;
    case Json::Bool: o += j.as_bool() ? "true" : "false"; break;
    case Json::Number: {
      char b[40];
      if (j.is_int()) snprintf(b, sizeof b, "%lld", (long long)j.as_int());
      else snprintf(b, sizeof b, "%.17g", j.as_double());
      o += b;
      break;
    }
    case Json::String: dump_string(j.as_string(), o); break;
    case Json::Array: {
      o += '[';
      bool first = true;
      for (auto &v : j.items()) {
        if (!first) o += ',';
        first = false;
        dump_rec(v, o);
      }
      o += ']';
      break;
    }
    case Json::Object: {
      o += '{';
      bool first = true;
      for (auto &m : j.members()) {
        if (!first) o += ',';
        first = false;
        dump_string(m.first, o);
        o += ':';
        dump_rec(m.second, o);
      }
      o += '}';
      break;
    }
  }
}

std::string Json::dump() const {
  std::string o;
  dump_rec(*this, o);
  return o;
}

class JsonParser {
 public:
  JsonParser(const std::string &t) : s_(t.c_str()), n_(t.size()) {}

  bool parse(Json *out, std::string *err) {
    skip();
    if (!value(out, 0)) {
      *err = err_ + " at offset " + std::to_string(p_);
      return false;
    }
    skip();
    if (p_ != n_) {
      *err = "trailing characters at offset " + std::to_string(p_);
      return false;
    }
    return true;
  }

 private:
  const char *s_;
  size_t n_, p_ = 0;
  std::string err_;

  void skip() {
    while (p_ < n_ && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t')) ++p_;
  }
  bool bad(const char *m) {
    err_ = m;
    return false;
  }
  bool lit(const char *w) {
    size_t l = strlen(w);
    if (p_ + l > n_ || strncmp(s_ + p_, w, l) != 0) return false;
    p_ += l;
    return true;
  }
  static void utf8(uint32_t c, std::string &o) {
    if (c < 0x80) {
      o += (char)c;
    } else if (c < 0x800) {
      o += (char)(0xc0 | (c >> 6));
      o += (char)(0x80 | (c & 0x3f));
    } else {
      o += (char)(0xe0 | (c >> 12));
      o += (char)(0x80 | ((c >> 6) & 0x3f));
      o += (char)(0x80 | (c & 0x3f));
    }
  }
  bool str(std::string *o) {
    ++p_;  // opening quote
    while (p_ < n_) {
      char c = s_[p_++];
      if (c == '"') return true;
      if (c != '\\') {
        *o += c;
        continue;
      }
      if (p_ >= n_) break;
      char e = s_[p_++];
      switch (e) {
        case '"': *o += '"'; break;
        case '\\': *o += '\\'; break;
        case '/': *o += '/'; break;
        case 'b': *o += '\b'; break;
        case 'f': *o += '\f'; break;
        case 'n': *o += '\n'; break;
        case 'r': *o += '\r'; break;
        case 't': *o += '\t'; break;
        case 'u': {
          if (p_ + 4 > n_) return bad("short \\u escape");
          char h[5] = {s_[p_], s_[p_ + 1], s_[p_ + 2], s_[p_ + 3], 0};
          p_ += 4;
          utf8((uint32_t)strtoul(h, nullptr, 16), *o);
          break;
        }
        default: return bad("bad escape");
      }
    }
    return bad("unterminated string");
  }
  bool num(Json *out) {
    size_t b = p_;
    bool isint = true;
    if (s_[p_] == '-') ++p_;
    while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    if (p_ < n_ && s_[p_] == '.') {
      isint = false;
      ++p_;
      while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    }
    if (p_ < n_ && (s_[p_] == 'e' || s_[p_] == 'E')) {
      isint = false;
      ++p_;
      if (p_ < n_ && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
      while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    }
    std::string t(s_ + b, p_ - b);
    if (t.empty() || t == "-") return bad("bad number");
    if (isint) {
      errno = 0;
      long long v = strtoll(t.c_str(), nullptr, 10);
      if (errno == 0) {
        *out = Json::integer(v);
        return true;
      }
    }
    const double d = strtod(t.c_str(), nullptr);
    if (!std::isfinite(d)) return bad("number out of range");  // no JSON spelling for inf
    *out = Json::real(d);
    return true;
  }
  bool value(Json *out, int depth) {
    if (depth > 256) return bad("nesting too deep");
    if (p_ >= n_) return bad("unexpected end");
    char c = s_[p_];
    if (c == '{') {
      ++p_;
      *out = Json::object();
      skip();
      if (p_ < n_ && s_[p_] == '}') {
        ++p_;
        return true;
      }
      while (true) {
        skip();
        if (p_ >= n_ || s_[p_] != '"') return bad("expected member name");
        std::string k;
        if (!str(&k)) return false;
        skip();
        if (p_ >= n_ || s_[p_] != ':') return bad("expected ':'");
        ++p_;
        skip();
        Json v;
        if (!value(&v, depth + 1)) return false;
        out->obj_.emplace_back(std::move(k), std::move(v));
        skip();
        if (p_ < n_ && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < n_ && s_[p_] == '}') {
          ++p_;
          return true;
        }
        return bad("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p_;
      *out = Json::array();
      skip();
      if (p_ < n_ && s_[p_] == ']') {
        ++p_;
        return true;
      }
      while (true) {
        skip();
        Json v;
        if (!value(&v, depth + 1)) return false;
        out->arr_.push_back(std::move(v));
        skip();
        if (p_ < n_ && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < n_ && s_[p_] == ']') {
          ++p_;
          return true;
        }
        return bad("expected ',' or ']'");
      }
    }
    if (c == '"') {
      std::string s;
      if (!str(&s)) return false;
      *out = Json::string(std::move(s));
      return true;
    }
    if (lit("true")) {
      *out = Json::boolean(true);
      return true;
    }
    if (lit("false")) {
      *out = Json::boolean(false);
      return true;
    }
    if (lit("null")) {
      *out = Json();
      return true;
    }
    if (c == '-' || (c >= '0' && c <= '9')) return num(out);
    return bad("unexpected character");
  }
};

bool json_parse(const std::string &text, Json *out, std::string *err) {
  JsonParser p(text);
  return p.parse(out, err);
}

}  // namespace engine
}  // namespace pj
