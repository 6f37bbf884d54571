// Minimal JSON document model for physical plans: parse (RFC 8259 subset:
// no surrogate-pair decoding beyond the BMP) and serialise.  Objects keep
// member order.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace pj {
namespace engine {

class Json {
 public:
  enum Kind { Null, Bool, Number, String, Array, Object };

  Json() = default;
  static Json boolean(bool b);
  static Json integer(int64_t v);
  static Json real(double v);
  static Json string(std::string s);
  static Json array();
  static Json object();

  Kind kind() const { return kind_; }
  bool is_null() const { return kind_ == Null; }
  bool is_bool() const { return kind_ == Bool; }
  bool is_number() const { return kind_ == Number; }
  bool is_int() const { return kind_ == Number && is_int_; }
  bool is_string() const { return kind_ == String; }
  bool is_array() const { return kind_ == Array; }
  bool is_object() const { return kind_ == Object; }

  bool as_bool() const { return b_; }
  int64_t as_int() const { return is_int_ ? i_ : (int64_t)d_; }
  double as_double() const { return is_int_ ? (double)i_ : d_; }
  const std::string &as_string() const { return s_; }

  // arrays
  size_t size() const { return kind_ == Array ? arr_.size() : kind_ == Object ? obj_.size() : 0; }
  const Json &operator[](size_t i) const { return arr_[i]; }
  const std::vector<Json> &items() const { return arr_; }
  Json &push(Json v);

  // objects
  bool has(const char *k) const;
  const Json &operator[](const char *k) const;  // null Json when absent
  const Json *get(const char *k) const;
  Json &set(const std::string &k, Json v);
  const std::vector<std::pair<std::string, Json>> &members() const { return obj_; }

  std::string dump() const;

 private:
  Kind kind_ = Null;
  bool b_ = false;
  bool is_int_ = false;
  int64_t i_ = 0;
  double d_ = 0;
  std::string s_;
  std::vector<Json> arr_;
  std::vector<std::pair<std::string, Json>> obj_;
  friend class JsonParser;
};

// Returns false and fills *err on malformed input.
bool json_parse(const std::string &text, Json *out, std::string *err);

}  // namespace engine
}  // namespace pj
