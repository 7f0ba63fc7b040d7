// Minimal JSON reader for the one-line parameter and survey files that
// FwiFlow's writers produce (src/Utils.jl:21-109).  Objects, arrays, strings,
// numbers (with an is-integer flag, so "nz": 134 and "dt": 0.0025 can be told
// apart as rapidjson's IsInt()/IsDouble() do in Parameter.cpp), true/false/null.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fwiflow {
namespace json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  bool is_int = false;  // number literal had no fraction/exponent
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  const Value* get(const std::string& key) const {
    if (kind != Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

// Parses `text`; on failure returns false and sets `err`.
bool parse(const std::string& text, Value& out, std::string& err);

}  // namespace json
}  // namespace fwiflow
