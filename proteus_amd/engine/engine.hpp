// Host engine of the plan API.  Talks to the device only through the C ABI
// of include/proteus_amd.h.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/proteus_amd.h"

extern "C" int pj_engine_fail(int code, const char *fmt, ...);
extern "C" void pj_set_last_error(const char *msg);
#define pj_fail pj_engine_fail

namespace pj {
namespace engine {

// Column types of the plan language (core/olap/lib/values/expressionTypes.cpp):
// "int" = int32, "int64", "date" = int64, "float" = FP64, "bool",
// "dstring" = dictionary-encoded int32; "float32" is a storage type for
// single-precision columns (read as FP64).
enum class Type { I32, I64, F64, BOOL, F32 };
bool parse_type(const char *s, Type *t);
int type_size(Type t);

struct Column {
  Type type = Type::I32;
  uint64_t n = 0;
  void *dev = nullptr;         // device pointer (caller's or staged)
  const void *host = nullptr;  // caller's host pointer (staged on first use)
  bool staged = false;         // dev was allocated by the engine
  uint64_t ticket = 0;         // staging copy not yet waited for (pj_stage_wait)
};

struct Relation {
  std::map<std::string, Column> cols;
};

// A column file of a prepared relation (<rel>.<attr>, headerless), staged to
// HBM on first use and kept there while the file is unchanged.
struct FileColumn {
  Column col;
  uint64_t size = 0;
  int64_t mtime_ns = 0;
  std::shared_ptr<std::vector<std::string>> dict;  // dstring: code -> string
};

struct Catalog {
  std::map<std::string, Relation> rels;
  std::map<std::string, FileColumn> files;  // by path
};

Catalog &catalog_of(pj_ctx *ctx);
void release_staged(pj_ctx *ctx, Catalog &cat);
void release_column(pj_ctx *ctx, Column &c);

}  // namespace engine
}  // namespace pj
