// Relation catalog of the plan API: named columns registered by the caller
// (the role of Catalog / VectorPlugin in the reference,
// core/olap/lib/util/catalog.hpp and core/olap/lib/plugins/vector-plugin.hpp).
#include <cstdio>
#include <cstring>

#include "engine.hpp"

namespace pj {
namespace engine {

static std::mutex g_mu;
static std::map<pj_ctx *, Catalog *> g_catalogs;

Catalog &catalog_of(pj_ctx *ctx) {
  std::lock_guard<std::mutex> g(g_mu);
  Catalog *&c = g_catalogs[ctx];
  if (!c) c = new Catalog();
  return *c;
}

bool parse_type(const char *s, Type *t) {
  if (!strcmp(s, "int") || !strcmp(s, "int32") || !strcmp(s, "dstring")) {
    *t = Type::I32;
  } else if (!strcmp(s, "int64") || !strcmp(s, "date")) {
    *t = Type::I64;
  } else if (!strcmp(s, "float") || !strcmp(s, "double")) {
    *t = Type::F64;
  } else if (!strcmp(s, "bool")) {
    *t = Type::BOOL;
  } else if (!strcmp(s, "float32")) {
    *t = Type::F32;
  } else {
    return false;
  }
  return true;
}

int type_size(Type t) {
  switch (t) {
    case Type::I32: return 4;
    case Type::I64: return 8;
    case Type::F64: return 8;
    case Type::BOOL: return 1;
    case Type::F32: return 4;
  }
  return 0;
}

// Frees a column the engine staged; a copy still in flight is waited for
// first, so the memory is not reused under it.
void release_column(pj_ctx *ctx, Column &c) {
  if (c.ticket) {
    pj_stage_wait(ctx, c.ticket);
    c.ticket = 0;
  }
  if (c.staged && c.dev) pj_free(ctx, c.dev);
  c.dev = nullptr;
  c.staged = false;
}

void release_staged(pj_ctx *ctx, Catalog &cat) {
  for (auto &r : cat.rels)
    for (auto &c : r.second.cols) release_column(ctx, c.second);
  for (auto &f : cat.files) release_column(ctx, f.second.col);
  cat.files.clear();
}

}  // namespace engine
}  // namespace pj

using namespace pj::engine;

static thread_local char g_engine_err[1024];

extern "C" int pj_engine_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_engine_err, sizeof(g_engine_err), fmt, ap);
  va_end(ap);
  pj_set_last_error(g_engine_err);
  return code;
}

extern "C" int pj_register_column(pj_ctx *ctx, const char *rel, const char *attr,
                                  const char *type, const void *ptr, uint64_t n,
                                  int on_device) {
  if (!rel || !attr || !type || (!ptr && n))
    return pj_fail(PJ_E_ARG, "pj_register_column: null argument");
  Type t;
  if (!parse_type(type, &t))
    return pj_fail(PJ_E_UNSUPPORTED, "pj_register_column: unsupported type '%s'", type);
  Catalog &cat = catalog_of(ctx);
  Column &old = cat.rels[rel].cols[attr];
  release_column(ctx, old);
  Column c;
  c.type = t;
  c.n = n;
  if (on_device)
    c.dev = const_cast<void *>(ptr);
  else
    c.host = ptr;
  old = c;
  return PJ_OK;
}

extern "C" int pj_clear_catalog(pj_ctx *ctx) {
  Catalog &cat = catalog_of(ctx);
  release_staged(ctx, cat);
  cat.rels.clear();
  return PJ_OK;
}

extern "C" void pj_engine_release_catalog(pj_ctx *ctx) {
  Catalog *cat = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_catalogs.find(ctx);
    if (it == g_catalogs.end()) return;
    cat = it->second;
    g_catalogs.erase(it);
  }
  release_staged(ctx, *cat);
  delete cat;
}
