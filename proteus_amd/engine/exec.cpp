// Plan executor: runs the reference's JSON physical plans
// (core/olap/lib/plan/plan-parser.cpp:238-1660) on one MI355X.
//
// Execution is operator-at-a-time over device columns: every operator
// returns a relation (named columns in HBM plus a row count).  Placement
// operators of the reference (cpu-to-gpu, gpu-to-cpu, unpack, pack,
// mem-move-device, router, ...) carry no work here because every column
// already lives in HBM.  Expressions compile to the stack VM of rel.hip;
// joins and group-bys go to the hot-path kernels; a reduce directly over a
// join is fused into the join's probe (the reference's pipelines emit
// matches straight into the parent's consume, hash-join-chained.cpp:321-566).
// Only the attributes that some operator above a join or select reads are
// carried through it.
#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sys/stat.h>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../csrc/rel_api.hpp"
#include "engine.hpp"
#include "json.hpp"

namespace pj {
namespace engine {

namespace R = ::pj::rel;

struct EngineError {
  int code;
  std::string msg;
};

[[noreturn]] static void raise(int code, const std::string &msg) { throw EngineError{code, msg}; }

static void ck(int rc) {
  if (rc != PJ_OK) raise(rc, pj_last_error());
}

struct DBuf {
  pj_ctx *ctx;
  void *p;
  DBuf(pj_ctx *c, void *q) : ctx(c), p(q) {}
  ~DBuf() {
    if (p) pj_free(ctx, p);
  }
};
using BufPtr = std::shared_ptr<DBuf>;
using Dict = std::vector<std::string>;  // code -> string

struct Col {
  int t = R::I32;
  const void *dev = nullptr;
  BufPtr owner;               // keeps dev alive when the engine allocated it
  std::shared_ptr<Dict> dict; // string columns: dictionary codes
};

struct Attr {
  std::string rel, name;
};

struct Rel {
  uint64_t n = 0;
  std::vector<std::pair<Attr, Col>> cols;

  const Col *find(const std::string &r, const std::string &a) const {
    for (auto &c : cols)
      if (c.first.rel == r && c.first.name == a) return &c.second;
    for (auto &c : cols)
      if (c.first.name == a) return &c.second;
    return nullptr;
  }
  void add(const Attr &a, const Col &c) {
    for (auto &x : cols)
      if (x.first.rel == a.rel && x.first.name == a.name) {
        x.second = c;
        return;
      }
    cols.emplace_back(a, c);
  }
};

// Attribute names read by the operators above the current one.
struct Need {
  bool all = false;
  std::set<std::string> names;
  bool has(const std::string &n) const { return all || names.count(n) != 0; }
};

static const char *kPassThrough[] = {"unpack", "block-to-tuples", "pack", "tuples-to-block",
                                     "cpu-to-gpu", "gpu-to-cpu", "mem-move-device",
                                     "mem-move-local-to", "mem-broadcast-device", "router",
                                     "exchange", "split"};

static bool is_pass_through(const std::string &op) {
  for (const char *p : kPassThrough)
    if (op == p) return true;
  return false;
}

// Join operators.  out-of-gpu-join (plan-parser.cpp:531-850) bucket-joins
// inputs that live in host memory; here every column of a plan is already in
// HBM, so it runs as the radix-partitioned join (pj_join_out_of_gpu_reduce
// is the call for host-resident inputs).
static bool is_join_op(const std::string &k) {
  return k == "hashjoin-chained" || k == "partitioned-hashjoin-chained" || k == "join" ||
         k == "out-of-gpu-join";
}

static const char *kInputs[] = {"input", "build_input", "probe_input", "leftInput", "rightInput"};

// Every attrName mentioned in j outside operator inputs.
static void collect_names(const Json &j, std::set<std::string> &out, bool top = true) {
  if (j.is_object()) {
    for (auto &m : j.members()) {
      bool is_input = false;
      for (const char *k : kInputs)
        if (top && m.first == k) is_input = true;
      if (is_input) continue;
      if (m.first == "attrName" && m.second.is_string()) out.insert(m.second.as_string());
      if (m.first == "attribute" && m.second.is_string()) out.insert(m.second.as_string());
      collect_names(m.second, out, false);
    }
  } else if (j.is_array()) {
    for (auto &v : j.items()) collect_names(v, out, false);
  }
}

static Need need_below(const Json &op, const Need &above) {
  Need n = above;
  if (!n.all) collect_names(op, n.names);
  return n;
}

static int parse_col_type(const Json &t, bool *is_string) {
  *is_string = false;
  const Json &tt = t.is_object() ? t["type"] : t;
  if (!tt.is_string()) raise(PJ_E_PARSE, "expression type is not a string");
  const std::string &s = tt.as_string();
  if (s == "int") return R::I32;
  if (s == "int64" || s == "date" || s == "datetime") return R::I64;
  if (s == "float") return R::F64;
  if (s == "bool") return R::BOOL;
  if (s == "string" || s == "dstring") {
    *is_string = true;
    return R::I32;
  }
  raise(PJ_E_UNSUPPORTED, "type '" + s + "' is not supported on the device path");
}

static int promote(int a, int b) {
  if (a == R::F64 || b == R::F64 || a == R::F32 || b == R::F32) return R::F64;
  if (a == R::I64 || b == R::I64) return R::I64;
  return R::I32;
}

static std::string fmt_double(double v) {
  char b[40];
  for (int p = 1; p <= 17; ++p) {
    snprintf(b, sizeof b, "%.*g", p, v);
    if (strtod(b, nullptr) == v) break;
  }
  return b;
}

class Exec {
 public:
  // PROTEUS_AMD_NO_PIPELINE=1 runs every join operator-at-a-time (the
  // fused probe pipelines off), for cross-checking the two executions.
  Exec(pj_ctx *ctx, std::string base) : ctx_(ctx), base_(std::move(base)) {
    const char *e = getenv("PROTEUS_AMD_NO_PIPELINE");
    no_pipeline_ = e && *e && strcmp(e, "0") != 0;
  }

  std::string run(const Json &plan) {
    Need top;
    top.all = true;
    prefetch(plan);
    Rel r = exec(plan, top);
    if (!printed_) print_rel(r, nullptr, "pm-csv");
    return out_;
  }

 private:
  pj_ctx *ctx_;
  std::string base_;
  std::string out_;
  bool printed_ = false;
  std::unordered_map<std::string, std::vector<std::string>> file_cache_;

  // ---- memory --------------------------------------------------------
  BufPtr alloc(uint64_t bytes) {
    void *p = nullptr;
    ck(pj_malloc(ctx_, bytes ? bytes : 1, &p));
    return std::make_shared<DBuf>(ctx_, p);
  }
  Col upload(int t, const void *host, uint64_t n) {
    Col c;
    c.t = t;
    c.owner = alloc(n * R::col_size(t));
    if (n) ck(pj_memcpy_h2d(ctx_, c.owner->p, host, n * R::col_size(t)));
    c.dev = c.owner->p;
    return c;
  }
  std::string path(const std::string &p) const {
    if (!p.empty() && p[0] == '/') return p;
    return base_.empty() ? p : base_ + "/" + p;
  }

  // ---- expressions ---------------------------------------------------
  struct Compiler {
    Exec &ex;
    const Rel &in;
    R::Prog p;
    int sp = 0;

    Compiler(Exec &e, const Rel &r) : ex(e), in(r) {}

    const Col &column(const Json &e) {
      const Json &a = e["attribute"];
      std::string rel, name;
      if (a.is_object()) {
        rel = a["relName"].is_string() ? a["relName"].as_string() : "";
        name = a["attrName"].as_string();
      } else if (a.is_string()) {
        name = a.as_string();
      } else {
        raise(PJ_E_PARSE, "recordProjection without attribute");
      }
      const Col *c = in.find(rel, name);
      if (!c) raise(PJ_E_PARSE, "attribute " + rel + "." + name + " not found in the input");
      return *c;
    }

    bool is_col(const Json &e) const { return e["expression"].as_string() == "recordProjection"; }

    int infer(const Json &e) {
      const std::string &k = e["expression"].as_string();
      if (k == "int" || k == "dstring") return R::I32;
      if (k == "int64" || k == "date" || k == "datetime") return R::I64;
      if (k == "float") return R::F64;
      if (k == "bool") return R::BOOL;
      if (k == "string") return R::I32;
      if (k == "recordProjection") {
        int t = column(e).t;
        return t == R::F32 ? R::F64 : t;
      }
      if (k == "cast") {
        bool s;
        return parse_col_type(e["type"], &s);
      }
      if (k == "add" || k == "sub" || k == "multiply" || k == "div" || k == "mod")
        return promote(infer(e["left"]), infer(e["right"]));
      if (k == "neg") return promote(infer(e["e"]), R::I32);
      if (k == "eq" || k == "neq" || k == "lt" || k == "le" || k == "gt" || k == "ge" ||
          k == "and" || k == "or")
        return R::BOOL;
      if (k == "if") return promote(infer(e["then"]), infer(e["else"]));
      raise(PJ_E_UNSUPPORTED, "expression '" + k + "' is not supported on the device path");
    }

    void push(int op, int t, int a, int64_t imm) {
      if (p.n >= R::kMaxInstr) raise(PJ_E_UNSUPPORTED, "expression longer than 64 instructions");
      R::Instr &i = p.ins[p.n++];
      i.op = op;
      i.t = t;
      i.a = a;
      i.pad = 0;
      i.imm = imm;
      if (op == R::OP_COL || op == R::OP_CONST) {
        ++sp;
        if (sp > p.depth) p.depth = sp;
      } else if (op == R::OP_IF) {
        sp -= 2;
      } else if (op != R::OP_NEG && op != R::OP_NOT && op != R::OP_CAST && op != R::OP_BLOOM) {
        --sp;
      }
    }

    int bind(const Col &c) {
      for (int i = 0; i < p.ncols; ++i)
        if (p.cols[i] == c.dev) return i;
      if (p.ncols >= R::kMaxCols) raise(PJ_E_UNSUPPORTED, "expression reads more than 16 columns");
      p.cols[p.ncols] = c.dev;
      p.ctype[p.ncols] = c.t;
      return p.ncols++;
    }

    static int64_t fbits(double d) {
      int64_t v;
      memcpy(&v, &d, 8);
      return v;
    }

    // String constant compared with a dictionary column: its code (-1 when
    // absent, which matches nothing).
    int64_t string_code(const Json &s, const Json &other) {
      if (!is_col(other)) raise(PJ_E_UNSUPPORTED, "string constant outside a comparison with a column");
      const Col &c = column(other);
      if (!c.dict) raise(PJ_E_UNSUPPORTED, "string constant compared with a non-string column");
      const std::string &v = s["v"].as_string();
      for (size_t i = 0; i < c.dict->size(); ++i)
        if ((*c.dict)[i] == v) return (int64_t)i;
      return -1;
    }

    void emit(const Json &e, int want, const Json *peer = nullptr) {
      int t = emit_natural(e, peer);
      if (t != want) push(R::OP_CAST, want, t, 0);
    }

    int emit_natural(const Json &e, const Json *peer) {
      const std::string &k = e["expression"].as_string();
      if (k == "int") {
        push(R::OP_CONST, R::I32, 0, (int32_t)e["v"].as_int());
        return R::I32;
      }
      if (k == "int64" || k == "date" || k == "datetime") {
        if (!e["v"].is_number()) raise(PJ_E_UNSUPPORTED, "date constants must be integers");
        push(R::OP_CONST, R::I64, 0, e["v"].as_int());
        return R::I64;
      }
      if (k == "float") {
        push(R::OP_CONST, R::F64, 0, fbits(e["v"].as_double()));
        return R::F64;
      }
      if (k == "bool") {
        push(R::OP_CONST, R::BOOL, 0, e["v"].as_bool() ? 1 : 0);
        return R::BOOL;
      }
      if (k == "dstring") {
        int64_t code;
        if (e["v"].is_number()) code = e["v"].as_int();
        else code = ex.dict_lookup(e["v"].as_string(), e["dict"]);
        push(R::OP_CONST, R::I32, 0, code);
        return R::I32;
      }
      if (k == "string") {
        if (!peer) raise(PJ_E_UNSUPPORTED, "string constant outside a comparison");
        push(R::OP_CONST, R::I32, 0, string_code(e, *peer));
        return R::I32;
      }
      if (k == "recordProjection") {
        const Col &c = column(e);
        push(R::OP_COL, c.t, bind(c), 0);
        return c.t == R::F32 ? R::F64 : c.t;
      }
      if (k == "cast") {
        bool s;
        int to = parse_col_type(e["type"], &s);
        int from = emit_natural(e["e"], nullptr);
        if (from != to) push(R::OP_CAST, to, from, 0);
        return to;
      }
      static const struct { const char *n; int op; } bin[] = {
          {"add", R::OP_ADD}, {"sub", R::OP_SUB}, {"multiply", R::OP_MUL},
          {"div", R::OP_DIV}, {"mod", R::OP_MOD}};
      for (auto &b : bin)
        if (k == b.n) {
          int t = promote(infer(e["left"]), infer(e["right"]));
          emit(e["left"], t);
          emit(e["right"], t);
          push(b.op, t, 0, 0);
          return t;
        }
      static const struct { const char *n; int op; } cmp[] = {
          {"eq", R::OP_EQ}, {"neq", R::OP_NE}, {"lt", R::OP_LT},
          {"le", R::OP_LE}, {"gt", R::OP_GT}, {"ge", R::OP_GE}};
      for (auto &c : cmp)
        if (k == c.n) {
          const Json &l = e["left"], &r = e["right"];
          int t = promote(infer(l), infer(r));
          emit(l, t, &r);
          emit(r, t, &l);
          push(c.op, t, 0, 0);
          return R::BOOL;
        }
      if (k == "and" || k == "or") {
        emit(e["left"], R::BOOL);
        emit(e["right"], R::BOOL);
        push(k == "and" ? R::OP_AND : R::OP_OR, R::BOOL, 0, 0);
        return R::BOOL;
      }
      if (k == "neg") {
        int t = promote(infer(e["e"]), R::I32);
        emit(e["e"], t);
        push(R::OP_NEG, t, 0, 0);
        return t;
      }
      if (k == "if") {
        int t = promote(infer(e["then"]), infer(e["else"]));
        emit(e["cond"], R::BOOL);
        emit(e["then"], t);
        emit(e["else"], t);
        push(R::OP_IF, t, 0, 0);
        return t;
      }
      raise(PJ_E_UNSUPPORTED, "expression '" + k + "' is not supported on the device path");
    }
  };

  int64_t dict_lookup(const std::string &v, const Json &dict) {
    if (!dict["path"].is_string()) raise(PJ_E_PARSE, "dstring constant without dict.path");
    std::shared_ptr<Dict> d = read_dict(path(dict["path"].as_string()));
    for (size_t i = 0; i < d->size(); ++i)
      if ((*d)[i] == v) return (int64_t)i;
    raise(PJ_E_PARSE, "value '" + v + "' not in dictionary " + dict["path"].as_string());
  }

  // Column holding e over every row of in (no copy for a plain projection).
  Col eval(const Json &e, const Rel &in, int *type_out = nullptr) {
    Compiler c(*this, in);
    if (c.is_col(e)) {
      Col col = c.column(e);
      if (type_out) *type_out = col.t;
      return col;
    }
    int t = c.infer(e);
    c.emit(e, t);
    Col out;
    out.t = t;
    out.owner = alloc(in.n * R::col_size(t));
    out.dev = out.owner->p;
    ck(R::eval(ctx_, c.p, in.n, out.owner->p, t));
    if (type_out) *type_out = t;
    return out;
  }

  Col eval_as(const Json &e, const Rel &in, int want) {
    Compiler c(*this, in);
    if (c.is_col(e) && c.column(e).t == want) return c.column(e);
    c.emit(e, want);
    Col out;
    out.t = want;
    out.owner = alloc(in.n * R::col_size(want));
    out.dev = out.owner->p;
    ck(R::eval(ctx_, c.p, in.n, out.owner->p, want));
    return out;
  }

  static bool is_const(const Json &e) {
    const std::string &k = e["expression"].as_string();
    return k == "int" || k == "int64" || k == "float" || k == "bool";
  }

  static Attr reg_attr(const Json &e, const std::string &def_rel, const std::string &def_name) {
    const Json &r = e["register_as"];
    if (r.is_object())
      return Attr{r["relName"].is_string() ? r["relName"].as_string() : def_rel,
                  r["attrName"].as_string()};
    if (e["expression"].as_string() == "recordProjection") {
      const Json &a = e["attribute"];
      if (a.is_object())
        return Attr{a["relName"].is_string() ? a["relName"].as_string() : def_rel,
                    a["attrName"].as_string()};
      if (a.is_string()) return Attr{def_rel, a.as_string()};
    }
    return Attr{def_rel, def_name};
  }

  Rel gather_rel(const Rel &in, const Col &idx, uint64_t n, const Need &need) {
    Rel out;
    out.n = n;
    for (auto &c : in.cols) {
      if (!need.has(c.first.name)) continue;
      Col g;
      g.t = c.second.t;
      g.dict = c.second.dict;
      g.owner = alloc(n * R::col_size(g.t));
      g.dev = g.owner->p;
      ck(R::gather(ctx_, c.second.dev, R::col_size(g.t), (const uint32_t *)idx.dev, n, g.owner->p));
      out.cols.emplace_back(c.first, g);
    }
    return out;
  }

  // ---- operators -----------------------------------------------------
  Rel exec(const Json &op, const Need &need) {
    if (!op.is_object() || !op["operator"].is_string()) raise(PJ_E_PARSE, "plan node without operator");
    const std::string &k = op["operator"].as_string();
    if (is_pass_through(k)) return exec(op["input"], need);
    if (k == "scan") return scan(op);
    if (k == "select" || k == "project") {
      Rel r;
      if (!no_pipeline_ && pipeline(op, need, &r)) return r;
      return k == "select" ? select(op, need) : project(op, need);
    }
    if (k == "reduce") return reduce(op, need);
    if (k == "groupby" || k == "hashgroupby-chained") return groupby(op, need);
    if (is_join_op(k)) {
      Rel r;
      if (!no_pipeline_ && pipeline(op, need, &r)) return r;
      return join(op, need);
    }
    if (k == "hash-rearrange" || k == "hash-pack") return rearrange(op, need);
    if (k == "bloomfilter-build") return bloom_build(op, need);
    if (k == "bloomfilter-probe" || k == "bloomfilter-repack") {
      Rel r;
      if (!no_pipeline_ && pipeline(op, need, &r)) return r;
      return bloom_filter(op, need);
    }
    if (k == "sort") return sort(op, need);
    if (k == "print") return print(op, need);
    raise(PJ_E_UNSUPPORTED, "operator '" + k + "' is not supported on the device path");
  }

  // scan: registered relations, csv files, bincol files
  Rel scan(const Json &op) {
    const Json &pl = op["plugin"];
    if (!pl.is_object()) raise(PJ_E_PARSE, "scan without plugin");
    const std::string name = pl["name"].as_string();
    const std::string type = pl["type"].is_string() ? pl["type"].as_string() : "block";
    const Json &projs = pl["projections"];
    if (type == "csv" || type == "pm-csv") return scan_csv(pl, name);
    if (type == "bincol") return scan_bincol(pl, name);
    // every other plugin type resolves to the caller's registered columns,
    // else to the column files of a prepared relation
    Catalog &cat = catalog_of(ctx_);
    auto it = cat.rels.find(name);
    if (it == cat.rels.end()) return scan_block(pl, name);
    Rel out;
    bool have_n = false;
    for (auto &c : it->second.cols) {
      if (!have_n) {
        out.n = c.second.n;
        have_n = true;
      }
    }
    for (auto &p : projs.items()) {
      const std::string attr = p["attrName"].as_string();
      auto ci = it->second.cols.find(attr);
      if (ci == it->second.cols.end()) raise(PJ_E_PARSE, "column " + name + "." + attr + " is not registered");
      Column &cc = ci->second;
      stage_registered(cc);
      if (cc.ticket) {
        const uint64_t t = cc.ticket;
        cc.ticket = 0;
        ck(pj_stage_wait(ctx_, t));
      }
      Col col;
      col.t = cc.type == Type::I32 ? R::I32 : cc.type == Type::I64 ? R::I64
              : cc.type == Type::F64 ? R::F64 : cc.type == Type::F32 ? R::F32 : R::BOOL;
      col.dev = cc.dev;
      out.n = cc.n;
      out.cols.emplace_back(Attr{name, attr}, col);
    }
    for (auto &c : it->second.cols)
      if (c.second.n != out.n) raise(PJ_E_ARG, "columns of relation '" + name + "' differ in length");
    return out;
  }

  // ---- host -> HBM staging ---------------------------------------------
  // A registered host column is copied to HBM on first use through the
  // pinned staging ring (pj_stage_begin); the scan that reads it orders the
  // context's stream after the copy.
  void stage_registered(Column &cc) {
    if (cc.dev || !cc.n) return;
    void *d = nullptr;
    ck(pj_malloc(ctx_, cc.n * type_size(cc.type), &d));
    cc.dev = d;
    cc.staged = true;
    ck(pj_stage_begin(ctx_, nullptr, cc.host, 0, cc.n * type_size(cc.type), d, &cc.ticket));
  }

  // Column file <name>.<attr> of a prepared relation (the layout of
  // tools/prepare_data.py:374-443: native little-endian int32 for "int",
  // int64 for "int64" / "date" / "datetime", FP64 for "float", int32 codes
  // for "dstring" with the dictionary <file>.dict of "value:code" lines).
  // Staged to HBM on first use and kept while the file's size and mtime are
  // unchanged.
  FileColumn &file_column(const std::string &name, const Json &proj) {
    const std::string attr = proj["attrName"].as_string();
    bool is_str;
    const int t = parse_col_type(proj_type(proj, name), &is_str);
    if (t == R::BOOL) raise(PJ_E_UNSUPPORTED, "block column " + name + "." + attr + " of type bool");
    const std::string file = path(name + "." + attr);
    struct stat sb;
    if (stat(file.c_str(), &sb) != 0)
      raise(PJ_E_PARSE, "relation '" + name + "' is not registered and has no column file " + file);
    const int64_t mt = (int64_t)sb.st_mtim.tv_sec * 1000000000LL + sb.st_mtim.tv_nsec;
    const uint64_t esz = R::col_size(t);
    if ((uint64_t)sb.st_size % esz)
      raise(PJ_E_PARSE, "column file " + file + " is not a whole number of " + std::to_string(esz) + "-byte values");
    Catalog &cat = catalog_of(ctx_);
    auto it = cat.files.find(file);
    if (it != cat.files.end()) {
      FileColumn &f = it->second;
      if (f.size == (uint64_t)sb.st_size && f.mtime_ns == mt && f.col.type == col_type(t)) return f;
      release_column(ctx_, f.col);
      cat.files.erase(it);
    }
    FileColumn f;
    f.size = (uint64_t)sb.st_size;
    f.mtime_ns = mt;
    f.col.type = col_type(t);
    f.col.n = f.size / esz;
    if (is_str) f.dict = read_dict(file + ".dict");
    void *d = nullptr;
    ck(pj_malloc(ctx_, std::max<uint64_t>(f.size, 1), &d));
    uint64_t ticket = 0;
    if (pj_stage_begin(ctx_, file.c_str(), nullptr, 0, f.size, d, &ticket) != PJ_OK) {
      const std::string msg = pj_last_error();
      pj_free(ctx_, d);
      raise(PJ_E_PARSE, msg);
    }
    f.col.dev = d;
    f.col.staged = true;
    f.col.ticket = ticket;
    return cat.files[file] = f;
  }

  // Orders the context's stream after a file column's copy; a failed copy
  // drops the column from the cache.
  void wait_file(const std::string &file, FileColumn &f) {
    if (!f.col.ticket) return;
    const uint64_t t = f.col.ticket;
    f.col.ticket = 0;
    if (pj_stage_wait(ctx_, t) != PJ_OK) {
      const std::string msg = pj_last_error();
      release_column(ctx_, f.col);
      catalog_of(ctx_).files.erase(file);
      raise(PJ_E_PARSE, msg);
    }
  }

  static Type col_type(int t) {
    return t == R::I64 ? Type::I64 : t == R::F64 ? Type::F64 : t == R::F32 ? Type::F32 : Type::I32;
  }

  // "value:code" lines, the value ending at the last ':' (storage-manager.cpp:314-330)
  std::shared_ptr<Dict> read_dict(const std::string &file) {
    std::ifstream is(file);
    if (!is) raise(PJ_E_PARSE, "cannot open dictionary " + file);
    auto d = std::make_shared<Dict>();
    std::string line;
    while (std::getline(is, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      if (line.empty()) continue;
      const size_t k = line.find_last_of(':');
      if (k == std::string::npos) raise(PJ_E_PARSE, "dictionary line without ':' in " + file);
      char *end = nullptr;
      const long long code = strtoll(line.c_str() + k + 1, &end, 10);
      if (code < 0 || code > INT32_MAX || *end) raise(PJ_E_PARSE, "bad dictionary code in " + file);
      if ((size_t)code >= d->size()) d->resize((size_t)code + 1);
      (*d)[(size_t)code] = line.substr(0, k);
    }
    return d;
  }

  Rel scan_block(const Json &pl, const std::string &name) {
    Rel out;
    bool first = true;
    for (auto &p : pl["projections"].items()) {
      FileColumn &f = file_column(name, p);
      wait_file(path(name + "." + p["attrName"].as_string()), f);
      if (!first && f.col.n != out.n) raise(PJ_E_PARSE, "column files of " + name + " differ in length");
      out.n = f.col.n;
      first = false;
      Col c;
      c.t = f.col.type == Type::I64 ? R::I64 : f.col.type == Type::F64 ? R::F64 : R::I32;
      c.dev = f.col.dev;
      c.dict = f.dict;
      out.cols.emplace_back(Attr{name, p["attrName"].as_string()}, c);
    }
    if (first) raise(PJ_E_PARSE, "relation '" + name + "' is not registered, and a scan of its column files needs projections");
    return out;
  }

  // Starts the staging of every host-resident column the plan scans, in the
  // order the operators run (a join's build side before its probe side), so
  // the later scans' copies overlap the earlier operators' kernels.
  void prefetch(const Json &op) {
    if (!op.is_object() || !op["operator"].is_string()) return;
    const std::string &k = op["operator"].as_string();
    if (k == "scan") {
      const Json &pl = op["plugin"];
      if (!pl.is_object() || !pl["name"].is_string()) return;
      const std::string type = pl["type"].is_string() ? pl["type"].as_string() : "block";
      if (type == "csv" || type == "pm-csv" || type == "bincol") return;
      const std::string name = pl["name"].as_string();
      Catalog &cat = catalog_of(ctx_);
      auto it = cat.rels.find(name);
      for (auto &p : pl["projections"].items()) {
        if (!p["attrName"].is_string()) continue;
        if (it != cat.rels.end()) {
          auto ci = it->second.cols.find(p["attrName"].as_string());
          if (ci != it->second.cols.end()) stage_registered(ci->second);
        } else {
          file_column(name, p);
        }
      }
      return;
    }
    if (k == "join") {
      prefetch(op["leftInput"]);
      prefetch(op["rightInput"]);
      return;
    }
    if (is_join_op(k)) {
      prefetch(op["build_input"]);
      prefetch(op["probe_input"]);
      return;
    }
    prefetch(op["input"]);
  }

  const std::vector<std::string> &lines_of(const std::string &file) {
    auto it = file_cache_.find(file);
    if (it != file_cache_.end()) return it->second;
    std::ifstream is(path(file));
    if (!is) raise(PJ_E_PARSE, "cannot open " + path(file));
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(is, l)) {
      if (!l.empty() && l.back() == '\r') l.pop_back();
      if (!l.empty()) lines.push_back(l);
    }
    return file_cache_[file] = std::move(lines);
  }

  // Attribute rel.attr (type, attrNo) from a catalog.json next to the data
  // or under inputs/.
  bool catalog_attr(const std::string &rel, const std::string &attr, Json *out) {
    std::vector<std::string> cands;
    size_t sl = rel.find_last_of('/');
    if (sl != std::string::npos) cands.push_back(rel.substr(0, sl) + "/catalog.json");
    cands.push_back("inputs/plans/catalog.json");
    cands.push_back("inputs/catalog.json");
    for (auto &c : cands) {
      std::ifstream is(path(c));
      if (!is) continue;
      std::stringstream ss;
      ss << is.rdbuf();
      Json j;
      std::string err;
      if (!json_parse(ss.str(), &j, &err)) continue;
      for (auto &m : j.members()) {
        const Json &e = m.second;
        if (!e["path"].is_string() || e["path"].as_string() != rel) continue;
        const Json &at = e["type"]["inner"]["attributes"];
        for (auto &a : at.items())
          if (a["attrName"].is_string() && a["attrName"].as_string() == attr) {
            *out = a;
            return true;
          }
      }
    }
    return false;
  }

  Json proj_type(const Json &p, const std::string &rel) {
    if (p["type"].is_object()) return p["type"];
    Json a;
    if (catalog_attr(rel, p["attrName"].as_string(), &a) && a["type"].is_object()) return a["type"];
    raise(PJ_E_PARSE, "no type for " + rel + "." + p["attrName"].as_string());
  }

  Rel scan_csv(const Json &pl, const std::string &name) {
    const auto &lines = lines_of(name);
    char delim = ';';
    if (pl["delimiter"].is_string() && !pl["delimiter"].as_string().empty())
      delim = pl["delimiter"].as_string()[0];
    Rel out;
    out.n = lines.size();
    std::vector<std::vector<std::string>> fields;
    fields.reserve(lines.size());
    for (auto &l : lines) {
      std::vector<std::string> f;
      std::string cur;
      for (char ch : l) {
        if (ch == delim) {
          f.push_back(cur);
          cur.clear();
        } else {
          cur += ch;
        }
      }
      f.push_back(cur);
      fields.push_back(std::move(f));
    }
    for (auto &p : pl["projections"].items()) {
      const std::string attr = p["attrName"].as_string();
      Json tj = proj_type(p, name);
      int no = p["attrNo"].is_number() ? (int)p["attrNo"].as_int() : 0;
      if (no <= 0) {
        Json a;
        if (catalog_attr(name, attr, &a) && a["attrNo"].is_number()) no = (int)a["attrNo"].as_int();
      }
      if (no <= 0) raise(PJ_E_PARSE, "csv projection " + attr + " without attrNo");
      bool is_str;
      int t = parse_col_type(tj, &is_str);
      Col c;
      if (is_str) {
        auto dict = std::make_shared<Dict>();
        std::unordered_map<std::string, int32_t> codes;
        std::vector<int32_t> v(out.n);
        for (uint64_t i = 0; i < out.n; ++i) {
          const std::string s = (size_t)no <= fields[i].size() ? fields[i][no - 1] : "";
          auto it = codes.find(s);
          if (it == codes.end()) {
            it = codes.emplace(s, (int32_t)dict->size()).first;
            dict->push_back(s);
          }
          v[i] = it->second;
        }
        c = upload(R::I32, v.data(), out.n);
        c.dict = dict;
      } else if (t == R::F64) {
        std::vector<double> v(out.n);
        for (uint64_t i = 0; i < out.n; ++i)
          v[i] = (size_t)no <= fields[i].size() ? strtod(fields[i][no - 1].c_str(), nullptr) : 0.0;
        c = upload(t, v.data(), out.n);
      } else if (t == R::I64) {
        std::vector<int64_t> v(out.n);
        for (uint64_t i = 0; i < out.n; ++i)
          v[i] = (size_t)no <= fields[i].size() ? strtoll(fields[i][no - 1].c_str(), nullptr, 10) : 0;
        c = upload(t, v.data(), out.n);
      } else if (t == R::BOOL) {
        std::vector<uint8_t> v(out.n);
        for (uint64_t i = 0; i < out.n; ++i) {
          const std::string s = (size_t)no <= fields[i].size() ? fields[i][no - 1] : "";
          v[i] = s == "true" || s == "1" || s == "TRUE";
        }
        c = upload(t, v.data(), out.n);
      } else {
        std::vector<int32_t> v(out.n);
        for (uint64_t i = 0; i < out.n; ++i)
          v[i] = (size_t)no <= fields[i].size() ? (int32_t)strtol(fields[i][no - 1].c_str(), nullptr, 10) : 0;
        c = upload(t, v.data(), out.n);
      }
      out.cols.emplace_back(Attr{name, attr}, c);
    }
    return out;
  }

  // bincol: one file per column, "<name>.<attrName>", an int64 row count
  // followed by the values.
  Rel scan_bincol(const Json &pl, const std::string &name) {
    Rel out;
    bool first = true;
    for (auto &p : pl["projections"].items()) {
      const std::string attr = p["attrName"].as_string();
      bool is_str;
      int t = parse_col_type(proj_type(p, name), &is_str);
      std::ifstream is(path(name + "." + attr), std::ios::binary);
      if (!is) raise(PJ_E_PARSE, "cannot open " + path(name + "." + attr));
      int64_t cnt = 0;
      is.read((char *)&cnt, 8);
      if (!is || cnt < 0) raise(PJ_E_PARSE, "bad bincol header in " + name + "." + attr);
      std::vector<char> buf((size_t)cnt * R::col_size(t));
      is.read(buf.data(), (std::streamsize)buf.size());
      if (!is) raise(PJ_E_PARSE, "short bincol file " + name + "." + attr);
      if (!first && (uint64_t)cnt != out.n) raise(PJ_E_PARSE, "bincol columns of " + name + " differ in length");
      out.n = (uint64_t)cnt;
      first = false;
      out.cols.emplace_back(Attr{name, attr}, upload(t, buf.data(), out.n));
    }
    if (first) raise(PJ_E_UNSUPPORTED, "bincol scan without projections has no row count");
    return out;
  }

  Rel select(const Json &op, const Need &need) {
    Need below = need_below(op, need);
    Rel in = exec(op["input"], below);
    return filter_rel(in, op["p"], need);
  }

  // The rows of `in` where `pred` holds, in order.
  Rel filter_rel(const Rel &in, const Json &pred, const Need &need) {
    Compiler c(*this, in);
    c.emit(pred, R::BOOL);
    Col idx;
    idx.t = R::I32;
    idx.owner = alloc(std::max<uint64_t>(in.n, 1) * 4);
    idx.dev = idx.owner->p;
    uint64_t m = 0;
    ck(R::select(ctx_, c.p, in.n, (uint32_t *)idx.owner->p, &m));
    return gather_rel(in, idx, m, need);
  }

  Rel project(const Json &op, const Need &need) {
    Rel in = exec(op["input"], need_below(op, need));
    const std::string rel = op["relName"].is_string() ? op["relName"].as_string() : "project";
    Rel out;
    out.n = in.n;
    int i = 0;
    for (auto &e : op["e"].items()) {
      Col c = eval(e, in);
      if (Compiler(*this, in).is_col(e)) c.dict = Compiler(*this, in).column(e).dict;
      out.add(reg_attr(e, rel, "expr" + std::to_string(i++)), c);
    }
    return out;
  }

  // ---- bloom filters -------------------------------------------------
  // bloomfilter-build / -probe / -repack {input, e, filterSize, bloomId
  // [, projections]}: the JSON form of RelBuilder::bloomfilter_build /
  // _probe / _repack (relbuilder.cpp:957-975; the reference's plan parser has
  // no bloom operators).  The key expression is int32, as the reference's
  // `e & (int32)(filterSize - 1)` (bloom-filter.cpp:85-107).
  static uint64_t bloom_id(const Json &op) {
    if (!op["bloomId"].is_number()) raise(PJ_E_PARSE, "bloom filter operator without bloomId");
    return (uint64_t)op["bloomId"].as_int();
  }
  static uint64_t bloom_size(const Json &op) {
    if (!op["filterSize"].is_number() || op["filterSize"].as_int() <= 0)
      raise(PJ_E_PARSE, "bloom filter operator without a positive filterSize");
    return (uint64_t)op["filterSize"].as_int();
  }
  // The filter of a probe or repack, checked against its filterSize.
  void bloom_of(const Json &op, const uint32_t **words, int *lg) {
    ck(R::bloom_lookup(ctx_, bloom_id(op), words, lg));
    const uint64_t size = bloom_size(op);
    if ((*lg >= 64 ? 0 : (1ULL << *lg)) != size)
      raise(PJ_E_ARG, "bloom filter " + std::to_string(bloom_id(op)) + " was built with " +
                          std::to_string(1ULL << *lg) + " entries, probed with " +
                          std::to_string(size));
  }

  Rel bloom_build(const Json &op, const Need &need) {
    Rel in = exec(op["input"], need_below(op, need));
    if (Compiler(*this, in).infer(op["e"]) != R::I32)
      raise(PJ_E_UNSUPPORTED, "bloom filter key must be an int32 expression");
    Col k = eval_as(op["e"], in, R::I32);
    ck(pj_bloom_build_i32(ctx_, bloom_id(op), bloom_size(op), (const int32_t *)k.dev, in.n));
    return in;
  }

  // The rows whose key's filter entry is set; repack keeps the projections.
  Rel bloom_filter(const Json &op, const Need &need) {
    // a repack without projections keeps every input attribute, as
    // RelBuilder::bloomfilter_repack(pred, filterSize, bloomId) (relbuilder.hpp:427-436)
    const bool repack = op["operator"].as_string() == "bloomfilter-repack" && op["projections"].is_array();
    Need below = need_below(op, need);
    Rel in = exec(op["input"], below);
    const uint32_t *words = nullptr;
    int lg = 0;
    bloom_of(op, &words, &lg);
    Compiler c(*this, in);
    if (c.infer(op["e"]) != R::I32) raise(PJ_E_UNSUPPORTED, "bloom filter key must be an int32 expression");
    c.emit(op["e"], R::I32);
    c.push(R::OP_BLOOM, R::BOOL, lg, (int64_t)(uintptr_t)words);
    Col idx;
    idx.t = R::I32;
    idx.owner = alloc(std::max<uint64_t>(in.n, 1) * 4);
    idx.dev = idx.owner->p;
    uint64_t m = 0;
    ck(R::select(ctx_, c.p, in.n, (uint32_t *)idx.owner->p, &m));
    if (!repack) return gather_rel(in, idx, m, need);
    Rel sel = gather_rel(in, idx, m, below);
    Rel out;
    out.n = m;
    int i = 0;
    for (auto &e : op["projections"].items()) {
      Col col = eval(e, sel);
      if (Compiler(*this, sel).is_col(e)) col.dict = Compiler(*this, sel).column(e).dict;
      out.add(reg_attr(e, "repack", "expr" + std::to_string(i++)), col);
    }
    return out;
  }

  // ---- reduce --------------------------------------------------------
  static int monoid_of(const std::string &m) {
    if (m == "sum") return R::M_SUM;
    if (m == "max") return R::M_MAX;
    if (m == "min") return R::M_MIN;
    if (m == "and") return R::M_AND;
    if (m == "or") return R::M_OR;
    raise(PJ_E_UNSUPPORTED, "accumulator '" + m + "' is not supported on the device path");
  }

  // A scalar result, stored as the expression's type.
  struct Scalar {
    int t;
    int64_t i;
    double f;
  };

  static Scalar finish(int t, int64_t i, double f) {
    Scalar s{t, i, f};
    if (t == R::I32) s.i = (int32_t)(uint32_t)(uint64_t)i;
    if (t == R::BOOL) s.i = i != 0;
    return s;
  }

  Rel one_row(const std::vector<std::pair<Attr, Scalar>> &vals) {
    Rel out;
    out.n = 1;
    for (auto &v : vals) {
      const Scalar &s = v.second;
      Col c;
      if (s.t == R::F64) c = upload(R::F64, &s.f, 1);
      else if (s.t == R::I64) c = upload(R::I64, &s.i, 1);
      else if (s.t == R::BOOL) {
        uint8_t b = s.i != 0;
        c = upload(R::BOOL, &b, 1);
      } else {
        int32_t x = (int32_t)s.i;
        c = upload(R::I32, &x, 1);
      }
      out.cols.emplace_back(v.first, c);
    }
    return out;
  }

  // Reduce of a constant over n rows, without touching the device.
  static Scalar const_reduce(const Json &e, int t, int m, uint64_t n) {
    const std::string &k = e["expression"].as_string();
    bool f = k == "float";
    int64_t v = k == "bool" ? (e["v"].as_bool() ? 1 : 0) : (f ? 0 : e["v"].as_int());
    double d = f ? e["v"].as_double() : (double)v;
    switch (m) {
      case R::M_SUM:
        return finish(t, (int64_t)((uint64_t)v * n), d * (double)n);
      case R::M_MIN:
      case R::M_MAX:
        if (n) return finish(t, v, d);
        return finish(t, m == R::M_MIN ? INT64_MAX : INT64_MIN,
                      m == R::M_MIN ? HUGE_VAL : -HUGE_VAL);
      case R::M_AND: return finish(t, n ? v != 0 : 1, 0);
      default: return finish(t, n ? v != 0 : 0, 0);
    }
  }

  Rel reduce(const Json &op, const Need &need) {
    // p: only the rows where it holds are accumulated (opt::Reduce's
    // predicate, reduce-opt.cpp); a missing p or literal true takes them all
    const Json &p = op["p"];
    const bool filtered =
        p.is_object() && !(p["expression"].as_string() == "bool" && p["v"].as_bool());
    const Json &es = op["e"], &accs = op["accumulator"];
    if (!es.is_array() || !accs.is_array() || es.size() != accs.size())
      raise(PJ_E_PARSE, "reduce needs as many accumulators as expressions");
    // fused join -> reduce
    const Json *below = &op["input"];
    while (below->is_object() && is_pass_through((*below)["operator"].as_string()))
      below = &(*below)["input"];
    const std::string bk = (*below)["operator"].is_string() ? (*below)["operator"].as_string() : "";
    if (!filtered && is_join_op(bk)) {
      Rel r;
      if (fused_join_reduce(op, *below, need, &r)) return r;
    }
    const Need nb = need_below(op, need);
    Rel in = exec(op["input"], nb);
    if (filtered) in = filter_rel(in, p, nb);
    std::vector<std::pair<Attr, Scalar>> vals;
    for (size_t i = 0; i < es.size(); ++i) {
      const Json &e = es[i];
      int m = monoid_of(accs[i].as_string());
      Compiler c(*this, in);
      int t = c.infer(e);
      Scalar s;
      if (is_const(e)) {
        s = const_reduce(e, t, m, in.n);
      } else {
        Col col = eval(e, in);
        int64_t vi = 0;
        double vf = 0;
        ck(R::reduce(ctx_, col.dev, col.t, in.n, m, &vi, &vf));
        s = finish(t, vi, vf);
      }
      vals.emplace_back(reg_attr(e, "reduce", "agg" + std::to_string(i)), s);
    }
    return one_row(vals);
  }

  struct JoinSides {
    const Json *build_in, *probe_in, *build_k, *probe_k;
  };

  static JoinSides join_sides(const Json &j) {
    const std::string &k = j["operator"].as_string();
    if (k == "join") {
      const Json &p = j["p"];
      if (p["expression"].as_string() != "eq") raise(PJ_E_UNSUPPORTED, "join predicate is not an equality");
      return JoinSides{&j["leftInput"], &j["rightInput"], &p["left"], &p["right"]};
    }
    if (k == "out-of-gpu-join") {  // the key is the first of build_e / probe_e
      const Json &be = j["build_e"], &pe = j["probe_e"];
      if (!be.is_array() || !pe.is_array() || !be.size() || !pe.size())
        raise(PJ_E_PARSE, "out-of-gpu-join without build_e / probe_e");
      return JoinSides{&j["build_input"], &j["probe_input"], &be[(size_t)0]["original"],
                       &pe[(size_t)0]["original"]};
    }
    return JoinSides{&j["build_input"], &j["probe_input"], &j["build_k"], &j["probe_k"]};
  }

  static int bits_for(uint64_t n) {
    int b = 0;
    while ((1ULL << b) < n && b < 32) ++b;
    return b + 1;
  }

  // Reduce over a join whose expressions are COUNT (sum of a constant),
  // sum/min/max of one int32 column of either side, or sum of a build plus
  // a probe column: one probe kernel with the aggregates fused.
  bool fused_join_reduce(const Json &op, const Json &j, const Need &need, Rel *out) {
    JoinSides js = join_sides(j);
    Need nb = need_below(j, need_below(op, need));
    Rel b = exec(*js.build_in, nb), p = exec(*js.probe_in, nb);
    struct Term {
      int kind;  // 0 const, 1 build col, 2 probe col, 3 build+probe
      const Json *e;
    };
    const Col *bcol = nullptr, *pcol = nullptr;
    std::vector<Term> terms;
    auto side_col = [&](const Json &e, int *side) -> const Col * {
      const Json *x = &e;
      while ((*x)["expression"].as_string() == "cast") x = &(*x)["e"];
      if ((*x)["expression"].as_string() != "recordProjection") return nullptr;
      Compiler cb(*this, b), cp(*this, p);
      const Json &a = (*x)["attribute"];
      std::string rel = a.is_object() && a["relName"].is_string() ? a["relName"].as_string() : "";
      std::string name = a.is_object() ? a["attrName"].as_string() : a.as_string();
      bool in_b = false, in_p = false;
      for (auto &c : b.cols) in_b |= c.first.rel == rel && c.first.name == name;
      for (auto &c : p.cols) in_p |= c.first.rel == rel && c.first.name == name;
      if (!in_b && !in_p) {
        in_b = b.find("", name) != nullptr;
        in_p = p.find("", name) != nullptr;
      }
      if (in_b == in_p) return nullptr;
      const Col *c = in_b ? b.find(rel, name) : p.find(rel, name);
      if (!c || c->t != R::I32) return nullptr;
      *side = in_b ? 1 : 2;
      return c;
    };
    auto claim = [&](const Col *c, int side) {
      const Col *&slot = side == 1 ? bcol : pcol;
      if (slot && slot->dev != c->dev) return false;
      slot = c;
      return true;
    };
    const Json &es = op["e"], &accs = op["accumulator"];
    for (size_t i = 0; i < es.size(); ++i) {
      const Json &e = es[i];
      const std::string m = accs[i].as_string();
      if (is_const(e) && e["expression"].as_string() != "float" && m == "sum") {
        terms.push_back({0, &e});
        continue;
      }
      int side = 0;
      const Col *c = side_col(e, &side);
      if (c && (m == "sum" || m == "min" || m == "max")) {
        if (!claim(c, side)) return false;
        terms.push_back({side, &e});
        continue;
      }
      if (m == "sum" && e["expression"].as_string() == "add") {
        int s1 = 0, s2 = 0;
        const Col *c1 = side_col(e["left"], &s1), *c2 = side_col(e["right"], &s2);
        if (c1 && c2 && s1 != s2 && claim(c1, s1) && claim(c2, s2)) {
          terms.push_back({3, &e});
          continue;
        }
      }
      return false;
    }
    Col bkc = eval_as(*js.build_k, b, R::I32), pkc = eval_as(*js.probe_k, p, R::I32);
    Compiler kb(*this, b), kp(*this, p);
    if (kb.infer(*js.build_k) != R::I32 || kp.infer(*js.probe_k) != R::I32)
      raise(PJ_E_UNSUPPORTED, "join keys must be int32 expressions");
    pj_join_agg agg{};
    const std::string &k = j["operator"].as_string();
    const uint64_t max_build = j["maxBuildInputSize"].is_number() ? (uint64_t)j["maxBuildInputSize"].as_int() : b.n;
    if (k == "hashjoin-chained") {
      int hb = j["hash_bits"].is_number() ? (int)j["hash_bits"].as_int() : bits_for(b.n);
      ck(pj_join_chained_reduce(ctx_, (const int32_t *)bkc.dev, bcol ? (const int32_t *)bcol->dev : nullptr,
                                b.n, (const int32_t *)pkc.dev, pcol ? (const int32_t *)pcol->dev : nullptr,
                                p.n, hb, max_build, &agg));
    } else {
      const uint64_t max_probe = j["maxProbeInputSize"].is_number() ? (uint64_t)j["maxProbeInputSize"].as_int() : p.n;
      ck(pj_join_partitioned_reduce(ctx_, (const int32_t *)bkc.dev, bcol ? (const int32_t *)bcol->dev : nullptr,
                                    b.n, (const int32_t *)pkc.dev, pcol ? (const int32_t *)pcol->dev : nullptr,
                                    p.n, max_build, max_probe, nullptr, &agg));
    }
    std::vector<std::pair<Attr, Scalar>> vals;
    for (size_t i = 0; i < es.size(); ++i) {
      const Json &e = *terms[i].e;
      const std::string m = accs[i].as_string();
      Compiler cc(*this, terms[i].kind == 2 ? p : b);
      int t = terms[i].kind == 0 ? (e["expression"].as_string() == "int" ? R::I32
                                    : e["expression"].as_string() == "bool" ? R::BOOL : R::I64)
              : terms[i].kind == 3 ? promote(infer_side(e["left"], b, p), infer_side(e["right"], b, p))
                                   : infer_side(e, b, p);
      int64_t v = 0;
      const bool any = agg.count > 0;
      switch (terms[i].kind) {
        case 0: {
          int64_t c = e["expression"].as_string() == "bool" ? (e["v"].as_bool() ? 1 : 0) : e["v"].as_int();
          v = (int64_t)((uint64_t)c * (uint64_t)agg.count);
          break;
        }
        case 1:
          v = m == "sum" ? agg.sum_b : m == "min" ? (any ? agg.min_b : INT64_MAX) : (any ? agg.max_b : INT64_MIN);
          break;
        case 2:
          v = m == "sum" ? agg.sum_p : m == "min" ? (any ? agg.min_p : INT64_MAX) : (any ? agg.max_p : INT64_MIN);
          break;
        default: v = (int64_t)((uint64_t)agg.sum_b + (uint64_t)agg.sum_p); break;
      }
      vals.emplace_back(reg_attr(e, "reduce", "agg" + std::to_string(i)), finish(t, v, (double)v));
    }
    *out = one_row(vals);
    return true;
  }

  int infer_side(const Json &e, const Rel &b, const Rel &p) {
    try {
      return Compiler(*this, b).infer(e);
    } catch (const EngineError &) {
      return Compiler(*this, p).infer(e);
    }
  }

  // ---- group-by ------------------------------------------------------
  Rel groupby(const Json &op, const Need &need) {
    Rel in = exec(op["input"], need_below(op, need));
    const Json &ks = op["k"], &es = op["e"];
    if (!ks.is_array() || !es.is_array()) raise(PJ_E_PARSE, "groupby needs k and e arrays");
    uint64_t max_groups = op["maxInputSize"].is_number() ? (uint64_t)op["maxInputSize"].as_int() : 0;
    if (!max_groups && op["hash_bits"].is_number()) max_groups = 1ULL << op["hash_bits"].as_int();
    if (!max_groups) max_groups = in.n;
    std::vector<Col> keys;
    std::vector<Attr> knames;
    for (size_t i = 0; i < ks.size(); ++i) {
      Col c = eval(ks[i], in);
      Compiler cc(*this, in);
      if (cc.is_col(ks[i])) c.dict = cc.column(ks[i]).dict;
      keys.push_back(c);
      knames.push_back(reg_attr(ks[i], "groupby", "key" + std::to_string(i)));
    }
    struct AggIn {
      Col col;
      int m;
      int t;         // output type
      int64_t mult;  // sum of a constant: count * mult
      Attr name;
    };
    std::vector<AggIn> aggs;
    for (size_t i = 0; i < es.size(); ++i) {
      const Json &e = es[i]["e"];
      int m = monoid_of(es[i]["m"].as_string());
      Compiler cc(*this, in);
      int t = cc.infer(e);
      AggIn a;
      a.t = t;
      a.m = m;
      a.mult = 1;
      a.name = reg_attr(e, "groupby", "agg" + std::to_string(i));
      if (is_const(e) && m == R::M_SUM && e["expression"].as_string() != "float") {
        a.m = R::M_CNT;
        a.mult = e["expression"].as_string() == "bool" ? (e["v"].as_bool() ? 1 : 0) : e["v"].as_int();
      } else {
        a.col = eval(e, in);
      }
      aggs.push_back(a);
    }
    Col kout_lds;
    std::vector<Col> aout_lds;
    uint64_t ng_lds = 0;
    if (groupby_lds(in, keys, max_groups, aggs, kout_lds, aout_lds, &ng_lds)) {
      Rel out;
      out.n = ng_lds;
      out.cols.emplace_back(knames[0], kout_lds);
      for (size_t i = 0; i < aggs.size(); ++i)
        out.cols.emplace_back(aggs[i].name, finish_agg(aout_lds[i], aggs[i].mult, aggs[i].t, ng_lds));
      return out;
    }
    if (keys.size() > (size_t)R::kGbMaxKeys) raise(PJ_E_UNSUPPORTED, "group-by over more than 4 keys");
    if (aggs.size() > (size_t)R::kGbMaxAggs) raise(PJ_E_UNSUPPORTED, "group-by with more than 16 aggregates");
    R::GbKey gk[R::kGbMaxKeys];
    R::GbAgg ga[R::kGbMaxAggs];
    R::GbOut go{};
    std::vector<Col> kout, aout;
    for (size_t i = 0; i < keys.size(); ++i) {
      gk[i] = R::GbKey{keys[i].dev, keys[i].t};
      Col c;
      c.t = keys[i].t;
      c.dict = keys[i].dict;
      c.owner = alloc(std::min(max_groups, std::max<uint64_t>(in.n, 1)) * R::col_size(c.t));
      c.dev = c.owner->p;
      go.keys[i] = c.owner->p;
      kout.push_back(c);
    }
    for (size_t i = 0; i < aggs.size(); ++i) {
      ga[i] = R::GbAgg{aggs[i].col.dev, aggs[i].col.t, aggs[i].m};
      const bool f = aggs[i].m != R::M_CNT && (aggs[i].col.t == R::F64 || aggs[i].col.t == R::F32);
      Col c;
      c.t = f ? R::F64 : R::I64;
      c.owner = alloc(std::min(max_groups, std::max<uint64_t>(in.n, 1)) * 8);
      c.dev = c.owner->p;
      go.aggs[i] = c.owner->p;
      aout.push_back(c);
    }
    uint64_t ng = 0;
    if (in.n) ck(R::groupby(ctx_, gk, (int)keys.size(), ga, (int)aggs.size(), in.n, max_groups, go, &ng));
    Rel out;
    out.n = ng;
    for (size_t i = 0; i < keys.size(); ++i) out.cols.emplace_back(knames[i], kout[i]);
    for (size_t i = 0; i < aggs.size(); ++i)
      out.cols.emplace_back(aggs[i].name, finish_agg(aout[i], aggs[i].mult, aggs[i].t, ng));
    return out;
  }

  // Accumulators are int64 / FP64: scale counts by `mult` (a sum of a
  // constant), narrow to the expression type t.
  Col finish_agg(Col c, int64_t mult, int t, uint64_t ng) {
    if (mult == 1 && t == c.t) return c;
    Rel tmp;
    tmp.n = ng;
    tmp.cols.emplace_back(Attr{"", "v"}, c);
    Compiler cc(*this, tmp);
    cc.push(R::OP_COL, c.t, cc.bind(c), 0);
    if (mult != 1) {
      cc.push(R::OP_CONST, R::I64, 0, mult);
      cc.push(R::OP_MUL, R::I64, 0, 0);
    }
    if (t != c.t) cc.push(R::OP_CAST, t, c.t, 0);
    Col o;
    o.t = t;
    o.owner = alloc(std::max<uint64_t>(ng, 1) * R::col_size(t));
    o.dev = o.owner->p;
    if (ng) ck(R::eval(ctx_, cc.p, ng, o.owner->p, t));
    return o;
  }

  // The LDS group-by (pj_groupby_i32: per-workgroup LDS tables, radix
  // partitioning above one table's groups, no global atomics) for the shape
  // it computes: one int32 key; COUNT; SUM / MIN / MAX of one integer column;
  // SUM of one float32 column.  Returns false (nothing run) for any other
  // group-by, which the general kernel takes.
  template <class AggIn>
  bool groupby_lds(const Rel &in, const std::vector<Col> &keys, uint64_t max_groups,
                   const std::vector<AggIn> &aggs, Col &kout, std::vector<Col> &aout,
                   uint64_t *n_groups) {
    if (getenv("PROTEUS_AMD_NO_LDS_GROUPBY")) return false;
    if (keys.size() != 1 || keys[0].t != R::I32 || aggs.empty()) return false;
    const void *vcol = nullptr, *wcol = nullptr;
    int vt = -1;
    enum { K_CNT, K_SUM, K_MIN, K_MAX, K_FSUM };
    std::vector<int> kind;
    for (const AggIn &a : aggs) {
      if (a.m == R::M_CNT) {
        kind.push_back(K_CNT);
        continue;
      }
      const bool integer = a.col.t == R::I32 || a.col.t == R::I64;
      if (integer && (a.m == R::M_SUM || a.m == R::M_MIN || a.m == R::M_MAX)) {
        if (vcol && vcol != a.col.dev) return false;  // one value column only
        vcol = a.col.dev;
        vt = a.col.t;
        kind.push_back(a.m == R::M_SUM ? K_SUM : a.m == R::M_MIN ? K_MIN : K_MAX);
      } else if (a.col.t == R::F32 && a.m == R::M_SUM) {
        if (wcol && wcol != a.col.dev) return false;
        wcol = a.col.dev;
        kind.push_back(K_FSUM);
      } else {
        return false;
      }
    }
    const uint64_t cap = std::min(max_groups, std::max<uint64_t>(in.n, 1));
    // int32 values are widened to int64 (the kernel's value column)
    Col v64;
    if (vcol && vt == R::I32) {
      Rel tmp;
      tmp.n = in.n;
      Col c;
      c.t = R::I32;
      c.dev = vcol;
      tmp.cols.emplace_back(Attr{"", "v"}, c);
      Compiler cc(*this, tmp);
      cc.push(R::OP_COL, R::I32, cc.bind(c), 0);
      cc.push(R::OP_CAST, R::I64, R::I32, 0);
      v64.t = R::I64;
      v64.owner = alloc(std::max<uint64_t>(in.n, 1) * 8);
      v64.dev = v64.owner->p;
      if (in.n) ck(R::eval(ctx_, cc.p, in.n, v64.owner->p, R::I64));
      vcol = v64.dev;
    }
    kout.t = R::I32;
    kout.dict = keys[0].dict;
    kout.owner = alloc(cap * 4);
    kout.dev = kout.owner->p;
    BufPtr outs[5];
    void *ptr[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int k : kind)
      if (!outs[k]) {
        outs[k] = alloc(cap * 8);
        ptr[k] = outs[k]->p;
      }
    *n_groups = 0;
    if (in.n)
      ck(pj_groupby_i32(ctx_, (const int32_t *)keys[0].dev, (const int64_t *)vcol,
                        (const float *)wcol, in.n, 0, cap, (int32_t *)kout.owner->p,
                        (int64_t *)ptr[K_CNT], (int64_t *)ptr[K_SUM], (int64_t *)ptr[K_MIN],
                        (int64_t *)ptr[K_MAX], (double *)ptr[K_FSUM], n_groups));
    aout.clear();
    for (int k : kind) {
      Col c;
      c.t = k == K_FSUM ? R::F64 : R::I64;
      c.owner = outs[k];
      c.dev = outs[k]->p;
      aout.push_back(c);
    }
    return true;
  }

  // ---- joins ---------------------------------------------------------
  Rel join(const Json &op, const Need &need) {
    JoinSides js = join_sides(op);
    Need nb = need_below(op, need);
    Rel b = exec(*js.build_in, nb), p = exec(*js.probe_in, nb);
    Compiler kb(*this, b), kp(*this, p);
    if (kb.infer(*js.build_k) != R::I32 || kp.infer(*js.probe_k) != R::I32)
      raise(PJ_E_UNSUPPORTED, "join keys must be int32 expressions");
    Col bk = eval_as(*js.build_k, b, R::I32), pk = eval_as(*js.probe_k, p, R::I32);
    const std::string &k = op["operator"].as_string();
    int hb = k == "hashjoin-chained" && op["hash_bits"].is_number() ? (int)op["hash_bits"].as_int()
                                                                     : bits_for(b.n);
    const uint64_t max_build = op["maxBuildInputSize"].is_number() ? (uint64_t)op["maxBuildInputSize"].as_int() : b.n;
    uint64_t cap = std::max<uint64_t>(std::max(p.n, b.n), 1);
    Col ob, opr;
    uint64_t m = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
      ob = Col();
      opr = Col();
      ob.owner = alloc(cap * 4);
      opr.owner = alloc(cap * 4);
      ob.dev = ob.owner->p;
      opr.dev = opr.owner->p;
      int rc = pj_join_chained_pairs(ctx_, (const int32_t *)bk.dev, b.n, (const int32_t *)pk.dev, p.n, hb,
                                     max_build, (uint32_t *)ob.owner->p, (uint32_t *)opr.owner->p, cap, &m);
      if (rc == PJ_E_CAPACITY && m > cap) {
        cap = m;
        continue;
      }
      ck(rc);
      break;
    }
    Rel out = gather_rel(b, ob, m, need);
    Rel pr = gather_rel(p, opr, m, need);
    for (auto &c : pr.cols) out.add(c.first, c.second);
    out.n = m;
    return out;
  }

  // ---- fused probe pipelines -----------------------------------------
  // A chain of joins along the probe side, with select / project /
  // placement operators between them, runs as one kernel: the source rows
  // stream once; each join probes a unique-key table of its build side
  // (key -> build row id) and later expressions read build columns through
  // the matched row ids; only the attributes read above the chain are
  // written out.  Builds with duplicate keys (or the reserved key INT32_MIN)
  // fall back to operator-at-a-time execution.
  struct Binding {
    int level = 0;
    Col col;
    const Json *expr = nullptr;  // project output: expression in `scope`
    int scope = 0;
    int type = R::I32;
  };
  struct Scope {
    std::vector<std::pair<Attr, Binding>> attrs;
    const Binding *find(const std::string &r, const std::string &a) const {
      for (auto &x : attrs)
        if (x.first.rel == r && x.first.name == a) return &x.second;
      for (auto &x : attrs)
        if (x.first.name == a) return &x.second;
      return nullptr;
    }
  };

  struct PipeCompiler {
    Exec &ex;
    std::vector<Scope> &scopes;
    R::PipeProg p;
    int sp = 0;
    std::vector<std::pair<const void *, int>> bound;  // (column, level) -> index

    const Binding &resolve(const Json &e, int sc) {
      const Json &a = e["attribute"];
      std::string rel, name;
      if (a.is_object()) {
        rel = a["relName"].is_string() ? a["relName"].as_string() : "";
        name = a["attrName"].as_string();
      } else if (a.is_string()) {
        name = a.as_string();
      } else {
        raise(PJ_E_PARSE, "recordProjection without attribute");
      }
      const Binding *b = scopes[sc].find(rel, name);
      if (!b) raise(PJ_E_PARSE, "attribute " + rel + "." + name + " not found in the pipeline");
      return *b;
    }

    int infer(const Json &e, int sc) {
      const std::string &k = e["expression"].as_string();
      if (k == "recordProjection") {
        const Binding &b = resolve(e, sc);
        if (b.expr) return infer(*b.expr, b.scope);
        return b.col.t == R::F32 ? R::F64 : b.col.t;
      }
      if (k == "int" || k == "dstring" || k == "string") return R::I32;
      if (k == "int64" || k == "date" || k == "datetime") return R::I64;
      if (k == "float") return R::F64;
      if (k == "bool") return R::BOOL;
      if (k == "cast") {
        bool s;
        return parse_col_type(e["type"], &s);
      }
      if (k == "add" || k == "sub" || k == "multiply" || k == "div" || k == "mod")
        return promote(infer(e["left"], sc), infer(e["right"], sc));
      if (k == "neg") return promote(infer(e["e"], sc), R::I32);
      if (k == "eq" || k == "neq" || k == "lt" || k == "le" || k == "gt" || k == "ge" || k == "and" ||
          k == "or")
        return R::BOOL;
      if (k == "if") return promote(infer(e["then"], sc), infer(e["else"], sc));
      raise(PJ_E_UNSUPPORTED, "expression '" + k + "' is not supported on the device path");
    }

    void push(int op, int t, int a, int64_t imm) {
      if (p.ins.size() >= 256) raise(PJ_E_UNSUPPORTED, "pipeline program longer than 256 instructions");
      R::Instr i{op, t, a, 0, imm};
      p.ins.push_back(i);
      if (op == R::OP_COL || op == R::OP_CONST) {
        if (++sp > p.depth) p.depth = sp;
      } else if (op == R::OP_IF) {
        sp -= 2;
      } else if (op == R::OP_PROBE || op == R::OP_FILTER || op == R::OP_OUT) {
        --sp;
      } else if (op != R::OP_NEG && op != R::OP_NOT && op != R::OP_CAST && op != R::OP_BLOOM) {
        --sp;
      }
    }

    int bind(const Col &c, int level) {
      for (size_t i = 0; i < bound.size(); ++i)
        if (bound[i].first == c.dev && bound[i].second == level) return (int)i;
      bound.emplace_back(c.dev, level);
      p.cols.push_back(c.dev);
      p.ctype.push_back(c.t);
      p.clevel.push_back(level);
      return (int)bound.size() - 1;
    }

    const std::shared_ptr<Dict> *dict_of(const Json &e, int sc) {
      if (e["expression"].as_string() != "recordProjection") return nullptr;
      const Binding &b = resolve(e, sc);
      if (b.expr) return dict_of(*b.expr, b.scope);
      return b.col.dict ? &b.col.dict : nullptr;
    }

    void emit(const Json &e, int sc, int want, const Json *peer = nullptr) {
      int t = emit_natural(e, sc, peer);
      if (t != want) push(R::OP_CAST, want, t, 0);
    }

    int emit_natural(const Json &e, int sc, const Json *peer) {
      const std::string &k = e["expression"].as_string();
      if (k == "recordProjection") {
        const Binding &b = resolve(e, sc);
        if (b.expr) return emit_natural(*b.expr, b.scope, peer);
        push(R::OP_COL, b.col.t, bind(b.col, b.level), 0);
        return b.col.t == R::F32 ? R::F64 : b.col.t;
      }
      if (k == "int") {
        push(R::OP_CONST, R::I32, 0, (int32_t)e["v"].as_int());
        return R::I32;
      }
      if (k == "int64" || k == "date" || k == "datetime") {
        push(R::OP_CONST, R::I64, 0, e["v"].as_int());
        return R::I64;
      }
      if (k == "float") {
        double d = e["v"].as_double();
        int64_t v;
        memcpy(&v, &d, 8);
        push(R::OP_CONST, R::F64, 0, v);
        return R::F64;
      }
      if (k == "bool") {
        push(R::OP_CONST, R::BOOL, 0, e["v"].as_bool() ? 1 : 0);
        return R::BOOL;
      }
      if (k == "dstring") {
        int64_t code = e["v"].is_number() ? e["v"].as_int() : ex.dict_lookup(e["v"].as_string(), e["dict"]);
        push(R::OP_CONST, R::I32, 0, code);
        return R::I32;
      }
      if (k == "string") {
        const std::shared_ptr<Dict> *d = peer ? dict_of(*peer, sc) : nullptr;
        if (!d) raise(PJ_E_UNSUPPORTED, "string constant outside a comparison with a string column");
        int64_t code = -1;
        for (size_t i = 0; i < (*d)->size(); ++i)
          if ((**d)[i] == e["v"].as_string()) code = (int64_t)i;
        push(R::OP_CONST, R::I32, 0, code);
        return R::I32;
      }
      if (k == "cast") {
        bool s;
        int to = parse_col_type(e["type"], &s);
        int from = emit_natural(e["e"], sc, nullptr);
        if (from != to) push(R::OP_CAST, to, from, 0);
        return to;
      }
      static const struct { const char *n; int op; } bin[] = {
          {"add", R::OP_ADD}, {"sub", R::OP_SUB}, {"multiply", R::OP_MUL},
          {"div", R::OP_DIV}, {"mod", R::OP_MOD}};
      for (auto &b : bin)
        if (k == b.n) {
          int t = promote(infer(e["left"], sc), infer(e["right"], sc));
          emit(e["left"], sc, t);
          emit(e["right"], sc, t);
          push(b.op, t, 0, 0);
          return t;
        }
      static const struct { const char *n; int op; } cmp[] = {
          {"eq", R::OP_EQ}, {"neq", R::OP_NE}, {"lt", R::OP_LT},
          {"le", R::OP_LE}, {"gt", R::OP_GT}, {"ge", R::OP_GE}};
      for (auto &c : cmp)
        if (k == c.n) {
          const Json &l = e["left"], &r = e["right"];
          int t = promote(infer(l, sc), infer(r, sc));
          emit(l, sc, t, &r);
          emit(r, sc, t, &l);
          push(c.op, t, 0, 0);
          return R::BOOL;
        }
      if (k == "and" || k == "or") {
        emit(e["left"], sc, R::BOOL);
        emit(e["right"], sc, R::BOOL);
        push(k == "and" ? R::OP_AND : R::OP_OR, R::BOOL, 0, 0);
        return R::BOOL;
      }
      if (k == "neg") {
        int t = promote(infer(e["e"], sc), R::I32);
        emit(e["e"], sc, t);
        push(R::OP_NEG, t, 0, 0);
        return t;
      }
      if (k == "if") {
        int t = promote(infer(e["then"], sc), infer(e["else"], sc));
        emit(e["cond"], sc, R::BOOL);
        emit(e["then"], sc, t);
        emit(e["else"], sc, t);
        push(R::OP_IF, t, 0, 0);
        return t;
      }
      raise(PJ_E_UNSUPPORTED, "expression '" + k + "' is not supported on the device path");
    }
  };

  bool no_pipeline_ = false;

  bool pipeline(const Json &top, const Need &need, Rel *out) {
    // stages top -> bottom along the probe side
    std::vector<const Json *> stages;
    const Json *node = &top;
    int joins = 0;
    for (;;) {
      const std::string k = (*node)["operator"].is_string() ? (*node)["operator"].as_string() : "";
      if (is_pass_through(k)) {
        node = &(*node)["input"];
        continue;
      }
      if (is_join_op(k)) {
        stages.push_back(node);
        ++joins;
        node = join_sides(*node).probe_in;
        continue;
      }
      if (k == "select" || k == "project" || k == "bloomfilter-probe" || k == "bloomfilter-repack") {
        stages.push_back(node);
        node = &(*node)["input"];
        continue;
      }
      break;
    }
    if (!joins || joins >= R::kPipeMaxLevels) return false;
    Need na = need;
    for (const Json *st : stages) na = need_below(*st, na);
    // Every build side runs before the source, as the reference runs build
    // pipelines before the probe pipeline: a bloom filter built on a build
    // side can be probed below.
    // A build side select(p, x) is not compacted: x's rows are kept with p as
    // a byte mask, and the dimension table leaves out the rows p rejects
    // (the probe finds no row for their keys, as after the select).  This
    // saves the select's compaction, its count read-back and the gathers of
    // x's columns; the pipeline reads them at x's row ids.
    std::map<const Json *, Rel> built;
    std::map<const Json *, Col> keep;
    for (auto it = stages.rbegin(); it != stages.rend(); ++it) {
      const std::string k = (**it)["operator"].as_string();
      if (!is_join_op(k)) continue;
      const Json *bi = join_sides(**it).build_in;
      while ((*bi)["operator"].is_string() && is_pass_through((*bi)["operator"].as_string()))
        bi = &(*bi)["input"];
      if ((*bi)["operator"].is_string() && (*bi)["operator"].as_string() == "select") {
        Rel x = exec((*bi)["input"], need_below(*bi, na));
        keep.emplace(*it, eval_as((*bi)["p"], x, R::BOOL));
        if (!na.all) {  // the columns the select would have passed on
          std::vector<std::pair<Attr, Col>> cols;
          for (auto &c : x.cols)
            if (na.has(c.first.name)) cols.push_back(c);
          x.cols = std::move(cols);
        }
        built.emplace(*it, std::move(x));
      } else {
        built.emplace(*it, exec(*join_sides(**it).build_in, na));
      }
    }
    Rel src = exec(*node, na);
    std::vector<Rel> builds;
    std::vector<pj_dim *> dims;
    auto drop = [&]() {
      for (pj_dim *d : dims) pj_dim_destroy(ctx_, d);
      dims.clear();
    };
    std::vector<Scope> scopes(1);
    for (auto &c : src.cols) {
      Binding b;
      b.level = 0;
      b.col = c.second;
      scopes[0].attrs.emplace_back(c.first, b);
    }
    PipeCompiler pc{*this, scopes, R::PipeProg(), 0, {}};
    int cur = 0;
    bool ok = true;
    std::vector<Col> bkeys;  // per join: build keys, keep mask, build rows
    std::vector<const uint8_t *> bkeep;
    std::vector<uint64_t> brows;
    for (auto it = stages.rbegin(); it != stages.rend() && ok; ++it) {
      const Json &st = **it;
      const std::string k = st["operator"].as_string();
      if (k == "select") {
        pc.emit(st["p"], cur, R::BOOL);
        pc.push(R::OP_FILTER, R::BOOL, 0, 0);
      } else if (k == "project") {
        Scope ns;
        const std::string rel = st["relName"].is_string() ? st["relName"].as_string() : "project";
        int i = 0;
        for (auto &e : st["e"].items()) {
          Binding b;
          b.expr = &e;
          b.scope = cur;
          ns.attrs.emplace_back(reg_attr(e, rel, "expr" + std::to_string(i++)), b);
        }
        scopes.push_back(ns);
        cur = (int)scopes.size() - 1;
      } else if (k == "bloomfilter-probe" || k == "bloomfilter-repack") {
        const uint32_t *words = nullptr;
        int lg = 0;
        bloom_of(st, &words, &lg);
        if (pc.infer(st["e"], cur) != R::I32)
          raise(PJ_E_UNSUPPORTED, "bloom filter key must be an int32 expression");
        pc.emit(st["e"], cur, R::I32);
        pc.push(R::OP_BLOOM, R::BOOL, lg, (int64_t)(uintptr_t)words);
        pc.push(R::OP_FILTER, R::BOOL, 0, 0);
        if (k == "bloomfilter-repack" && st["projections"].is_array()) {  // the projections, by their names
          Scope ns;
          int i = 0;
          for (auto &e : st["projections"].items()) {
            Binding b;
            b.expr = &e;
            b.scope = cur;
            ns.attrs.emplace_back(reg_attr(e, "repack", "expr" + std::to_string(i++)), b);
          }
          scopes.push_back(ns);
          cur = (int)scopes.size() - 1;
        }
      } else {
        JoinSides js = join_sides(st);
        Rel b = std::move(built.at(*it));
        Compiler kb(*this, b);
        if (kb.infer(*js.build_k) != R::I32 || pc.infer(*js.probe_k, cur) != R::I32) {
          ok = false;
          break;
        }
        // the tables are built together after the loop
        bkeys.push_back(eval_as(*js.build_k, b, R::I32));
        auto km = keep.find(*it);
        bkeep.push_back(km == keep.end() ? nullptr : (const uint8_t *)km->second.dev);
        brows.push_back(b.n);
        const int tj = (int)pc.p.tables.size();
        pc.p.tables.push_back(R::PipeTable{});
        pc.emit(*js.probe_k, cur, R::I32);
        pc.push(R::OP_PROBE, R::I32, tj, 0);
        // the join's output = probe attributes + build attributes
        Scope ns = scopes[cur];
        for (auto &c : b.cols) {
          Binding bb;
          bb.level = tj + 1;
          bb.col = c.second;
          ns.attrs.emplace_back(c.first, bb);
        }
        scopes.push_back(ns);
        cur = (int)scopes.size() - 1;
        builds.push_back(std::move(b));
      }
    }
    if (ok) {  // the dimension tables (values = build rows), built together
      const int m = (int)bkeys.size();
      std::vector<const int32_t *> kp(m);
      for (int i = 0; i < m; ++i) kp[i] = (const int32_t *)bkeys[i].dev;
      dims.assign(m, nullptr);
      int rc = R::dims_create_rows(ctx_, m, kp.data(), bkeep.data(), brows.data(), dims.data());
      if (rc == PJ_E_ARG) {  // duplicate or reserved build keys
        dims.clear();
        ok = false;
      } else {
        ck(rc);
        for (int i = 0; i < m; ++i) ck(R::dim_table(dims[i], &pc.p.tables[i]));
      }
    }
    if (!ok) {
      drop();
      no_pipeline_ = true;
      Rel r = exec(top, need);
      no_pipeline_ = false;
      *out = std::move(r);
      return true;
    }
    // outputs: the attributes of the final scope read above the chain
    std::vector<std::pair<Attr, Col>> outs;
    std::vector<void *> optr;
    int j = 0;
    for (auto &a : scopes[cur].attrs) {
      if (!need.has(a.first.name)) continue;
      bool dup = false;
      for (auto &o : outs) dup |= o.first.rel == a.first.rel && o.first.name == a.first.name;
      if (dup) continue;
      Json ref = Json::object();
      Json attr = Json::object();
      attr.set("relName", Json::string(a.first.rel));
      attr.set("attrName", Json::string(a.first.name));
      ref.set("expression", Json::string("recordProjection"));
      ref.set("attribute", attr);
      int t = pc.infer(ref, cur);
      pc.emit(ref, cur, t);
      pc.push(R::OP_OUT, t, j++, 0);
      Col c;
      c.t = t;
      const std::shared_ptr<Dict> *d = pc.dict_of(ref, cur);
      if (d) c.dict = *d;
      c.owner = alloc(std::max<uint64_t>(src.n, 1) * R::col_size(t));
      c.dev = c.owner->p;
      optr.push_back(c.owner->p);
      pc.p.out_t.push_back(t);
      outs.emplace_back(a.first, c);
    }
    if (outs.empty()) {  // nothing read above (e.g. COUNT only): one dummy column
      pc.push(R::OP_CONST, R::BOOL, 0, 1);
      pc.push(R::OP_OUT, R::BOOL, 0, 0);
      Col c;
      c.t = R::BOOL;
      c.owner = alloc(std::max<uint64_t>(src.n, 1));
      c.dev = c.owner->p;
      optr.push_back(c.owner->p);
      pc.p.out_t.push_back(R::BOOL);
      outs.emplace_back(Attr{"", "__row"}, c);
    }
    uint64_t m = 0;
    int rc = R::pipeline(ctx_, pc.p, src.n, optr.data(), &m);
    drop();
    ck(rc);
    out->n = m;
    out->cols = std::move(outs);
    return true;
  }

  // ---- hash-rearrange ------------------------------------------------
  Rel rearrange(const Json &op, const Need &need) {
    Rel in = exec(op["input"], need_below(op, need));
    int buckets = op["buckets"].is_number() ? (int)op["buckets"].as_int() : 2;
    Compiler kc(*this, in);
    if (kc.infer(op["e"]) != R::I32) raise(PJ_E_UNSUPPORTED, "hash-rearrange key must be an int32 expression");
    Col key = eval_as(op["e"], in, R::I32);
    Col iota, ok, perm;
    iota.owner = alloc(in.n * 4);
    ok.owner = alloc(in.n * 4);
    perm.owner = alloc(in.n * 4);
    iota.dev = iota.owner->p;
    ok.dev = ok.owner->p;
    perm.dev = perm.owner->p;
    ok.t = R::I32;
    ck(R::iota(ctx_, (uint32_t *)iota.owner->p, in.n));
    std::vector<uint64_t> offs((size_t)buckets + 1);
    ck(pj_hash_rearrange_i32(ctx_, (const int32_t *)key.dev, (const int32_t *)iota.dev, in.n, (uint32_t)buckets,
                             (int32_t *)ok.owner->p, (int32_t *)perm.owner->p, offs.data()));
    Rel out;
    out.n = in.n;
    std::vector<const Json *> projs;
    for (auto &pj : op["projections"].items()) {
      if (pj["expression"].as_string() == "recordConstruction") {
        for (auto &a : pj["attributes"].items()) projs.push_back(&a["e"]);
      } else {
        projs.push_back(&pj);
      }
    }
    int i = 0;
    for (const Json *e : projs) {
      Col c = eval(*e, in);
      Compiler cc(*this, in);
      if (cc.is_col(*e)) c.dict = cc.column(*e).dict;
      Col g;
      g.t = c.t;
      g.dict = c.dict;
      g.owner = alloc(in.n * R::col_size(c.t));
      g.dev = g.owner->p;
      ck(R::gather(ctx_, c.dev, R::col_size(c.t), (const uint32_t *)perm.dev, in.n, g.owner->p));
      out.add(reg_attr(*e, "rearrange", "expr" + std::to_string(i++)), g);
    }
    if (op["hashProject"].is_object()) {
      const Json &h = op["hashProject"];
      out.add(Attr{h["relName"].is_string() ? h["relName"].as_string() : "", h["attrName"].as_string()}, ok);
    }
    return out;
  }

  // ---- sort (host side: it sits above gpu-to-cpu in the plans) --------
  Rel sort(const Json &op, const Need &need) {
    Rel in = exec(op["input"], need_below(op, need));
    const Json &es = op["e"];
    std::vector<Col> cols;
    std::vector<int> dir;
    for (auto &x : es.items()) {
      Col c = eval(x["expression"], in);
      Compiler cc(*this, in);
      if (cc.is_col(x["expression"])) c.dict = cc.column(x["expression"]).dict;
      cols.push_back(c);
      const std::string d = x["direction"].is_string() ? x["direction"].as_string() : "ASC";
      dir.push_back(d == "DESC" ? -1 : d == "NONE" ? 0 : 1);
    }
    // Stable device sort of a row permutation (sort.hip).  Dictionary
    // columns sort by their strings: each code maps to its rank in string
    // order (equal strings, equal ranks).
    std::vector<Col> ranks(cols.size());
    std::vector<R::SortKey> keys;
    for (size_t k = 0; k < cols.size(); ++k) {
      R::SortKey sk{cols[k].dev, cols[k].t, dir[k], nullptr, 0};
      if (cols[k].dict && dir[k]) {
        const Dict &d = *cols[k].dict;
        std::vector<uint32_t> order(d.size());
        for (size_t c = 0; c < d.size(); ++c) order[c] = (uint32_t)c;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return d[a] < d[b]; });
        std::vector<int32_t> rank(std::max<size_t>(d.size(), 1), 0);
        int32_t r = 0;
        for (size_t c = 0; c < order.size(); ++c) {
          if (c && d[order[c]] != d[order[c - 1]]) ++r;
          rank[order[c]] = r;
        }
        ranks[k] = upload(R::I32, rank.data(), rank.size());
        sk.rank = (const int32_t *)ranks[k].dev;
        sk.nrank = (uint32_t)d.size();
      }
      keys.push_back(sk);
    }
    Col idx;
    idx.t = R::I32;
    idx.owner = alloc(std::max<uint64_t>(in.n, 1) * 4);
    idx.dev = idx.owner->p;
    // a plain int32 first key comes back sorted from the passes themselves
    Col key0;
    bool key0_done = false;
    if (!cols.empty() && cols[0].t == R::I32 && !cols[0].dict) {
      key0.t = R::I32;
      key0.owner = alloc(std::max<uint64_t>(in.n, 1) * 4);
      key0.dev = key0.owner->p;
    }
    ck(R::sort_perm(ctx_, keys.data(), (int)keys.size(), in.n, (uint32_t *)idx.owner->p,
                    (int32_t *)key0.dev, &key0_done));
    Rel out;
    out.n = in.n;
    const Json &rt = op["rowType"];
    for (size_t k = 0; k < cols.size(); ++k) {
      Col g;
      g.t = cols[k].t;
      g.dict = cols[k].dict;
      if (k == 0 && key0_done) {
        g.owner = key0.owner;
        g.dev = key0.dev;
      } else {
        g.owner = alloc(in.n * R::col_size(g.t));
        g.dev = g.owner->p;
        ck(R::gather(ctx_, cols[k].dev, R::col_size(g.t), (const uint32_t *)idx.dev, in.n,
                     g.owner->p));
      }
      Attr a = reg_attr(es[k]["expression"], "sort", "key" + std::to_string(k));
      if (rt.is_array() && k < rt.size())
        a = Attr{rt[k]["relName"].as_string(), rt[k]["attrName"].as_string()};
      out.add(a, g);
    }
    return out;
  }

  // Every column's n rows to the host with one synchronisation: integers
  // and bools widened into iv[k], floats into fv[k].
  void load_host(const std::vector<Col> &cols, uint64_t n, std::vector<std::vector<int64_t>> *iv,
                 std::vector<std::vector<double>> *fv) {
    const size_t m = cols.size();
    std::vector<std::vector<uint8_t>> raw(m);
    std::vector<void *> dst;
    std::vector<const void *> src;
    std::vector<uint64_t> by;
    for (size_t k = 0; k < m; ++k) {
      raw[k].resize(n * R::col_size(cols[k].t));
      if (n) {
        dst.push_back(raw[k].data());
        src.push_back(cols[k].dev);
        by.push_back(raw[k].size());
      }
    }
    if (!dst.empty()) ck(R::d2h_many(ctx_, (int)dst.size(), dst.data(), src.data(), by.data()));
    iv->assign(m, {});
    fv->assign(m, {});
    for (size_t k = 0; k < m; ++k) {
      auto &I = (*iv)[k];
      auto &F = (*fv)[k];
      I.assign(n, 0);
      F.assign(n, 0.0);
      const uint8_t *p = raw[k].data();
      for (uint64_t i = 0; i < n; ++i) {
        switch (cols[k].t) {
          case R::I32: { int32_t x; memcpy(&x, p + 4 * i, 4); I[i] = x; break; }
          case R::I64: { int64_t x; memcpy(&x, p + 8 * i, 8); I[i] = x; break; }
          case R::F64: { double x; memcpy(&x, p + 8 * i, 8); F[i] = x; break; }
          case R::F32: { float x; memcpy(&x, p + 4 * i, 4); F[i] = x; break; }
          default: I[i] = p[i]; break;
        }
      }
    }
  }

  // ---- print ---------------------------------------------------------
  Rel print(const Json &op, const Need &need) {
    Need below = need;
    if (op["e"].is_array()) {
      below = Need();
      collect_names(op, below.names);
    }
    Rel in = exec(op["input"], below);
    std::string fmt = op["plugin"]["type"].is_string() ? op["plugin"]["type"].as_string() : "pm-csv";
    print_rel(in, op["e"].is_array() ? &op["e"] : nullptr, fmt);
    return in;
  }

  void print_rel(const Rel &in, const Json *es, const std::string &fmt) {
    std::vector<Col> cols;
    std::vector<std::string> names;
    if (es) {
      int i = 0;
      for (auto &e : es->items()) {
        Col c = eval(e, in);
        Compiler cc(*this, in);
        if (cc.is_col(e)) c.dict = cc.column(e).dict;
        cols.push_back(c);
        names.push_back(reg_attr(e, "", "f" + std::to_string(i++)).name);
      }
    } else {
      for (auto &c : in.cols) {
        cols.push_back(c.second);
        names.push_back(c.first.name);
      }
    }
    std::vector<std::vector<int64_t>> iv;
    std::vector<std::vector<double>> fv;
    load_host(cols, in.n, &iv, &fv);
    auto field = [&](size_t k, uint64_t r, bool json) -> std::string {
      const Col &c = cols[k];
      if (c.dict) {
        const std::string &s = (*c.dict)[iv[k][r]];
        return json ? "\"" + s + "\"" : s;
      }
      if (c.t == R::F64 || c.t == R::F32) return fmt_double(fv[k][r]);
      if (c.t == R::BOOL) return iv[k][r] ? "true" : "false";
      return std::to_string(iv[k][r]);
    };
    std::string o;
    if (fmt == "json") {
      o += '[';
      for (uint64_t r = 0; r < in.n; ++r) {
        if (r) o += ',';
        o += '{';
        for (size_t k = 0; k < cols.size(); ++k) {
          if (k) o += ',';
          o += "\"" + names[k] + "\":" + field(k, r, true);
        }
        o += '}';
      }
      o += ']';
    } else {
      for (uint64_t r = 0; r < in.n; ++r) {
        for (size_t k = 0; k < cols.size(); ++k) {
          if (k) o += ',';
          o += field(k, r, false);
        }
        o += '\n';
      }
    }
    out_ += o;
    printed_ = true;
  }
};

}  // namespace engine
}  // namespace pj

using namespace pj::engine;

extern "C" int pj_plan_execute(pj_ctx *ctx, const char *plan_json, const char *base_dir, char *out,
                               uint64_t out_cap) {
  if (!ctx || !plan_json) return pj_fail(PJ_E_ARG, "pj_plan_execute: null argument");
  if (out && out_cap) out[0] = 0;
  Json plan;
  std::string err;
  if (!json_parse(plan_json, &plan, &err)) return pj_fail(PJ_E_PARSE, "plan JSON: %s", err.c_str());
  std::string text;
  try {
    Exec ex(ctx, base_dir ? base_dir : "");
    text = ex.run(plan);
  } catch (const EngineError &e) {
    return pj_fail(e.code, "%s", e.msg.c_str());
  } catch (const std::exception &e) {
    return pj_fail(PJ_E_PARSE, "plan: %s", e.what());
  }
  if (!out || text.size() + 1 > out_cap) {
    if (out && out_cap) {
      memcpy(out, text.data(), out_cap - 1);
      out[out_cap - 1] = 0;
    }
    return pj_fail(PJ_E_CAPACITY, "plan output needs %llu bytes", (unsigned long long)text.size() + 1);
  }
  memcpy(out, text.data(), text.size());
  out[text.size()] = 0;
  return PJ_OK;
}
