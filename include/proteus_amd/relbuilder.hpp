/*
 * proteus_amd/relbuilder.hpp -- C++ RelBuilder over the C ABI.
 *
 * The call shapes of the reference's RelBuilder
 * (core/olap/include/olap/operators/relbuilder.hpp:110-609,
 * core/olap/lib/operators/relbuilder.cpp:304-349,432-443,653-755): scan,
 * to_gpu / to_cpu, unpack / pack, filter, project, join, groupby, reduce,
 * sort, print, hintRowCount, prepare() -> PreparedStatement::execute() ->
 * QueryResult.  Operators take lambdas over an input argument whose
 * operator[] names an attribute of the operator's input, as in the
 * reference (`[](const auto &arg) { return arg["sid"]; }`).  Instead of
 * generating code, prepare() renders the plan in the reference's JSON plan
 * language (core/olap/lib/plan/plan-parser.cpp) and execute() runs it through
 * pj_plan_execute.  Errors are thrown as std::runtime_error, as the
 * reference does.
 *
 * Header-only; needs C++17 and proteus_amd.h.
 */
#ifndef PROTEUS_AMD_RELBUILDER_HPP
#define PROTEUS_AMD_RELBUILDER_HPP

#include <cmath>
#include <cstdint>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../proteus_amd.h"

namespace proteus_amd {

namespace detail {
inline std::string quote(const std::string &s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}
// the plan language's name of a RelBuilder plugin (plan-parser.cpp:1679-1765)
inline std::string plan_plugin(const std::string &pg) { return pg == "pm-csv" ? "csv" : pg; }
inline std::string attr_json(const std::string &rel, const std::string &attr) {
  return "{\"relName\":" + quote(rel) + ",\"attrName\":" + quote(attr) + "}";
}
}  // namespace detail

// ---- expressions (core/olap/lib/plan/expression-parser.cpp) --------------
// An expression of the plan language, kept as its JSON text.  as(rel, attr)
// registers its value as rel.attr (the reference's expression_t::as).
class Expr {
 public:
  Expr() = default;
  explicit Expr(std::string json) : j_(std::move(json)) {}
  Expr(int32_t v) : j_("{\"expression\":\"int\",\"v\":" + std::to_string(v) + "}") {}
  Expr(int64_t v) : j_("{\"expression\":\"int64\",\"v\":" + std::to_string(v) + "}") {}
  Expr(double v) {
    std::ostringstream s;
    s.precision(17);
    s << v;
    j_ = "{\"expression\":\"float\",\"v\":" + s.str() + "}";
  }
  Expr(bool v) : j_(std::string("{\"expression\":\"bool\",\"v\":") + (v ? "true" : "false") + "}") {}

  Expr as(const std::string &rel, const std::string &attr) const {
    Expr e = *this;
    e.j_.pop_back();  // the closing brace
    e.j_ += ",\"register_as\":" + detail::attr_json(rel, attr) + "}";
    e.reg_rel_ = rel;
    e.reg_attr_ = attr;
    return e;
  }
  const std::string &json() const { return j_; }
  // the attribute this expression is registered as, or projects
  const std::string &rel() const { return reg_rel_.empty() ? col_rel_ : reg_rel_; }
  const std::string &attr() const { return reg_attr_.empty() ? col_attr_ : reg_attr_; }

  static Expr column(const std::string &rel, const std::string &attr) {
    Expr e("{\"expression\":\"recordProjection\",\"e\":{\"expression\":\"argument\"},"
           "\"attribute\":" + detail::attr_json(rel, attr) + "}");
    e.col_rel_ = rel;
    e.col_attr_ = attr;
    return e;
  }
  static Expr binary(const char *op, const Expr &l, const Expr &r) {
    return Expr(std::string("{\"expression\":\"") + op + "\",\"left\":" + l.j_ +
                ",\"right\":" + r.j_ + "}");
  }

 private:
  std::string j_;
  std::string reg_rel_, reg_attr_, col_rel_, col_attr_;
};

inline Expr operator+(const Expr &a, const Expr &b) { return Expr::binary("add", a, b); }
inline Expr operator-(const Expr &a, const Expr &b) { return Expr::binary("sub", a, b); }
inline Expr operator*(const Expr &a, const Expr &b) { return Expr::binary("multiply", a, b); }
inline Expr operator/(const Expr &a, const Expr &b) { return Expr::binary("div", a, b); }
inline Expr operator%(const Expr &a, const Expr &b) { return Expr::binary("mod", a, b); }
inline Expr eq(const Expr &a, const Expr &b) { return Expr::binary("eq", a, b); }
inline Expr ne(const Expr &a, const Expr &b) { return Expr::binary("neq", a, b); }
inline Expr lt(const Expr &a, const Expr &b) { return Expr::binary("lt", a, b); }
inline Expr le(const Expr &a, const Expr &b) { return Expr::binary("le", a, b); }
inline Expr gt(const Expr &a, const Expr &b) { return Expr::binary("gt", a, b); }
inline Expr ge(const Expr &a, const Expr &b) { return Expr::binary("ge", a, b); }
inline Expr operator&&(const Expr &a, const Expr &b) { return Expr::binary("and", a, b); }
inline Expr operator||(const Expr &a, const Expr &b) { return Expr::binary("or", a, b); }
inline Expr operator-(const Expr &a) { return Expr("{\"expression\":\"neg\",\"e\":" + a.json() + "}"); }
inline Expr cast(const Expr &e, const std::string &type) {
  return Expr("{\"expression\":\"cast\",\"type\":{\"type\":" + detail::quote(type) + "},\"e\":" +
              e.json() + "}");
}
inline Expr cond(const Expr &c, const Expr &t, const Expr &e) {
  return Expr("{\"expression\":\"if\",\"cond\":" + c.json() + ",\"then\":" + t.json() +
              ",\"else\":" + e.json() + "}");
}

// Aggregate monoids (the reference's Monoid, core/olap/include/olap/expressions/
// expressions.hpp), sort directions and the routing vocabulary.
enum Monoid { SUM, MAX, MIN, AND, OR };
enum class direction { ASC, DESC, NONE };
enum class DeviceType { CPU, GPU };
enum class RoutingPolicy { RANDOM, LOCAL, HASH_BASED };
struct DegreeOfParallelism {
  size_t dop;
  explicit DegreeOfParallelism(size_t d) : dop(d) {}
};

inline const char *monoid_name(Monoid m) {
  switch (m) {
    case SUM: return "sum";
    case MAX: return "max";
    case MIN: return "min";
    case AND: return "and";
    default: return "or";
  }
}

// One aggregate of a group-by: expression, packet and offset of the
// reference's GpuAggrMatExpr (gpu-aggr-mat-expr.hpp), and its monoid.
struct GpuAggrMatExpr {
  Expr e;
  size_t packet;
  size_t offset;
  Monoid m;
  GpuAggrMatExpr(Expr e_, size_t packet_, size_t offset_, Monoid m_)
      : e(std::move(e_)), packet(packet_), offset(offset_), m(m_) {}
};

// ---- schemas of CSV relations (the reference's rel(...)(Int(...), ...)) ---
struct AttrSpec {
  std::string name, type;
};
inline AttrSpec Int(const std::string &n) { return {n, "int"}; }
inline AttrSpec Int64(const std::string &n) { return {n, "int64"}; }
inline AttrSpec Float(const std::string &n) { return {n, "float"}; }
inline AttrSpec String(const std::string &n) { return {n, "string"}; }

struct Relation {
  std::string name;
  std::vector<AttrSpec> attrs;
  template <class... A>
  Relation operator()(A... a) const {
    Relation r{name, {a...}};
    return r;
  }
};
inline Relation rel(const std::string &name) { return Relation{name, {}}; }

// The attributes an operator produces; operator[] is the reference's
// InputArgument lookup (expressions.hpp:1469-1530): the first attribute of
// that name, or "$i" for the i-th attribute.
class Arg {
 public:
  std::vector<std::pair<std::string, std::string>> attrs;  // (rel, attr)
  Expr operator[](const std::string &name) const {
    for (auto &a : attrs)
      if (a.second == name) return Expr::column(a.first, a.second);
    if (name.size() > 1 && name[0] == '$') {
      size_t i = std::stoul(name.substr(1));
      if (i < attrs.size()) return Expr::column(attrs[i].first, attrs[i].second);
    }
    throw std::runtime_error("Invalid record projection: " + name);
  }
  // a projection (rel, name) of the input
  Expr operator()(const std::string &rel, const std::string &name) const {
    return Expr::column(rel, name);
  }
};

class QueryResult {
 public:
  explicit QueryResult(std::string text) : text_(std::move(text)) {}
  const std::string &str() const { return text_; }
  std::vector<std::vector<std::string>> rows() const {
    std::vector<std::vector<std::string>> out;
    std::istringstream in(text_);
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty()) continue;
      std::vector<std::string> f;
      std::string cell;
      std::istringstream ls(line);
      while (std::getline(ls, cell, ',')) f.push_back(cell);
      out.push_back(f);
    }
    return out;
  }

 private:
  std::string text_;
};

class PreparedStatement {
 public:
  PreparedStatement(pj_ctx *ctx, std::string plan, std::string base_dir)
      : ctx_(ctx), plan_(std::move(plan)), base_(std::move(base_dir)) {}
  const std::string &plan() const { return plan_; }
  QueryResult execute() const {
    if (!ctx_) throw std::runtime_error("PreparedStatement without a pj_ctx");
    std::vector<char> buf(1 << 20);
    for (;;) {
      int rc = pj_plan_execute(ctx_, plan_.c_str(), base_.c_str(), buf.data(), buf.size());
      if (rc == PJ_OK) return QueryResult(std::string(buf.data()));
      std::string msg = pj_last_error();
      const std::string need = "plan output needs ";
      if (rc == PJ_E_CAPACITY && msg.compare(0, need.size(), need) == 0) {
        buf.resize(std::stoull(msg.substr(need.size())) + 1);
        continue;
      }
      throw std::runtime_error("pj_plan_execute: " + msg);
    }
  }

 private:
  pj_ctx *ctx_;
  std::string plan_, base_;
};

class RelBuilder {
 public:
  explicit RelBuilder(pj_ctx *ctx = nullptr, std::string base_dir = ".")
      : ctx_(ctx), base_(std::move(base_dir)) {}

  // ---- sources ----------------------------------------------------------
  // Relation `rel` whose columns were registered with pj_register_column
  // (any plugin type; "block" by default), or a CSV file with a Relation
  // schema (plugin "pm-csv" / "csv", fields separated by `delimiter`).
  RelBuilder scan(const std::string &rel, const std::vector<std::string> &attrs,
                  const std::string &plugin = "block") const {
    std::string pr;
    for (auto &a : attrs) pr += (pr.empty() ? "" : ",") + detail::attr_json(rel, a);
    RelBuilder r = wrap("{\"operator\":\"scan\",\"plugin\":{\"name\":" + detail::quote(rel) +
                        ",\"type\":" + detail::quote(detail::plan_plugin(plugin)) + ",\"projections\":[" + pr + "]}}");
    for (auto &a : attrs) r.out_.attrs.push_back({rel, a});
    return r;
  }
  RelBuilder scan(const Relation &schema, const std::vector<std::string> &attrs,
                  const std::string &plugin = "pm-csv", char delimiter = ';') const {
    std::string pr;
    RelBuilder r;
    for (auto &a : attrs) {
      size_t no = 0;
      std::string type;
      for (size_t i = 0; i < schema.attrs.size(); ++i)
        if (schema.attrs[i].name == a) {
          no = i + 1;
          type = schema.attrs[i].type;
        }
      if (!no) throw std::runtime_error("no attribute " + a + " in " + schema.name);
      pr += std::string(pr.empty() ? "" : ",") + "{\"type\":{\"type\":" + detail::quote(type) +
            "},\"relName\":" + detail::quote(schema.name) + ",\"attrName\":" + detail::quote(a) +
            ",\"attrNo\":" + std::to_string(no) + "}";
    }
    r = wrap("{\"operator\":\"scan\",\"plugin\":{\"name\":" + detail::quote(schema.name) +
             ",\"type\":" + detail::quote(detail::plan_plugin(plugin)) +
             ",\"delimiter\":" + detail::quote(std::string(1, delimiter)) +
             ",\"brackets\":false,\"projections\":[" + pr + "]}}");
    for (auto &a : attrs) r.out_.attrs.push_back({schema.name, a});
    return r;
  }

  // The expected output rows of this operator (used by join() without an
  // explicit table size, as the reference's hintRowCount).
  RelBuilder hintRowCount(double rows) const {
    RelBuilder r = *this;
    r.rows_ = rows;
    return r;
  }

  // ---- placement and routing: with every column in HBM of the one device
  // these pass their input through (the plan keeps the operator) --------
  RelBuilder to_gpu() const { return on("cpu-to-gpu", ""); }
  RelBuilder to_cpu() const { return on("gpu-to-cpu", ""); }
  RelBuilder unpack() const { return on("unpack", ""); }
  RelBuilder pack() const { return on("pack", ""); }
  RelBuilder memmove(size_t slack, DeviceType to) const {
    return on("mem-move-device", ",\"slack\":" + std::to_string(slack) + ",\"to_cpu\":" +
                                     (to == DeviceType::CPU ? "true" : "false"));
  }
  RelBuilder membrdcst(DegreeOfParallelism fanout, bool to_cpu, bool always_share = false) const {
    return on("mem-broadcast-device", ",\"num_of_targets\":" + std::to_string(fanout.dop) +
                                          ",\"to_cpu\":" + (to_cpu ? "true" : "false") +
                                          ",\"always_share\":" + (always_share ? "true" : "false"));
  }
  RelBuilder router(DegreeOfParallelism fanout, size_t slack, RoutingPolicy p,
                    DeviceType target) const {
    const char *pol = p == RoutingPolicy::RANDOM ? "RANDOM" : p == RoutingPolicy::LOCAL ? "LOCAL"
                                                                                       : "HASH_BASED";
    return on("router", ",\"numOfParents\":" + std::to_string(fanout.dop) + ",\"slack\":" +
                            std::to_string(slack) + ",\"policy\":\"" + pol + "\",\"target\":\"" +
                            (target == DeviceType::CPU ? "CPU" : "GPU") + "\"");
  }
  RelBuilder split(size_t alternatives, size_t slack, RoutingPolicy p) const {
    (void)p;
    return on("split", ",\"numOfParents\":" + std::to_string(alternatives) + ",\"slack\":" +
                           std::to_string(slack));
  }
  // hash-rearrange of the tuples into numOfBuckets buckets by hashExpr
  template <class T, class Th>
  RelBuilder pack(T expr, Th hashExpr, size_t numOfBuckets) const {
    std::vector<Expr> es = expr(out_);
    return on("hash-rearrange", ",\"projections\":" + list(es) + ",\"e\":" + hashExpr(out_).json() +
                                    ",\"buckets\":" + std::to_string(numOfBuckets));
  }

  // ---- relational operators ---------------------------------------------
  template <class T>
  RelBuilder filter(T pred) const {
    return on("select", ",\"p\":" + pred(out_).json());
  }
  // bloom filters (relbuilder.hpp:427-462): pred is the int32 key expression;
  // build sets the key's entry of filter bloomId and passes every row on,
  // probe keeps the rows whose entry is set, repack keeps them and
  // compacts every input attribute
  template <class T>
  RelBuilder bloomfilter_build(T pred, size_t filterSize, uint64_t bloomId) const {
    return on("bloomfilter-build", bloom_fields(pred(out_), filterSize, bloomId));
  }
  template <class T>
  RelBuilder bloomfilter_probe(T pred, size_t filterSize, uint64_t bloomId) const {
    return on("bloomfilter-probe", bloom_fields(pred(out_), filterSize, bloomId));
  }
  template <class T>
  RelBuilder bloomfilter_repack(T pred, size_t filterSize, uint64_t bloomId) const {
    std::vector<Expr> attrs;
    for (auto &a : out_.attrs) attrs.push_back(Expr::column(a.first, a.second));
    return on("bloomfilter-repack",
              bloom_fields(pred(out_), filterSize, bloomId) + ",\"projections\":" + list(attrs));
  }
  template <class T>
  RelBuilder project(T expr) const {
    std::vector<Expr> es = expr(out_);
    std::string relname = es.empty() ? "project" : es[0].rel();
    RelBuilder r = on("project", ",\"relName\":" + detail::quote(relname) + ",\"e\":" + list(es));
    r.out_.attrs.clear();
    for (auto &e : es) r.out_.attrs.push_back({e.rel(), e.attr()});
    return r;
  }
  // hashjoin-chained; *this is the probe side
  template <class Tbk, class Tpk>
  RelBuilder join(const RelBuilder &build, Tbk build_k, Tpk probe_k, int hash_bits,
                  size_t maxBuildInputSize) const {
    Expr bk = build_k(build.out_), pk = probe_k(out_);
    RelBuilder r = wrap("{\"operator\":\"hashjoin-chained\",\"build_input\":" + build.node_ +
                        ",\"probe_input\":" + node_ + ",\"build_k\":" + bk.json() +
                        ",\"probe_k\":" + pk.json() + ",\"hash_bits\":" +
                        std::to_string(hash_bits) + ",\"maxBuildInputSize\":" +
                        std::to_string(maxBuildInputSize) + "}");
    r.out_.attrs = build.out_.attrs;
    r.out_.attrs.insert(r.out_.attrs.end(), out_.attrs.begin(), out_.attrs.end());
    return r;
  }
  // table sized from the build side's expected rows, as relbuilder.cpp:653-668:
  // hash_bits = ceil(log2(rows)) + 1, maxBuildInputSize = 1.1 * rows
  template <class Tbk, class Tpk>
  RelBuilder join(const RelBuilder &build, Tbk build_k, Tpk probe_k) const {
    if (build.rows_ <= 0)
      throw std::runtime_error("join without a table size needs the build side's row count "
                               "(hintRowCount)");
    return join(build, build_k, probe_k, (int)std::ceil(std::log2(build.rows_)) + 1,
                (size_t)(build.rows_ * 1.1));
  }
  // k: keys (vector<Expr>); e: aggregates (vector<GpuAggrMatExpr>)
  template <class Tk, class Te>
  RelBuilder groupby(Tk k, Te e, size_t hash_bits, size_t maxInputSize) const {
    std::vector<Expr> ks = k(out_);
    std::vector<GpuAggrMatExpr> es = e(out_);
    std::string agg;
    for (auto &a : es)
      agg += std::string(agg.empty() ? "" : ",") + "{\"e\":" + a.e.json() + ",\"m\":\"" +
             monoid_name(a.m) + "\",\"packet\":" + std::to_string(a.packet) +
             ",\"offset\":" + std::to_string(a.offset) + "}";
    RelBuilder r = on("groupby", ",\"k\":" + list(ks) + ",\"e\":[" + agg + "],\"hash_bits\":" +
                                     std::to_string(hash_bits) + ",\"maxInputSize\":" +
                                     std::to_string(maxInputSize));
    r.out_.attrs.clear();
    for (auto &x : ks) r.out_.attrs.push_back({x.rel(), x.attr()});
    for (auto &a : es) r.out_.attrs.push_back({a.e.rel(), a.e.attr()});
    return r;
  }
  // single-group aggregates; the attribute names are those the expressions
  // are registered as
  template <class T>
  RelBuilder reduce(T expr, const std::vector<Monoid> &accs) const {
    std::vector<Expr> es = expr(out_);
    if (es.size() != accs.size())
      throw std::runtime_error("reduce: one monoid per expression");
    std::string ms;
    for (Monoid m : accs) ms += std::string(ms.empty() ? "" : ",") + "\"" + monoid_name(m) + "\"";
    RelBuilder r = on("reduce", ",\"e\":" + list(es) + ",\"accumulator\":[" + ms +
                                    "],\"p\":{\"expression\":\"bool\",\"v\":true}");
    r.out_.attrs.clear();
    for (auto &x : es) r.out_.attrs.push_back({x.rel(), x.attr()});
    return r;
  }
  template <class T>
  RelBuilder sort(T e, const std::vector<direction> &dirs) const {
    std::vector<Expr> es = e(out_);
    if (es.size() != dirs.size()) throw std::runtime_error("sort: one direction per expression");
    std::string items;
    for (size_t i = 0; i < es.size(); ++i) {
      const char *d = dirs[i] == direction::ASC ? "ASC" : dirs[i] == direction::DESC ? "DESC" : "NONE";
      items += std::string(items.empty() ? "" : ",") + "{\"expression\":" + es[i].json() +
               ",\"direction\":\"" + d + "\"}";
    }
    return on("sort", ",\"e\":[" + items + "]");
  }
  // print the output rows ("pm-csv": comma-separated, "json": array of objects)
  RelBuilder print(const std::string &format = "pm-csv") const {
    return on("print", ",\"plugin\":{\"type\":" + detail::quote(format) + "}");
  }

  // ---- compile ----------------------------------------------------------
  PreparedStatement prepare() const { return PreparedStatement(ctx_, node_, base_); }
  const std::string &plan() const { return node_; }
  const Arg &getOutputArg() const { return out_; }

 private:
  RelBuilder wrap(std::string node) const {
    RelBuilder r(ctx_, base_);
    r.node_ = std::move(node);
    r.rows_ = rows_;
    return r;
  }
  RelBuilder on(const char *op, const std::string &fields) const {
    RelBuilder r = wrap(std::string("{\"operator\":\"") + op + "\"" + fields + ",\"input\":" + node_ + "}");
    r.out_ = out_;
    return r;
  }
  static std::string bloom_fields(const Expr &e, size_t size, uint64_t id) {
    return ",\"e\":" + e.json() + ",\"filterSize\":" + std::to_string(size) +
           ",\"bloomId\":" + std::to_string(id);
  }
  static std::string list(const std::vector<Expr> &es) {
    std::string o = "[";
    for (size_t i = 0; i < es.size(); ++i) o += (i ? "," : "") + es[i].json();
    return o + "]";
  }

  pj_ctx *ctx_;
  std::string base_;
  std::string node_ = "null";
  Arg out_;
  double rows_ = 0;
};

}  // namespace proteus_amd

#endif
