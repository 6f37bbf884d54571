// Queries written with the C++ RelBuilder (include/proteus_amd/relbuilder.hpp)
// in the reference's call shapes (tests-join-sailors.cpp:100-166 and the
// sailors scan / select tests).
//
//   relbuilder_driver plan                 one line per query: name TAB plan JSON
//   relbuilder_driver run BASE DATA        runs them on device 0: name TAB rows,
//                                          rows separated by '|'
//
// BASE holds inputs/{sailors,reserves,boats}.csv;
// DATA holds R.k, R.v, S.k, S.v (raw int32) for the registered-column join.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "proteus_amd/relbuilder.hpp"

using namespace proteus_amd;

namespace {
// schemas in C++, as the reference's test declares them
const auto sailors_csv =
    rel("inputs/sailors.csv")(Int("sid"), String("sname"), Int("rating"), Float("age"));
const auto boats_csv = rel("inputs/boats.csv")(Int("bid"), String("bname"), String("color"));
const auto reserves_csv = rel("inputs/reserves.csv")(Int("sid"), Int("bid"), String("day"));

struct Query {
  std::string name;
  std::function<RelBuilder(const RelBuilder &)> build;
};

std::vector<Query> queries() {
  std::vector<Query> q;
  // SailorsTest.JoinLeft3
  q.push_back({"join_left3", [](const RelBuilder &f) {
                 auto sailors_scan = f.scan(sailors_csv, {"sid", "age"}, "pm-csv");
                 auto reserves_scan = f.scan(reserves_csv, {"sid", "bid"}, "pm-csv");
                 auto boats_scan = f.scan(boats_csv, {"bid"}, "pm-csv");
                 return sailors_scan
                     .join(
                         reserves_scan, [&](const auto &build_arg) -> Expr { return build_arg["sid"]; },
                         [&](const auto &probe_arg) -> Expr { return probe_arg["sid"]; }, 16, 128)
                     .join(
                         boats_scan, [&](const auto &build_arg) -> Expr { return build_arg["bid"]; },
                         [&](const auto &probe_arg) -> Expr { return probe_arg["bid"]; }, 16, 128)
                     .reduce(
                         [&](const auto &arg) -> std::vector<Expr> {
                           return {arg["sid"].as("tmp4", "max_sid"), arg["sid"].as("tmp4", "sum_sid"),
                                   Expr(1).as("tmp4", "count_sid")};
                         },
                         {MAX, SUM, SUM})
                     .print("pm-csv");
               }});
  // the same with the tables sized from row-count hints
  q.push_back({"join_left3_hinted", [](const RelBuilder &f) {
                 auto reserves_scan = f.scan(reserves_csv, {"sid", "bid"}, "pm-csv").hintRowCount(10);
                 auto boats_scan = f.scan(boats_csv, {"bid"}, "pm-csv").hintRowCount(4);
                 return f.scan(sailors_csv, {"sid", "age"}, "pm-csv")
                     .to_gpu()
                     .unpack()
                     .join(reserves_scan, [](const auto &b) { return b["sid"]; },
                           [](const auto &p) { return p["sid"]; })
                     .join(boats_scan, [](const auto &b) { return b["bid"]; },
                           [](const auto &p) { return p["bid"]; })
                     .pack()
                     .to_cpu()
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {arg["sid"].as("t", "max_sid"), arg["sid"].as("t", "sum_sid"),
                                   Expr(1).as("t", "cnt")};
                         },
                         {MAX, SUM, SUM})
                     .print("pm-csv");
               }});
  // SailorsTest.Scan / Select / ScanBoats
  q.push_back({"scan", [](const RelBuilder &f) {
                 return f.scan(sailors_csv, {"sid"}, "pm-csv")
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {arg["sid"].as("t", "max_sid"), arg["sid"].as("t", "sum_sid"),
                                   Expr(1).as("t", "cnt")};
                         },
                         {MAX, SUM, SUM})
                     .print("pm-csv");
               }});
  q.push_back({"select", [](const RelBuilder &f) {
                 return f.scan(sailors_csv, {"sid", "age"}, "pm-csv")
                     .filter([](const auto &arg) { return gt(arg["age"], 40.0); })
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {arg["sid"].as("t", "max_sid"), arg["sid"].as("t", "sum_sid"),
                                   Expr(1).as("t", "cnt")};
                         },
                         {MAX, SUM, SUM})
                     .print("pm-csv");
               }});
  q.push_back({"scan_boats", [](const RelBuilder &f) {
                 return f.scan(boats_csv, {"bid"}, "pm-csv")
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {arg["bid"].as("t", "max_bid"), arg["bid"].as("t", "sum_bid"),
                                   Expr(1).as("t", "cnt")};
                         },
                         {MAX, SUM, SUM})
                     .print("pm-csv");
               }});
  // group-by rating: count, sum of sid, max age; sorted by rating
  q.push_back({"groupby_rating", [](const RelBuilder &f) {
                 return f.scan(sailors_csv, {"sid", "rating", "age"}, "pm-csv")
                     .groupby([](const auto &arg) -> std::vector<Expr> { return {arg["rating"]}; },
                              [](const auto &arg) -> std::vector<GpuAggrMatExpr> {
                                return {GpuAggrMatExpr(Expr(1).as("g", "cnt"), 1, 0, SUM),
                                        GpuAggrMatExpr(arg["sid"].as("g", "sum_sid"), 2, 0, SUM),
                                        GpuAggrMatExpr(arg["age"].as("g", "max_age"), 3, 0, MAX)};
                              },
                              10, 128)
                     .sort([](const auto &arg) -> std::vector<Expr> {
                             return {arg["rating"], arg["cnt"], arg["sum_sid"], arg["max_age"]};
                           },
                           {direction::ASC, direction::NONE, direction::NONE, direction::NONE})
                     .print("pm-csv");
               }});
  // project then sort by age descending, sid ascending
  q.push_back({"project_sort", [](const RelBuilder &f) {
                 return f.scan(sailors_csv, {"sid", "age", "rating"}, "pm-csv")
                     .project([](const auto &arg) -> std::vector<Expr> {
                       return {arg["sid"].as("p", "sid"), (arg["rating"] * Expr(2)).as("p", "r2"),
                               arg["age"].as("p", "age")};
                     })
                     .sort([](const auto &arg) -> std::vector<Expr> {
                             return {arg["age"], arg["sid"], arg["r2"]};
                           },
                           {direction::DESC, direction::ASC, direction::NONE})
                     .print("pm-csv");
               }});
  // boats through a bloom filter into a join with reserves
  q.push_back({"bloom_join", [](const RelBuilder &f) {
                 auto boats = f.scan(boats_csv, {"bid"}, "pm-csv")
                                  .filter([](const auto &a) { return lt(a["bid"], 103); })
                                  .bloomfilter_build([](const auto &a) { return a["bid"]; }, 1024, 3);
                 return f.scan(reserves_csv, {"sid", "bid"}, "pm-csv")
                     .bloomfilter_probe([](const auto &a) { return a["bid"]; }, 1024, 3)
                     .join(boats, [](const auto &b) { return b["bid"]; },
                           [](const auto &p) { return p["bid"]; }, 10, 16)
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {Expr(1).as("t", "cnt"), arg["sid"].as("t", "sum_sid")};
                         },
                         {SUM, SUM})
                     .print("pm-csv");
               }});
  // columns registered with pj_register_column (plugin "block")
  q.push_back({"registered_join", [](const RelBuilder &f) {
                 auto R = f.scan("R", {"k", "v"}).hintRowCount(1 << 16);
                 return f.scan("S", {"k", "v"})
                     .router(DegreeOfParallelism(1), 8, RoutingPolicy::LOCAL, DeviceType::GPU)
                     .memmove(8, DeviceType::GPU)
                     .to_gpu()
                     .unpack()
                     .join(R, [](const auto &b) { return b["k"]; }, [](const auto &p) { return p["k"]; })
                     .reduce(
                         [](const auto &arg) -> std::vector<Expr> {
                           return {Expr(int64_t(1)).as("t", "cnt"),
                                   (cast(arg("R", "v"), "int64") + cast(arg("S", "v"), "int64"))
                                       .as("t", "sum_v"),
                                   arg("R", "v").as("t", "max_rv")};
                         },
                         {SUM, SUM, MAX})
                     .print("pm-csv");
               }});
  return q;
}

std::vector<int32_t> read_i32(const std::string &path) {
  std::ifstream is(path, std::ios::binary);
  if (!is) throw std::runtime_error("cannot open " + path);
  std::vector<char> b((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
  std::vector<int32_t> v(b.size() / 4);
  std::memcpy(v.data(), b.data(), v.size() * 4);
  return v;
}

std::string one_line(std::string s) {
  while (!s.empty() && s.back() == '\n') s.pop_back();
  for (char &c : s)
    if (c == '\n') c = '|';
  return s;
}
}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s plan | run BASE DATA\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  try {
    if (mode == "plan") {
      RelBuilder f;
      for (auto &q : queries()) std::cout << q.name << '\t' << q.build(f).plan() << '\n';
      // errors the builder raises before any plan runs
      try {
        f.scan(sailors_csv, {"sid"}, "pm-csv").filter([](const auto &a) { return a["nope"]; });
        std::cout << "no_error\n";
      } catch (const std::runtime_error &e) {
        std::cout << "error\t" << e.what() << '\n';
      }
      try {
        auto b = f.scan(boats_csv, {"bid"}, "pm-csv");
        f.scan(reserves_csv, {"bid"}, "pm-csv")
            .join(b, [](const auto &x) { return x["bid"]; }, [](const auto &x) { return x["bid"]; });
        std::cout << "no_error\n";
      } catch (const std::runtime_error &e) {
        std::cout << "error\t" << e.what() << '\n';
      }
      return 0;
    }
    if (mode != "run" || argc < 4) {
      std::fprintf(stderr, "usage: %s plan | run BASE DATA\n", argv[0]);
      return 2;
    }
    const std::string data = argv[3];
    pj_ctx *ctx = nullptr;
    if (pj_ctx_create(0, 0, &ctx) != PJ_OK) throw std::runtime_error(pj_last_error());
    std::vector<int32_t> rk = read_i32(data + "/R.k"), rv = read_i32(data + "/R.v");
    std::vector<int32_t> sk = read_i32(data + "/S.k"), sv = read_i32(data + "/S.v");
    struct {
      const char *rel, *attr;
      std::vector<int32_t> *v;
    } cols[] = {{"R", "k", &rk}, {"R", "v", &rv}, {"S", "k", &sk}, {"S", "v", &sv}};
    for (auto &c : cols)
      if (pj_register_column(ctx, c.rel, c.attr, "int", c.v->data(), c.v->size(), 0) != PJ_OK)
        throw std::runtime_error(pj_last_error());
    RelBuilder f(ctx, argv[2]);
    for (auto &q : queries()) {
      std::cout << q.name << '\t' << std::flush;
      std::cout << one_line(q.build(f).prepare().execute().str()) << '\n' << std::flush;
    }
    pj_clear_catalog(ctx);
    pj_ctx_destroy(ctx);
  } catch (const std::exception &e) {
    std::cout << std::flush;
    std::fprintf(stderr, "relbuilder_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
