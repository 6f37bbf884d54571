// Device building blocks of the plan executor: an expression VM over
// columns, ordered compaction, gathers, typed reductions and a general hash
// group-by.  Plain C++ types only, so the host engine (built with g++) can
// call them; the kernels live in rel.hip.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/proteus_amd.h"

namespace pj {
namespace rel {

// Column types.  The plan language's "int" is int32, "int64"/"date" int64,
// "float" FP64 (core/olap/lib/values/expressionTypes.cpp:87-97); bool is one
// byte; F32 is a storage type for registered float32 columns, widened to FP64
// when read.
enum ColT : int32_t { I32 = 0, I64 = 1, F64 = 2, BOOL = 3, F32 = 4 };
inline int col_size(int t) { return t == I32 || t == F32 ? 4 : t == BOOL ? 1 : 8; }

// Postfix program of one expression.  Every value on the stack is a 64-bit
// slot: integers and bools sign-extended, FP64 by bit pattern.
enum OpCode : int32_t {
  OP_COL,    // push column a (typed by ctype[a])
  OP_CONST,  // push imm
  OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD,  // t = operand/result type (I32 wraps)
  OP_NEG,
  OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE,  // t = operand type, result bool
  OP_AND, OP_OR, OP_NOT,
  OP_CAST,   // a = from type, t = to type
  OP_IF,     // pops else, then, cond
  OP_MIN, OP_MAX,  // t = operand type
  // pipeline programs only
  OP_PROBE,   // pops an int32 key, looks it up in table a; a miss ends the row,
              // a hit sets the row id of level a + 1
  OP_FILTER,  // pops a bool; false ends the row
  OP_OUT,     // pops into output column a
  // any program
  OP_BLOOM,   // pops an int32 key k, pushes bit (uint32)k & (2^a - 1) of the
              // filter words at address imm (a = log2 of the filter size)
};

struct Instr {
  int32_t op;
  int32_t t;
  int32_t a;
  int32_t pad;
  int64_t imm;
};

constexpr int kMaxInstr = 64;
constexpr int kMaxCols = 16;
constexpr int kMaxDepth = 16;

struct Prog {
  Instr ins[kMaxInstr];
  const void *cols[kMaxCols];
  int32_t ctype[kMaxCols];
  int32_t n = 0;      // instructions
  int32_t ncols = 0;
  int32_t depth = 0;  // stack slots needed
};

// out[i] = prog(row i), stored as out_t, i < n.
int eval(pj_ctx *ctx, const Prog &p, uint64_t n, void *out, int out_t);
// Row ids (ascending) of the rows whose predicate holds; *n_out set.
// out_idx must hold n entries.
int select(pj_ctx *ctx, const Prog &pred, uint64_t n, uint32_t *out_idx, uint64_t *n_out);
// dst[i] = src[idx[i]] for elements of esize bytes (1, 4 or 8).
int gather(pj_ctx *ctx, const void *src, int esize, const uint32_t *idx, uint64_t n, void *dst);
int iota(pj_ctx *ctx, uint32_t *d, uint64_t n);
// n device-to-host copies with one synchronisation (through pinned memory
// when they fit it together): dst[i] <- src[i], bytes[i] bytes.
int d2h_many(pj_ctx *ctx, int n, void *const *dst, const void *const *src, const uint64_t *bytes);

enum Monoid : int32_t { M_SUM = 0, M_MIN = 1, M_MAX = 2, M_AND = 3, M_OR = 4, M_CNT = 5 };
// Reduction of a column: integer types and bool accumulate in int64, F64/F32
// in FP64 (fixed order: per-workgroup partials, then one workgroup).
int reduce(pj_ctx *ctx, const void *col, int t, uint64_t n, int monoid, int64_t *out_i,
           double *out_f);
// MIN and MAX of n >= 1 rows of an integer column, with one read-back.
int reduce_minmax(pj_ctx *ctx, const void *col, int t, uint64_t n, int64_t *mn, int64_t *mx);

// General hash group-by over up to 4 key columns and up to 16 aggregates.
struct GbKey { const void *col; int32_t t; };
struct GbAgg { const void *col; int32_t t; int32_t monoid; };  // col may be null for M_CNT
constexpr int kGbMaxKeys = 4;
constexpr int kGbMaxAggs = 16;
struct GbOut {
  void *keys[kGbMaxKeys];  // device, capacity max_groups each, key types
  void *aggs[kGbMaxAggs];  // device, capacity max_groups: int64 (integer/bool/count) or FP64
};
int groupby(pj_ctx *ctx, const GbKey *keys, int nk, const GbAgg *aggs, int na, uint64_t n,
            uint64_t max_groups, const GbOut &out, uint64_t *n_groups);

// Fused probe pipeline.  Columns have a level: 0 = the source row,
// j + 1 = the build row matched by table j.  Tables are resident unique-key
// dimension tables whose value is the build row id.  Surviving rows are
// compacted (order unspecified) into the output columns, which must hold n
// rows.
struct PipeTable {
  const unsigned long long *table;  // 64-bit slots key << 32 | build row
  const uint32_t *ids;  // or (direct tables of dims_create_rows): the build row of
                        // key lo + i at i, 0xffffffff where no key
  uint64_t mask;   // hash tables: slots - 1
  int32_t lo;      // direct tables: the smallest key
  uint32_t range;  // direct tables: slots (keys lo .. lo + range - 1); 0 = hash table
};
constexpr int kPipeMaxLevels = 8;  // source + 7 joins
constexpr int kPipeMaxOut = 24;
struct PipeProg {
  std::vector<Instr> ins;
  std::vector<const void *> cols;
  std::vector<int32_t> ctype, clevel;
  std::vector<PipeTable> tables;
  std::vector<int32_t> out_t;
  int32_t depth = 0;
};
int pipeline(pj_ctx *ctx, const PipeProg &p, uint64_t n, void *const *outs, uint64_t *n_out);
// The table of a dimension built with dims_create_rows.
int dim_table(const pj_dim *dim, PipeTable *t);
// Tables of build rows for m build sides at once: side i's keys d_keys[i]
// (n[i] rows) with the build row ids 0 .. n[i]-1 as values; rows whose
// d_keep[i] byte is 0 (d_keep[i] may be null) are left out, a semijoin
// filter folded into the build (their keys are neither checked nor found).
// Direct tables hold 4-byte ids (half the bytes a probe's cache lines cover).
// Two synchronisations in all: one for every side's key range, one for
// every side's error flags.  On error no table is returned.
int dims_create_rows(pj_ctx *ctx, int m, const int32_t *const *d_keys,
                     const uint8_t *const *d_keep, const uint64_t *n, pj_dim **out);

// Stable sort permutation of n rows by nk keys (the first most
// significant): dir 1 ascending, -1 descending, 0 ignored; rank (int32 keys
// only, may be null) maps a dictionary code in [0, nrank) to its position in
// string order (codes outside sort after all others).  perm (n entries)
// receives the row ids in sorted order.  key0_out (may be null, n entries):
// when key 0 is a plain int32 key with a direction and the radix passes ran,
// its values in sorted order, and *key0_done is set (so no gather of that
// column through perm is needed).
struct SortKey {
  const void *col;
  int32_t t;
  int32_t dir;
  const int32_t *rank;
  uint32_t nrank;
};
int sort_perm(pj_ctx *ctx, const SortKey *keys, int nk, uint64_t n, uint32_t *perm,
              int32_t *key0_out = nullptr, bool *key0_done = nullptr);

// The words and log2 size of the context's bloom filter `id` (built with
// pj_bloom_build_i32), for OP_BLOOM.
int bloom_lookup(pj_ctx *ctx, uint64_t id, const uint32_t **words, int *log2_size);

// bucket[i] = hashInt32(k[i]) mod buckets (the reference's hash-rearrange
// target, gpu-hash-rearrange.cpp).
int hash_buckets(pj_ctx *ctx, const int32_t *k, uint64_t n, uint32_t buckets, uint32_t *out);
// hashInt32(k[i]) as int64
int hash_values(pj_ctx *ctx, const int32_t *k, uint64_t n, int64_t *out);

}  // namespace rel
}  // namespace pj
