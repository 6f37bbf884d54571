// Device building blocks of the plan executor (declared in rel_api.hpp).
//
// Expressions run on a small stack VM: the program (postfix, at most 64
// instructions over at most 16 columns) travels in the kernel arguments, so
// every lane of a wave executes the same instruction and dispatch is a scalar
// branch; each lane keeps its stack in LDS (slot d of lane t at st[d][t],
// conflict-free).  This plays the role of the reference's expression code
// generation (core/olap/lib/expressions/expressions-generator.cpp) without a
// JIT.
#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>

#include "pj_internal.hpp"
#include "rel_api.hpp"

namespace pj {
namespace rel {

__device__ __forceinline__ int64_t ld_slot(const void *c, int t, uint64_t i) {
  switch (t) {
    case I32: return ((const int32_t *)c)[i];
    case I64: return ((const int64_t *)c)[i];
    case F64: return ((const int64_t *)c)[i];
    case BOOL: return ((const uint8_t *)c)[i] != 0;
    default: return __builtin_bit_cast(int64_t, (double)((const float *)c)[i]);
  }
}

__device__ __forceinline__ void st_slot(void *o, int t, uint64_t i, int64_t v) {
  switch (t) {
    case I32: ((int32_t *)o)[i] = (int32_t)v; break;
    case I64:
    case F64: ((int64_t *)o)[i] = v; break;
    case BOOL: ((uint8_t *)o)[i] = v != 0; break;
    default: ((float *)o)[i] = (float)__builtin_bit_cast(double, v); break;
  }
}

__device__ __forceinline__ double as_f(int64_t v) { return __builtin_bit_cast(double, v); }
__device__ __forceinline__ int64_t of_f(double v) { return __builtin_bit_cast(int64_t, v); }
__device__ __forceinline__ int64_t wrap32(int64_t v) { return (int64_t)(int32_t)(uint32_t)v; }

// Integer arithmetic in two's complement (wrapping as LLVM's add/sub/mul on
// i32/i64); division by zero yields 0 and INT_MIN / -1 wraps.
__device__ __forceinline__ int64_t arith(int op, int t, int64_t a, int64_t b) {
  if (t == F64) {
    double x = as_f(a), y = as_f(b), r;
    switch (op) {
      case OP_ADD: r = x + y; break;
      case OP_SUB: r = x - y; break;
      case OP_MUL: r = x * y; break;
      case OP_DIV: r = x / y; break;
      default: r = fmod(x, y); break;
    }
    return of_f(r);
  }
  uint64_t x = (uint64_t)a, y = (uint64_t)b, r;
  switch (op) {
    case OP_ADD: r = x + y; break;
    case OP_SUB: r = x - y; break;
    case OP_MUL: r = x * y; break;
    case OP_DIV:
      if (b == 0) r = 0;
      else if (b == -1) r = 0 - x;
      else r = (uint64_t)(a / b);
      break;
    default:
      if (b == 0 || b == -1) r = 0;
      else r = (uint64_t)(a % b);
      break;
  }
  return t == I32 ? wrap32((int64_t)r) : (int64_t)r;
}

__device__ __forceinline__ bool cmp(int op, int t, int64_t a, int64_t b) {
  if (t == F64) {
    double x = as_f(a), y = as_f(b);
    switch (op) {
      case OP_EQ: return x == y;
      case OP_NE: return x != y;
      case OP_LT: return x < y;
      case OP_LE: return x <= y;
      case OP_GT: return x > y;
      default: return x >= y;
    }
  }
  switch (op) {
    case OP_EQ: return a == b;
    case OP_NE: return a != b;
    case OP_LT: return a < b;
    case OP_LE: return a <= b;
    case OP_GT: return a > b;
    default: return a >= b;
  }
}

__device__ __forceinline__ int64_t cast(int from, int to, int64_t v) {
  if (from == F64 || from == F32) {
    double x = as_f(v);
    if (to == F64 || to == F32) return v;
    if (to == BOOL) return x != 0.0;
    // fptosi: truncation toward zero (out-of-range values saturate here)
    if (x != x) return 0;
    if (x >= 9.2233720368547758e18) return INT64_MAX;
    if (x <= -9.2233720368547758e18) return INT64_MIN;
    int64_t r = (int64_t)x;
    return to == I32 ? wrap32(r) : r;
  }
  switch (to) {
    case F64:
    case F32: return of_f((double)v);
    case BOOL: return v != 0;
    case I32: return wrap32(v);
    default: return v;
  }
}

// OP_BLOOM: bit (uint32)key & (2^log2 - 1) of a bit-packed filter.
__device__ __forceinline__ int64_t bloom_bit(int64_t words, int log2, int64_t key) {
  const uint32_t mask = log2 >= 32 ? 0xffffffffu : (1u << log2) - 1u;
  const uint32_t h = (uint32_t)key & mask;
  return (((const uint32_t *)words)[h >> 5] >> (h & 31)) & 1u;
}

// One row of the program; st points at this lane's slot 0, stride BLOCK.
template <int BLOCK>
__device__ __forceinline__ int64_t run_prog(const Prog &p, uint64_t row, int64_t *st) {
  int sp = 0;
  for (int k = 0; k < p.n; ++k) {
    const Instr in = p.ins[k];
    switch (in.op) {
      case OP_COL: st[sp++ * BLOCK] = ld_slot(p.cols[in.a], p.ctype[in.a], row); break;
      case OP_CONST: st[sp++ * BLOCK] = in.imm; break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MOD: {
        int64_t b = st[--sp * BLOCK], a = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] = arith(in.op, in.t, a, b);
        break;
      }
      case OP_NEG: {
        int64_t a = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] =
            in.t == F64 ? of_f(-as_f(a)) : (in.t == I32 ? wrap32(0 - a) : (int64_t)(0 - (uint64_t)a));
        break;
      }
      case OP_EQ: case OP_NE: case OP_LT: case OP_LE: case OP_GT: case OP_GE: {
        int64_t b = st[--sp * BLOCK], a = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] = cmp(in.op, in.t, a, b);
        break;
      }
      case OP_AND: {
        int64_t b = st[--sp * BLOCK], a = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] = (a != 0) & (b != 0);
        break;
      }
      case OP_OR: {
        int64_t b = st[--sp * BLOCK], a = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] = (a != 0) | (b != 0);
        break;
      }
      case OP_NOT: st[(sp - 1) * BLOCK] = st[(sp - 1) * BLOCK] == 0; break;
      case OP_CAST: st[(sp - 1) * BLOCK] = cast(in.a, in.t, st[(sp - 1) * BLOCK]); break;
      case OP_IF: {
        int64_t e = st[--sp * BLOCK], t = st[--sp * BLOCK], c = st[(sp - 1) * BLOCK];
        st[(sp - 1) * BLOCK] = c ? t : e;
        break;
      }
      case OP_MIN: case OP_MAX: {
        int64_t b = st[--sp * BLOCK], a = st[(sp - 1) * BLOCK];
        bool lt = in.t == F64 ? as_f(a) < as_f(b) : a < b;
        st[(sp - 1) * BLOCK] = (in.op == OP_MIN) == lt ? a : b;
        break;
      }
      case OP_BLOOM: st[(sp - 1) * BLOCK] = bloom_bit(in.imm, in.a, st[(sp - 1) * BLOCK]); break;
      default: break;
    }
  }
  return st[0];
}

constexpr int kVmBlock = 256;

__global__ __launch_bounds__(kVmBlock) void k_vm_eval(Prog p, uint64_t n, void *out, int out_t) {
  extern __shared__ int64_t vm_st[];
  int64_t *st = vm_st + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * kVmBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kVmBlock + threadIdx.x; i < n; i += stride)
    st_slot(out, out_t, i, run_prog<kVmBlock>(p, i, st));
}

// Ordered compaction.  Tiles of kSelTile rows: pass 1 evaluates the
// predicate into a byte mask and counts each tile, a one-workgroup scan turns
// counts into offsets, pass 2 writes the row ids of each tile in order.
constexpr int kSelTile = 4096;

__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t *tmp, uint32_t *tot) {
  constexpr int NW = kVmBlock / 64;
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t x = __shfl_up(incl, o, 64);
    if (lane >= o) incl += x;
  }
  if (lane == 63) tmp[wid] = incl;
  __syncthreads();
  uint32_t pre = 0, t = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    uint32_t x = tmp[w];
    pre += w < wid ? x : 0;
    t += x;
  }
  *tot = t;
  __syncthreads();
  return pre + incl - v;
}

__global__ __launch_bounds__(kVmBlock) void k_sel_mask(Prog p, uint64_t n, uint8_t *mask,
                                                        uint32_t *tile_cnt) {
  extern __shared__ int64_t vm_st[];
  __shared__ uint32_t red[kVmBlock / 64];
  int64_t *st = vm_st + threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kSelTile;
  uint32_t c = 0;
  for (int j = threadIdx.x; j < kSelTile; j += kVmBlock) {
    uint64_t i = base + j;
    if (i < n) {
      bool keep = run_prog<kVmBlock>(p, i, st) != 0;
      mask[i] = keep;
      c += keep;
    }
  }
  c = (uint32_t)wave_sum_i64(c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kVmBlock / 64; ++w) t += red[w];
    tile_cnt[blockIdx.x] = t;
  }
}

// Mask of non-zero u32 entries (occupied hash slots) with tile counts.
__global__ __launch_bounds__(kVmBlock) void k_nz_mask(const uint32_t *v, uint64_t n,
                                                       uint8_t *mask, uint32_t *tile_cnt) {
  __shared__ uint32_t red[kVmBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kSelTile;
  uint32_t c = 0;
  for (int j = threadIdx.x; j < kSelTile; j += kVmBlock) {
    uint64_t i = base + j;
    if (i < n) {
      bool keep = v[i] != 0;
      mask[i] = keep;
      c += keep;
    }
  }
  c = (uint32_t)wave_sum_i64(c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kVmBlock / 64; ++w) t += red[w];
    tile_cnt[blockIdx.x] = t;
  }
}

// Exclusive scan of m u32 counts into u64 offsets, one workgroup (each
// thread scans a contiguous run): off[i] = sum of cnt[0 .. i), off[m] = the
// total; *mx (mx may be null) = the largest count.
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t *cnt, uint64_t m,
                                                       unsigned long long *off,
                                                       unsigned long long *mx) {
  __shared__ unsigned long long part[1024];
  __shared__ uint32_t s_mx;
  if (threadIdx.x == 0) s_mx = 0;
  const uint64_t per = (m + 1023) / 1024;
  const uint64_t a = threadIdx.x * per, b = a + per < m ? a + per : m;
  unsigned long long s = 0;
  uint32_t lm = 0;
  for (uint64_t i = a; i < b; ++i) {
    s += cnt[i];
    lm = cnt[i] > lm ? cnt[i] : lm;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (mx && lm) atomicMax(&s_mx, lm);
  for (int o = 1; o < 1024; o <<= 1) {
    unsigned long long x = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned long long run = part[threadIdx.x] - s;
  for (uint64_t i = a; i < b; ++i) {
    off[i] = run;
    run += cnt[i];
  }
  if (threadIdx.x == 1023) off[m] = part[1023];
  // every thread's atomicMax came before the scan's barriers
  if (mx && threadIdx.x == 0) *mx = s_mx;
}

__global__ __launch_bounds__(kVmBlock) void k_sel_write(const uint8_t *mask, uint64_t n,
                                                         const unsigned long long *off,
                                                         uint32_t *out) {
  __shared__ uint32_t tmp[kVmBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kSelTile;
  unsigned long long pos = off[blockIdx.x];
  constexpr int IPT = kSelTile / kVmBlock;  // 16 consecutive rows per thread
  const uint64_t r0 = base + (uint64_t)threadIdx.x * IPT;
  uint32_t bits = 0, c = 0;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    uint64_t i = r0 + j;
    bool k = i < n && mask[i];
    bits |= (uint32_t)k << j;
    c += k;
  }
  uint32_t tot;
  uint32_t pre = block_scan_excl(c, tmp, &tot);
  unsigned long long p = pos + pre;
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (bits >> j & 1) out[p++] = (uint32_t)(r0 + j);
}

__global__ void k_gather(const void *src, int esize, const uint32_t *idx, uint64_t n, void *dst) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t j = idx[i];
    if (esize == 4) ((uint32_t *)dst)[i] = ((const uint32_t *)src)[j];
    else if (esize == 8) ((uint64_t *)dst)[i] = ((const uint64_t *)src)[j];
    else ((uint8_t *)dst)[i] = ((const uint8_t *)src)[j];
  }
}

__global__ void k_iota(uint32_t *d, uint64_t n) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    d[i] = (uint32_t)i;
}

// Reductions ---------------------------------------------------------------
__device__ __forceinline__ int64_t red_ident(int m, bool f) {
  switch (m) {
    case M_MIN: return f ? of_f(__builtin_inf()) : INT64_MAX;
    case M_MAX: return f ? of_f(-__builtin_inf()) : INT64_MIN;
    case M_AND: return 1;
    default: return f ? of_f(0.0) : 0;
  }
}

__device__ __forceinline__ int64_t red_op(int m, bool f, int64_t a, int64_t b) {
  if (f) {
    double x = as_f(a), y = as_f(b);
    switch (m) {
      case M_MIN: return of_f(x < y ? x : y);
      case M_MAX: return of_f(x > y ? x : y);
      default: return of_f(x + y);
    }
  }
  switch (m) {
    case M_MIN: return a < b ? a : b;
    case M_MAX: return a > b ? a : b;
    case M_AND: return (a != 0) & (b != 0);
    case M_OR: return (a != 0) | (b != 0);
    default: return (int64_t)((uint64_t)a + (uint64_t)b);
  }
}

__device__ __forceinline__ int64_t block_red(int m, bool f, int64_t v) {
  __shared__ int64_t tmp[1024 / 64];
  for (int o = 32; o > 0; o >>= 1) v = red_op(m, f, v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t r = red_ident(m, f);
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = red_op(m, f, r, tmp[w]);
  return r;
}

__global__ __launch_bounds__(256) void k_reduce_part(const void *col, int t, uint64_t n, int m,
                                                      int64_t *part) {
  const bool f = t == F64 || t == F32;
  int64_t acc = red_ident(m, f);
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc = red_op(m, f, acc, ld_slot(col, t, i));
  acc = block_red(m, f, acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(1024) void k_reduce_final(const int64_t *part, int np, int t, int m,
                                                        int64_t *out) {
  const bool f = t == F64 || t == F32;
  int64_t acc = red_ident(m, f);
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc = red_op(m, f, acc, part[i]);
  acc = block_red(m, f, acc);
  if (threadIdx.x == 0) *out = acc;
}

// General group-by ----------------------------------------------------------
struct GbArgs {
  const void *kc[kGbMaxKeys];
  int32_t kt[kGbMaxKeys];
  const void *ac[kGbMaxAggs];
  int32_t at[kGbMaxAggs];
  int32_t am[kGbMaxAggs];
  int32_t nk, na;
};

__device__ __forceinline__ uint64_t key_hash(const GbArgs &a, uint64_t row) {
  uint64_t h = 0x243f6a8885a308d3ULL;
  for (int k = 0; k < a.nk; ++k) {
    uint64_t v = (uint64_t)ld_slot(a.kc[k], a.kt[k], row);
    h = mix64(h ^ (v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2)));
  }
  return h;
}

__device__ __forceinline__ bool key_eq(const GbArgs &a, uint64_t r1, uint64_t r2) {
  for (int k = 0; k < a.nk; ++k)
    if (ld_slot(a.kc[k], a.kt[k], r1) != ld_slot(a.kc[k], a.kt[k], r2)) return false;
  return true;
}

__device__ __forceinline__ void atomic_min_f64(double *p, double v) {
  unsigned long long *q = (unsigned long long *)p;
  unsigned long long old = *q;
  while (v < __builtin_bit_cast(double, old)) {
    unsigned long long prev = atomicCAS(q, old, __builtin_bit_cast(unsigned long long, v));
    if (prev == old) break;
    old = prev;
  }
}
__device__ __forceinline__ void atomic_max_f64(double *p, double v) {
  unsigned long long *q = (unsigned long long *)p;
  unsigned long long old = *q;
  while (v > __builtin_bit_cast(double, old)) {
    unsigned long long prev = atomicCAS(q, old, __builtin_bit_cast(unsigned long long, v));
    if (prev == old) break;
    old = prev;
  }
}

// Insert every row's key (slot = representative row + 1, linear probing,
// at most S probes: a full table raises the overflow flag) and apply its
// aggregates to the slot's accumulators with atomics.
__global__ __launch_bounds__(256) void k_gb_generic(GbArgs a, uint64_t n, uint32_t *slots,
                                                     uint64_t mask, int64_t *acc,
                                                     unsigned int *overflow) {
  const uint64_t S = mask + 1;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n; row += stride) {
    uint64_t s = key_hash(a, row) & mask;
    bool found = false;
    for (uint64_t probes = 0; probes < S; ++probes) {
      uint32_t cur = __hip_atomic_load(&slots[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == 0) {
        uint32_t prev = atomicCAS(&slots[s], 0u, (uint32_t)(row + 1));
        if (prev == 0) {
          found = true;
          break;
        }
        cur = prev;
      }
      if (key_eq(a, row, cur - 1)) {
        found = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!found) {
      atomicOr(overflow, 1u);
      continue;
    }
    for (int j = 0; j < a.na; ++j) {
      int64_t *slot = acc + (uint64_t)j * S + s;
      const int m = a.am[j];
      if (m == M_CNT) {
        atomicAdd((unsigned long long *)slot, 1ULL);
        continue;
      }
      const int t = a.at[j];
      int64_t v = ld_slot(a.ac[j], t, row);
      if (t == F64 || t == F32) {
        double x = as_f(v);
        if (m == M_SUM) unsafeAtomicAdd((double *)slot, x);
        else if (m == M_MIN) atomic_min_f64((double *)slot, x);
        else atomic_max_f64((double *)slot, x);
      } else {
        switch (m) {
          case M_SUM: atomicAdd((unsigned long long *)slot, (unsigned long long)v); break;
          case M_MIN: atomicMin((long long *)slot, (long long)v); break;
          case M_MAX: atomicMax((long long *)slot, (long long)v); break;
          case M_AND: atomicAnd((unsigned long long *)slot, (unsigned long long)(v != 0)); break;
          default: atomicOr((unsigned long long *)slot, (unsigned long long)(v != 0)); break;
        }
      }
    }
  }
}

// The global table's insert of k_gb_generic for the row `row`: slot index,
// or -1 when all S slots are taken by other keys.
__device__ __forceinline__ int64_t gb_global_slot(const GbArgs &a, uint64_t row, uint32_t *slots,
                                                  uint64_t mask) {
  const uint64_t S = mask + 1;
  uint64_t s = key_hash(a, row) & mask;
  for (uint64_t probes = 0; probes < S; ++probes) {
    uint32_t cur = __hip_atomic_load(&slots[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0) {
      const uint32_t prev = atomicCAS(&slots[s], 0u, (uint32_t)(row + 1));
      if (prev == 0) return (int64_t)s;
      cur = prev;
    }
    if (key_eq(a, row, cur - 1)) return (int64_t)s;
    s = (s + 1) & mask;
  }
  return -1;
}

// One integer aggregate value (or a count) into an accumulator with the
// monoid's atomic; works on LDS and global addresses alike.
__device__ __forceinline__ void gb_apply_int(int64_t *slot, int m, int64_t v) {
  switch (m) {
    case M_CNT:
    case M_SUM: atomicAdd((unsigned long long *)slot, (unsigned long long)v); break;
    case M_MIN: atomicMin((long long *)slot, (long long)v); break;
    case M_MAX: atomicMax((long long *)slot, (long long)v); break;
    case M_AND: atomicAnd((unsigned long long *)slot, (unsigned long long)(v != 0)); break;
    default: atomicOr((unsigned long long *)slot, (unsigned long long)(v != 0)); break;
  }
}

// Group-by of integer aggregates (and counts) with few groups: every
// workgroup folds its rows into a private LDS table of LS slots (the
// representative row + 1, 0 = empty, and one accumulator per aggregate),
// with LDS atomics and no global traffic per row, then merges each occupied
// slot into the global table once: the global slot of its representative
// row, and the monoid's atomic on each accumulator (a count adds its local
// count).  Keys are compared through the representative rows, as in
// k_gb_generic.  A row whose key finds neither its slot nor a free one
// within kGbLdsProbes local probes (a workgroup that sees more groups than
// its table holds) goes to the global table directly.  The merge first lists
// the occupied slots (a ballot per wave), so that each thread merges its
// share of the workgroup's groups rather than scanning its share of the
// slots.  LDS: na * LS int64 accumulators, LS * 4 bytes of rows, then the
// list of LS 16-bit slot indices (LS <= 4096).
constexpr uint32_t kGbLdsProbes = 32;
constexpr uint32_t kGbLdsSlotBytes = 4 + 2;  // per slot beside the accumulators
constexpr uint64_t kGbLdsRows = 2048;          // rows per workgroup (grid <= 8 per CU)
constexpr uint64_t kGbMaxCopies = 16;          // global accumulator copies
constexpr uint64_t kGbCopyBytes = 16ULL << 20;
__global__ __launch_bounds__(256) void k_gb_lds(GbArgs a, uint64_t n, uint32_t *slots,
                                                uint64_t mask, int64_t *acc, uint32_t R,
                                                unsigned int *overflow, uint32_t LS) {
  extern __shared__ int64_t lacc[];  // na * LS accumulators, then LS rows
  uint32_t *lrow = reinterpret_cast<uint32_t *>(lacc + (uint64_t)a.na * LS);
  const uint64_t S = mask + 1;
  for (uint32_t i = threadIdx.x; i < LS; i += 256) {
    lrow[i] = 0;
    for (int j = 0; j < a.na; ++j) {
      const int m = a.am[j];
      lacc[(uint64_t)j * LS + i] = m == M_MIN ? INT64_MAX : m == M_MAX ? INT64_MIN : m == M_AND ? 1 : 0;
    }
  }
  __syncthreads();
  const uint32_t lmask = LS - 1;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t row = (uint64_t)blockIdx.x * 256 + threadIdx.x; row < n; row += stride) {
    uint32_t s = (uint32_t)key_hash(a, row) & lmask;
    int64_t *base = nullptr;
    uint64_t step = 0;
    for (uint32_t probes = 0; probes < kGbLdsProbes; ++probes) {
      uint32_t cur = __hip_atomic_load(&lrow[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (cur == 0) {
        const uint32_t prev = atomicCAS(&lrow[s], 0u, (uint32_t)(row + 1));
        if (prev == 0) {
          base = lacc + s;
          break;
        }
        cur = prev;
      }
      if (key_eq(a, row, cur - 1)) {
        base = lacc + s;
        break;
      }
      s = (s + 1) & lmask;
    }
    if (base) {
      step = LS;
    } else {  // the local table is full of other keys: straight to the global one
      const int64_t gs = gb_global_slot(a, row, slots, mask);
      if (gs < 0) {
        atomicOr(overflow, 1u);
        continue;
      }
      base = acc + gs;  // copy 0
      step = (uint64_t)R * S;
    }
    for (int j = 0; j < a.na; ++j) {
      const int m = a.am[j];
      gb_apply_int(base + (uint64_t)j * step, m, m == M_CNT ? 1 : ld_slot(a.ac[j], a.at[j], row));
    }
  }
  __shared__ uint32_t nocc;
  if (threadIdx.x == 0) nocc = 0;
  __syncthreads();
  uint16_t *list = reinterpret_cast<uint16_t *>(lrow + LS);
  const int lane = threadIdx.x & 63;
  for (uint32_t i = threadIdx.x; i < LS; i += 256) {  // LS is a multiple of 256
    const bool occ = lrow[i] != 0;
    const uint64_t bal = __ballot(occ);
    uint32_t base = 0;
    if (lane == 0 && bal) base = atomicAdd(&nocc, (uint32_t)__popcll(bal));
    base = __shfl(base, 0, 64);
    if (occ) list[base + __popcll(bal & ((1ULL << lane) - 1))] = (uint16_t)i;
  }
  __syncthreads();
  const uint32_t nl = nocc;
  const uint64_t copy = blockIdx.x % R;
  for (uint32_t x = threadIdx.x; x < nl; x += 256) {
    const uint32_t i = list[x];
    const uint32_t r = lrow[i];
    const int64_t gs = gb_global_slot(a, r - 1, slots, mask);
    if (gs < 0) {
      atomicOr(overflow, 1u);
      continue;
    }
    for (int j = 0; j < a.na; ++j)
      gb_apply_int(acc + ((uint64_t)j * R + copy) * S + gs, a.am[j], lacc[(uint64_t)j * LS + i]);
  }
}

__global__ void k_fill_i64(int64_t *d, int64_t v, uint64_t n) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) d[i] = v;
}

// Group g of the output = the g-th occupied slot (slot order).
// R > 1 (integer aggregates only): aggregate j of slot s is the monoid over
// copies acc[(j R + r) S + s], r < R.
__global__ void k_gb_emit(GbArgs a, const uint32_t *slots, const uint32_t *occ, uint64_t ng,
                          const int64_t *acc, uint64_t S, uint32_t R, GbOut out) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += stride) {
    const uint32_t s = occ[g];
    const uint64_t rep = slots[s] - 1;
    for (int k = 0; k < a.nk; ++k) st_slot(out.keys[k], a.kt[k], g, ld_slot(a.kc[k], a.kt[k], rep));
    for (int j = 0; j < a.na; ++j) {
      const int64_t *c = acc + (uint64_t)j * R * S + s;
      int64_t v = c[0];
      for (uint32_t r = 1; r < R; ++r) {
        const int64_t x = c[(uint64_t)r * S];
        switch (a.am[j]) {
          case M_MIN: v = x < v ? x : v; break;
          case M_MAX: v = x > v ? x : v; break;
          case M_AND: v &= x; break;
          case M_OR: v |= x; break;
          default: v += x; break;  // sums and counts
        }
      }
      ((int64_t *)out.aggs[j])[g] = v;
    }
  }
}

__global__ void k_hash_buckets(const int32_t *k, uint64_t n, uint32_t buckets, uint32_t *out) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (uint32_t)(hash_i32(k[i]) % buckets);
}

__global__ void k_hash_values(const int32_t *k, uint64_t n, int64_t *out) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (int64_t)hash_i32(k[i]);
}

// Fused probe pipeline ------------------------------------------------------
struct PipeArgs {
  const Instr *ins;
  const void *const *cols;
  const int32_t *ctype;
  const int32_t *clevel;
  void *const *outs;
  const int32_t *out_t;
  PipeTable tables[kPipeMaxLevels];
  uint64_t n;
  uint32_t *counts;  // survivors per workgroup
  uint64_t region;   // output slots per workgroup: block b writes from b * region
  int32_t nins, nout, depth, nlev;
};

constexpr unsigned long long kPipeEmpty = 0x8000000000000000ULL;  // key INT32_MIN
constexpr int kPipeTile = 256;  // rows per wave pass
constexpr int kPipeWave = 64;   // one wave per workgroup

// Tile-at-a-time interpretation, one wave per workgroup: each instruction
// runs over the surviving rows of a kPipeTile-row tile (a list of tile
// positions in LDS), so its dispatch is paid once per tile, and a CU keeps
// ~20 tiles in flight.  A lane owns the survivors lane + 64 j (j < PER); all
// of its loads of one instruction are issued before any result is used, and
// a probe advances all of the lane's keys one slot per round, so a tile
// waits for one memory round trip per probe round, not per row.  PROBE /
// FILTER compact the survivor list in order with one ballot per j (the
// stack is empty there: the host checks it); build row ids live per tile row
// and level; the outputs of a tile take consecutive positions from one
// atomic.
template <class T>
__device__ __forceinline__ void load_col(const T *c, int64_t *d, const uint16_t *sel,
                                         const uint32_t *rl, bool lv0, uint64_t base, uint32_t n,
                                         int lane) {
  constexpr int PER = kPipeTile / kPipeWave;
  T v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t i = lane + j * kPipeWave;
    if (i < n) {
      const uint32_t r = sel[i];
      v[j] = c[lv0 ? base + r : (uint64_t)rl[r]];
    }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t i = lane + j * kPipeWave;
    if (i < n) {
      if constexpr (sizeof(T) == 8) d[i] = __builtin_bit_cast(int64_t, v[j]);
      else if constexpr (std::is_same<T, float>::value) d[i] = of_f((double)v[j]);
      else if constexpr (std::is_same<T, uint8_t>::value) d[i] = v[j] != 0;
      else d[i] = (int64_t)v[j];
    }
  }
}

__global__ __launch_bounds__(kPipeWave) void k_pipe(PipeArgs a) {
  extern __shared__ int64_t sh[];
  int64_t *st = sh;                                                   // depth x TILE
  uint32_t *rid = (uint32_t *)(sh + (size_t)a.depth * kPipeTile);     // nlev x TILE (level 0 unused)
  uint16_t *sel = (uint16_t *)(rid + (size_t)a.nlev * kPipeTile);     // TILE
  const int lane = threadIdx.x;
  constexpr int PER = kPipeTile / kPipeWave;
  const unsigned long long lt = lane ? ((1ULL << lane) - 1) : 0ULL;
  uint64_t written = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * kPipeTile; base < a.n;
       base += (uint64_t)gridDim.x * kPipeTile) {
    const uint32_t rows = (uint32_t)(a.n - base < (uint64_t)kPipeTile ? a.n - base : kPipeTile);
#pragma unroll
    for (int j = 0; j < PER; ++j) sel[lane + j * kPipeWave] = (uint16_t)(lane + j * kPipeWave);
    __syncthreads();
    uint32_t n = rows;
    int sp = 0;
    bool reserved = false;
    unsigned long long p0 = 0;
    for (int k = 0; k < a.nins && n; ++k) {
      const Instr in = a.ins[k];
      switch (in.op) {
        case OP_COL: {
          const void *c = a.cols[in.a];
          const int lv = a.clevel[in.a];
          const uint32_t *rl = rid + (size_t)lv * kPipeTile;
          int64_t *d = st + (size_t)sp * kPipeTile;
          switch (a.ctype[in.a]) {
            case I32: load_col((const int32_t *)c, d, sel, rl, lv == 0, base, n, lane); break;
            case I64:
            case F64: load_col((const int64_t *)c, d, sel, rl, lv == 0, base, n, lane); break;
            case BOOL: load_col((const uint8_t *)c, d, sel, rl, lv == 0, base, n, lane); break;
            default: load_col((const float *)c, d, sel, rl, lv == 0, base, n, lane); break;
          }
          ++sp;
          break;
        }
        case OP_CONST: {
          int64_t *d = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) d[i] = in.imm;
          ++sp;
          break;
        }
        case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MOD: {
          --sp;
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile, *y = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) x[i] = arith(in.op, in.t, x[i], y[i]);
          break;
        }
        case OP_NEG: {
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave)
            x[i] = in.t == F64 ? of_f(-as_f(x[i])) : (in.t == I32 ? wrap32(0 - x[i]) : (int64_t)(0 - (uint64_t)x[i]));
          break;
        }
        case OP_EQ: case OP_NE: case OP_LT: case OP_LE: case OP_GT: case OP_GE: {
          --sp;
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile, *y = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) x[i] = cmp(in.op, in.t, x[i], y[i]);
          break;
        }
        case OP_AND: case OP_OR: {
          --sp;
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile, *y = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave)
            x[i] = in.op == OP_AND ? ((x[i] != 0) & (y[i] != 0)) : ((x[i] != 0) | (y[i] != 0));
          break;
        }
        case OP_NOT: {
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) x[i] = x[i] == 0;
          break;
        }
        case OP_CAST: {
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) x[i] = cast(in.a, in.t, x[i]);
          break;
        }
        case OP_IF: {
          sp -= 2;
          int64_t *c = st + (size_t)(sp - 1) * kPipeTile, *t = st + (size_t)sp * kPipeTile,
                  *e = st + (size_t)(sp + 1) * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) c[i] = c[i] ? t[i] : e[i];
          break;
        }
        case OP_MIN: case OP_MAX: {
          --sp;
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile, *y = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) {
            bool lt2 = in.t == F64 ? as_f(x[i]) < as_f(y[i]) : x[i] < y[i];
            x[i] = (in.op == OP_MIN) == lt2 ? x[i] : y[i];
          }
          break;
        }
        case OP_BLOOM: {
          int64_t *x = st + (size_t)(sp - 1) * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) x[i] = bloom_bit(in.imm, in.a, x[i]);
          break;
        }
        case OP_PROBE:
        case OP_FILTER: {
          --sp;  // the stack is empty after the pop
          bool hit[PER];
          uint16_t rr[PER];
#pragma unroll
          for (int j = 0; j < PER; ++j) {
            const uint32_t i = lane + j * kPipeWave;
            hit[j] = false;
            rr[j] = i < n ? sel[i] : 0;
          }
          if (in.op == OP_FILTER) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              const uint32_t i = lane + j * kPipeWave;
              hit[j] = i < n && st[i] != 0;
            }
          } else {
            const PipeTable tb = a.tables[in.a];
            uint32_t *rl = rid + (size_t)(in.a + 1) * kPipeTile;
            int32_t key[PER];
            uint64_t h[PER];
            bool act[PER];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              const uint32_t i = lane + j * kPipeWave;
              key[j] = i < n ? (int32_t)st[i] : 0;
              bool in;
              h[j] = dim_start(key[j], tb.mask, tb.lo, tb.range, &in);
              act[j] = i < n && in;
            }
            if (tb.ids) {  // direct table of 4-byte build rows: no probe sequence
#pragma unroll
              for (int j = 0; j < PER; ++j) {
                const uint32_t x = act[j] ? tb.ids[h[j]] : 0xffffffffu;
                if (x != 0xffffffffu) {
                  rl[rr[j]] = x;
                  hit[j] = true;
                }
              }
            } else for (;;) {
              // all of the lane's keys advance one slot per round
              bool any = false;
              unsigned long long sv[PER];
#pragma unroll
              for (int j = 0; j < PER; ++j) sv[j] = act[j] ? tb.table[h[j]] : 0ULL;
#pragma unroll
              for (int j = 0; j < PER; ++j) {
                if (!act[j]) continue;
                if (sv[j] == kPipeEmpty) {
                  act[j] = false;
                } else if ((int32_t)(sv[j] >> 32) == key[j]) {
                  rl[rr[j]] = (uint32_t)sv[j];
                  hit[j] = true;
                  act[j] = false;
                } else if (tb.range) {  // a direct slot holds only its own key
                  act[j] = false;
                } else {
                  h[j] = (h[j] + 1) & tb.mask;
                  any = true;
                }
              }
              if (!any) break;
            }
          }
          __syncthreads();  // every lane has read sel and st
          uint32_t o = 0;
#pragma unroll
          for (int j = 0; j < PER; ++j) {
            const unsigned long long m = __ballot(hit[j]);
            if (hit[j]) sel[o + __popcll(m & lt)] = rr[j];
            o += __popcll(m);
          }
          n = o;
          break;
        }
        case OP_OUT: {
          --sp;
          if (!reserved) {
            p0 = (unsigned long long)blockIdx.x * a.region + written;
            written += n;
            reserved = true;
          }
          const int64_t *x = st + (size_t)sp * kPipeTile;
          for (uint32_t i = lane; i < n; i += kPipeWave) st_slot(a.outs[in.a], a.out_t[in.a], p0 + i, x[i]);
          break;
        }
        default: break;
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (lane == 0) a.counts[blockIdx.x] = (uint32_t)written;
}

// Probe chains without interpretation.  The common pipeline shape -- int32
// key column -> probe, column cmp constant -> filter, then plain columns out
// -- runs row-at-a-time: each lane carries independent rows through the
// stages with no barriers (a row ends at its first miss; a wave stops when
// all its rows have), which keeps far more probes in flight than a
// tile-synchronous interpreter.  Stage 0 runs on kChainRowsA rows per lane;
// its survivors are compacted and the later stages run on kChainRowsB rows
// per lane (k_chain_q).
constexpr int kChainRowsA = 4;
constexpr int kChainRowsB = 1;
// Per wave, queue L (1 <= L <= D) holds rows that passed stages 0 .. L-1,
// with those stages' build rows.  Queue 1 takes one tile's stage-0
// survivors (64 RA) on top of fewer than 64 left over from the tiles
// before; queue L >= 2 takes 64 on top of fewer than 64.  D = the number of
// queues: stages 1 .. D-1 each feed the next queue, and level D runs the
// remaining stages and the outputs.
constexpr int kChainQueue = 64 * kChainRowsA + 64 * kChainRowsB;
constexpr int kChainQueueN = 128;
constexpr int kChainMaxQueues = 3;
__host__ __device__ constexpr int chain_queues(int nst) {
  return nst - 1 < 1 ? 1 : (nst - 1 > kChainMaxQueues ? kChainMaxQueues : nst - 1);
}
// LDS words of one wave's queues: row + L build rows per entry of queue L
__host__ __device__ constexpr int chain_queue_words(int nst) {
  int w = 2 * kChainQueue;
  for (int L = 2; L <= chain_queues(nst); ++L) w += (1 + L) * kChainQueueN;
  return w;
}
constexpr int kChainMaxStages = 8;
constexpr int kChainMaxCols = 32;
struct ChainStage {
  int32_t kind;  // 0 probe table `tab` with key column `col`; 1 filter col cmp imm;
                 // 2 bloom filter test of int32 column `col` (words at imm, log2 size op)
  int32_t col;
  int32_t tab;
  int32_t op;    // comparison (filters)
  int64_t imm;
};
struct ChainOut {
  void *out;
  int32_t col;
  int32_t out_t;
};
struct ChainArgs {
  const void *cols[kChainMaxCols];
  int32_t ctype[kChainMaxCols];
  int32_t clevel[kChainMaxCols];
  ChainStage st[kChainMaxStages];
  // Key column of stage s when it is a probe or a bloom filter test on a
  // source-row int32 column (pf_on[s] = 1); such keys are loaded one stage ahead, beside the
  // previous stage's probes.  Otherwise pf[s] points at a one-element
  // dummy read at index 0.
  const int32_t *pf[kChainMaxStages];
  int32_t pf_on[kChainMaxStages];
  ChainOut outs[kPipeMaxOut];
  PipeTable tables[kPipeMaxLevels];
  uint64_t n;
  uint32_t *counts;  // survivors per workgroup
  uint64_t region;   // output slots per workgroup: block b writes from b * region
  int32_t nst, nout;
  int32_t ntab;      // tables (LDS ridl slots per row)
};

// Loads of R values of one column at row indices ix (branch-free: the type
// switch is outside the per-row loop, so all loads are in flight together).
template <class T, int R>
__device__ __forceinline__ void gather_rows(const void *c, const uint32_t (&ix)[R], int64_t (&v)[R]) {
  const T *p = (const T *)c;
  T x[R];
#pragma unroll
  for (int k = 0; k < R; ++k) x[k] = p[ix[k]];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if constexpr (sizeof(T) == 8) v[k] = __builtin_bit_cast(int64_t, x[k]);
    else if constexpr (std::is_same<T, float>::value) v[k] = of_f((double)x[k]);
    else if constexpr (std::is_same<T, uint8_t>::value) v[k] = x[k] != 0;
    else v[k] = (int64_t)x[k];
  }
}

template <int R>
__device__ __forceinline__ void gather_col(const void *c, int ct, const uint32_t (&ix)[R], int64_t (&v)[R]) {
  switch (ct) {
    case I32: gather_rows<int32_t, R>(c, ix, v); break;
    case I64:
    case F64: gather_rows<int64_t, R>(c, ix, v); break;
    case BOOL: gather_rows<uint8_t, R>(c, ix, v); break;
    default: gather_rows<float, R>(c, ix, v); break;
  }
}

// Per-lane state of R rows.  Build row ids of level L >= 1 for row slot k of
// lane t live in LDS at ridl[((L - 1) * R + k) * 256 + t]; level 0 is the
// source row.
template <int R>
struct ChainRows {
  uint32_t row[R];
  bool alive[R];
  uint32_t *ridl;
  __device__ __forceinline__ void at(int lv, uint32_t (&ix)[R]) const {
    if (lv == 0) {
#pragma unroll
      for (int k = 0; k < R; ++k) ix[k] = alive[k] ? row[k] : 0u;
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t r = ridl[((lv - 1) * R + k) * 256 + threadIdx.x];
        ix[k] = alive[k] ? r : 0u;
      }
    }
  }
};

// Stage S of K (template recursion: every stage's descriptor is a
// compile-time index, so it stays in registers).
// Loads stage S's prefetchable keys (unconditionally: rows that are dead or
// not prefetchable read index 0 of the stage's column or dummy).
template <int S, int K, int R>
__device__ __forceinline__ void chain_prefetch(const ChainArgs &a, const ChainRows<R> &cr,
                                               int32_t (&pre)[R]) {
  if constexpr (S < K) {
    const int32_t *c = a.pf[S];
    const bool on = a.pf_on[S] != 0;
    uint32_t ix[R];
#pragma unroll
    for (int k = 0; k < R; ++k) ix[k] = (on && cr.alive[k]) ? cr.row[k] : 0u;
#pragma unroll
    for (int k = 0; k < R; ++k) pre[k] = c[ix[k]];
  }
}

template <int S, int K, int R>
__device__ __forceinline__ void chain_stages(const ChainArgs &a, ChainRows<R> &cr, bool &live,
                                             int32_t (&pre)[R]) {
  if constexpr (S < K) {
    int32_t nxt[R];
    if (live) {
      const ChainStage sg = a.st[S];
      const void *c = a.cols[sg.col];
      const int ct = a.ctype[sg.col];
      int64_t v[R];
      if (a.pf_on[S]) {
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = pre[k];
      } else {
        uint32_t ix[R];
        cr.at(a.clevel[sg.col], ix);
        gather_col<R>(c, ct, ix, v);
      }
      // the next stage's keys stream in beside this stage's probes
      chain_prefetch<S + 1, K, R>(a, cr, nxt);
      if (sg.kind == 1) {
#pragma unroll
        for (int k = 0; k < R; ++k)
          cr.alive[k] = cr.alive[k] && cmp(sg.op, ct == F64 || ct == F32 ? F64 : I64, v[k], sg.imm);
      } else if (sg.kind == 2) {
        // a dead row's key is still a valid int32, so every lane tests a bit
        int64_t bit[R];
#pragma unroll
        for (int k = 0; k < R; ++k) bit[k] = bloom_bit(sg.imm, sg.op, v[k]);
#pragma unroll
        for (int k = 0; k < R; ++k) cr.alive[k] = cr.alive[k] && bit[k] != 0;
      } else {
        const PipeTable tb = a.tables[sg.tab];
        const uint32_t mask = (uint32_t)tb.mask;
        const bool direct = tb.range != 0;
        uint32_t h[R], found[R];
        bool act[R], hit[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
          bool in;
          h[k] = dim_start((int32_t)v[k], tb.mask, tb.lo, tb.range, &in);
          act[k] = cr.alive[k] && in;
          hit[k] = false;
          found[k] = 0;
        }
        if (tb.ids) {  // direct table of 4-byte build rows: one load, no key compare
          uint32_t x[R];
#pragma unroll
          for (int k = 0; k < R; ++k) x[k] = tb.ids[act[k] ? h[k] : 0u];
#pragma unroll
          for (int k = 0; k < R; ++k) {
            hit[k] = act[k] && x[k] != 0xffffffffu;
            found[k] = x[k];
          }
        } else for (;;) {
          unsigned long long sv[R];
#pragma unroll
          for (int k = 0; k < R; ++k) sv[k] = tb.table[act[k] ? h[k] : 0u];
          bool any = false;
#pragma unroll
          for (int k = 0; k < R; ++k) {
            const bool empty = sv[k] == kPipeEmpty;
            const bool eq = (int32_t)(sv[k] >> 32) == (int32_t)v[k];
            const bool done = empty || eq || direct;  // a direct slot holds only its key
            const bool got = act[k] && !empty && eq;
            hit[k] = hit[k] || got;
            found[k] = got ? (uint32_t)sv[k] : found[k];
            h[k] = (act[k] && !done) ? ((h[k] + 1) & mask) : h[k];
            act[k] = act[k] && !done;
            any = any || act[k];
          }
          if (!any) break;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
          cr.alive[k] = hit[k];
          if (hit[k]) cr.ridl[(sg.tab * R + k) * 256 + threadIdx.x] = found[k];
        }
      }
      bool any_alive = false;
#pragma unroll
      for (int k = 0; k < R; ++k) any_alive |= cr.alive[k];
      live = __any(any_alive);
    }
    chain_stages<S + 1, K, R>(a, cr, live, nxt);
  }
}

// Loads R rows of an output column as raw 64-bit slots (I32 sign-extended,
// F32 widened to FP64 bits, I64/F64 as stored), with no branch around the
// loads: 4-byte columns read their word twice.
template <int R>
__device__ __forceinline__ void gather_out(const void *c, int ct, const uint32_t (&ix)[R],
                                           int64_t (&v)[R]) {
  const bool wide = ct == I64 || ct == F64;
  const uint32_t *w = (const uint32_t *)c;
  uint32_t lo[R], hi[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t e = wide ? 2ull * ix[k] : (uint64_t)ix[k];
    lo[k] = w[e];
    hi[k] = w[wide ? e + 1 : e];
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int64_t x64 = (int64_t)(((uint64_t)hi[k] << 32) | lo[k]);
    const int64_t x32 = ct == F32 ? of_f((double)__uint_as_float(lo[k])) : (int64_t)(int32_t)lo[k];
    v[k] = wide ? x64 : x32;
  }
}

// Writes the alive rows of cr to the workgroup's output segment: positions
// from one LDS atomic per wave, the output columns gathered four at a time
// (all their loads in flight together; past the last output the last one is
// loaded again and not stored).
template <int R>
__device__ __forceinline__ void chain_emit(const ChainArgs &a, const ChainRows<R> &cr,
                                           uint32_t *s_cnt, int lane) {
  const unsigned long long lt = lane ? ((1ULL << lane) - 1) : 0ULL;
  uint32_t cnt = 0;
  unsigned long long m[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    m[k] = __ballot(cr.alive[k]);
    cnt += __popcll(m[k]);
  }
  if (!cnt) return;
  unsigned long long p0 = 0;
  if (lane == 0) p0 = (unsigned long long)blockIdx.x * a.region + atomicAdd(s_cnt, cnt);
  p0 = __shfl(p0, 0, 64);
  uint64_t pos[R];
  uint32_t off = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    pos[k] = p0 + off + __popcll(m[k] & lt);
    off += __popcll(m[k]);
  }
  for (int j0 = 0; j0 < a.nout; j0 += 4) {
    int64_t v[4][R];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = min(j0 + u, a.nout - 1);
      const int col = a.outs[j].col;
      uint32_t ix[R];
      cr.at(a.clevel[col], ix);
      gather_out<R>(a.cols[col], a.ctype[col], ix, v[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (j0 + u < a.nout) {
        const ChainOut o = a.outs[j0 + u];
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (cr.alive[k]) st_slot(o.out, o.out_t, pos[k], v[u][k]);
      }
    }
  }
}

// One wave's queues and the state its levels share.
struct ChainWave {
  const ChainArgs *a;
  uint32_t *ridl;
  uint32_t *qrow[kChainMaxQueues + 1];  // index L = queue L
  uint32_t *qrid[kChainMaxQueues + 1];  // queue L: build row of stage s at [s * cap + e]
  uint32_t cap[kChainMaxQueues + 1];
  uint32_t n[kChainMaxQueues + 1];      // entries waiting (wave-uniform)
  int tab[kChainMaxQueues];             // table stage s probes, or -1
  uint32_t *s_cnt;
  int lane;
  unsigned long long lt;
};

// Level L over queue-L entries [e0, e1), one row per lane: the build rows of
// stages 0 .. L-1 come back into the lanes' ridl slots, then stage L runs
// and its survivors go onto queue L + 1 (which is worked off whenever 64
// rows wait there), or, at the last level D, the remaining stages and the
// outputs run.
template <int L, int D, int K>
__device__ __forceinline__ void chain_level(ChainWave &w, uint32_t e0, uint32_t e1) {
  const ChainArgs &a = *w.a;
  ChainRows<1> cr;
  cr.ridl = w.ridl;
  const uint32_t e = e0 + w.lane;
  cr.alive[0] = e < e1;
  const uint32_t ec = cr.alive[0] ? e : e0;
  cr.row[0] = w.qrow[L][ec];
#pragma unroll
  for (int s = 0; s < L; ++s)
    if (w.tab[s] >= 0)
      w.ridl[w.tab[s] * 256 + threadIdx.x] = cr.alive[0] ? w.qrid[L][s * w.cap[L] + ec] : 0u;
  bool live = true;
  int32_t pre[1];
  if constexpr (L < D) {
    constexpr int M = L + 1;
    chain_prefetch<L, M, 1>(a, cr, pre);
    chain_stages<L, M, 1>(a, cr, live, pre);
    const unsigned long long m = __ballot(cr.alive[0]);
    if (cr.alive[0]) {
      const uint32_t at = w.n[M] + __popcll(m & w.lt);
      w.qrow[M][at] = cr.row[0];
#pragma unroll
      for (int s = 0; s < M; ++s)
        if (w.tab[s] >= 0) w.qrid[M][s * w.cap[M] + at] = w.ridl[w.tab[s] * 256 + threadIdx.x];
    }
    w.n[M] += __popcll(m);
    __builtin_amdgcn_wave_barrier();
    if (w.n[M] >= 64) {
      chain_level<M, D, K>(w, 0, 64);
      // the rest (< 64 entries) moves to the front: no overlap
      for (uint32_t j = w.lane; j < w.n[M] - 64; j += 64) {
        w.qrow[M][j] = w.qrow[M][64 + j];
#pragma unroll
        for (int s = 0; s < M; ++s) w.qrid[M][s * w.cap[M] + j] = w.qrid[M][s * w.cap[M] + 64 + j];
      }
      w.n[M] -= 64;
    }
  } else {
    chain_prefetch<L, K, 1>(a, cr, pre);
    chain_stages<L, K, 1>(a, cr, live, pre);
    if (live) chain_emit<1>(a, cr, w.s_cnt, w.lane);
  }
  __builtin_amdgcn_wave_barrier();
}

// The probe chain with the survivors of the first stages compacted.  Each
// wave takes tiles of 64 * RA consecutive rows: stage 0 runs on RA rows per
// lane, and the rows still alive (with the build row stage 0 found, when it
// probes a table) go onto the wave's first LDS queue.  Whenever 64 rows wait
// in a queue, the next stage runs on them, one per lane, and its survivors
// go onto the next queue; the last queue's level runs the remaining stages
// and the outputs.  Rows left in the queues at the end are flushed, level by
// level, with fewer lanes.  Selective stages (the star plan's date join
// keeps 1/7 of the rows, the supplier join 1/5 of those, the part join 1/25
// of those) would otherwise run every later stage, a chain of dependent
// memory round trips, with most lanes dead.  The waves of a workgroup share
// only the output counter; a wave's queues and the ridl slots of its own
// lanes are written and read by that wave alone, in program order (the
// one-row slots of the levels alias phase A's, whose stage-0 rows are copied
// onto the first queue before a level runs).
template <int K, int RA, int RB>
__global__ __launch_bounds__(256) void k_chain_q(const ChainArgs *__restrict__ ap) {
  static_assert(RB == 1, "the levels run one row per lane");
  static_assert(RB <= RA, "the levels reuse phase A's ridl slots");
  constexpr int D = chain_queues(K);
  constexpr uint32_t QW = chain_queue_words(K);
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t s_cnt;
  const ChainArgs &a = *ap;
  const int ntab = a.ntab;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  ChainWave w;
  w.a = ap;
  w.ridl = lds;  // ntab * RA * 256
  w.s_cnt = &s_cnt;
  w.lane = lane;
  w.lt = lane ? ((1ULL << lane) - 1) : 0ULL;
  {
    uint32_t *q = lds + ntab * RA * 256 + wid * QW;
    w.cap[0] = 0;
    w.qrow[0] = w.qrid[0] = nullptr;
#pragma unroll
    for (int L = 1; L <= kChainMaxQueues; ++L) {
      w.cap[L] = L == 1 ? kChainQueue : kChainQueueN;
      w.qrow[L] = q;
      w.qrid[L] = q + w.cap[L];
      w.n[L] = 0;
      if (L <= D) q += (1 + L) * w.cap[L];
    }
#pragma unroll
    for (int st = 0; st < kChainMaxQueues; ++st)
      w.tab[st] = (st < K && a.st[st].kind == 0) ? a.st[st].tab : -1;
  }
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int t0 = w.tab[0];
  const uint64_t stride = (uint64_t)gridDim.x * 256u * RA;  // 64-bit: no wrap near n = 2^32
  auto tile_rows = [&](uint64_t base, ChainRows<RA> &c) {
    c.ridl = w.ridl;
#pragma unroll
    for (int k = 0; k < RA; ++k) {
      const uint64_t r = base + k * 64u + lane;
      c.alive[k] = r < a.n;
      c.row[k] = c.alive[k] ? (uint32_t)r : 0u;
    }
  };
  // stage 0's keys of the wave's next tile, loaded while the levels work
  // on the current one
  int32_t pre_next[RA];
  uint64_t base = (uint64_t)blockIdx.x * 256u * RA + wid * 64u * RA;
  {
    ChainRows<RA> cn;
    tile_rows(base, cn);
    chain_prefetch<0, 1, RA>(a, cn, pre_next);
  }
  for (; base < a.n; base += stride) {
    // ---- phase A: stage 0 on RA rows per lane --------------------------------
    ChainRows<RA> ca;
    tile_rows(base, ca);
    {
      bool live = true;
      int32_t pre[RA];
#pragma unroll
      for (int k = 0; k < RA; ++k) pre[k] = pre_next[k];
      chain_stages<0, 1, RA>(a, ca, live, pre);
    }
    // ---- compaction of the survivors onto the first queue -------------------
    uint32_t *qrow = w.qrow[1], *qrid = w.qrid[1];
    uint32_t qn = w.n[1];
#pragma unroll
    for (int k = 0; k < RA; ++k) {
      const unsigned long long m = __ballot(ca.alive[k]);
      if (ca.alive[k]) {
        const uint32_t at = qn + __popcll(m & w.lt);
        qrow[at] = ca.row[k];
        if (t0 >= 0) qrid[at] = w.ridl[(t0 * RA + k) * 256 + threadIdx.x];
      }
      qn += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    {
      ChainRows<RA> cn;
      tile_rows(base + stride, cn);
      chain_prefetch<0, 1, RA>(a, cn, pre_next);
    }
    // ---- level 1 on whole groups of 64 queued rows; the rest waits ----------
    uint32_t q0 = 0;
    for (; q0 + 64u <= qn; q0 += 64u) chain_level<1, D, K>(w, q0, q0 + 64u);
    // the rest (< 64 <= q0 entries, so source and target do not overlap)
    // moves to the front
    const uint32_t rem = qn - q0;
    if (q0)
      for (uint32_t j = lane; j < rem; j += 64) {
        qrow[j] = qrow[q0 + j];
        qrid[j] = qrid[q0 + j];
      }
    w.n[1] = rem;
    __builtin_amdgcn_wave_barrier();
  }
  // the rows left in the queues, level by level
  if (w.n[1]) chain_level<1, D, K>(w, 0, w.n[1]);
  if constexpr (D >= 2) {
    if (w.n[2]) chain_level<2, D, K>(w, 0, w.n[2]);
  }
  if constexpr (D >= 3) {
    if (w.n[3]) chain_level<3, D, K>(w, 0, w.n[3]);
  }
  __syncthreads();
  if (threadIdx.x == 0) a.counts[blockIdx.x] = s_cnt;
}

// Moves every workgroup's output segment to its place in the dense output,
// all columns in one launch: blockIdx.y = segment, blockIdx.z = column.
struct SegCols {
  const void *src[kPipeMaxOut];
  void *dst[kPipeMaxOut];
  int32_t esize[kPipeMaxOut];
};
__global__ void k_seg_compact(SegCols sc, const uint32_t *counts,
                              const unsigned long long *offs, uint64_t region) {
  const uint32_t b = blockIdx.y, col = blockIdx.z;
  const int esize = sc.esize[col];
  const uint32_t c = counts[b];
  const char *s = (const char *)sc.src[col] + (uint64_t)b * region * esize;
  char *d = (char *)sc.dst[col] + offs[b] * esize;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
    if (esize == 8) ((uint64_t *)d)[i] = ((const uint64_t *)s)[i];
    else if (esize == 4) ((uint32_t *)d)[i] = ((const uint32_t *)s)[i];
    else d[i] = s[i];
  }
}

// The program as a probe chain, or false.
static bool as_chain(const PipeProg &p, void *const *outs, ChainArgs *a) {
  if (p.cols.size() > (size_t)kChainMaxCols) return false;
  for (size_t c = 0; c < p.cols.size(); ++c) {
    a->cols[c] = p.cols[c];
    a->ctype[c] = p.ctype[c];
    a->clevel[c] = p.clevel[c];
  }
  a->nst = a->nout = 0;
  const auto &I = p.ins;
  size_t k = 0;
  while (k < I.size()) {
    if (k + 1 < I.size() && I[k].op == OP_COL && I[k + 1].op == OP_PROBE) {
      if (p.ctype[I[k].a] != I32 || a->nst >= kChainMaxStages) return false;
      a->st[a->nst++] = ChainStage{0, I[k].a, I[k + 1].a, 0, 0};
      k += 2;
    } else if (k + 2 < I.size() && I[k].op == OP_COL && I[k + 1].op == OP_BLOOM &&
               I[k + 2].op == OP_FILTER) {
      if (p.ctype[I[k].a] != I32 || a->nst >= kChainMaxStages) return false;
      a->st[a->nst++] = ChainStage{2, I[k].a, 0, I[k + 1].a, I[k + 1].imm};
      k += 3;
    } else if (k + 3 < I.size() && I[k].op == OP_COL && I[k + 1].op == OP_CONST &&
               I[k + 2].op >= OP_EQ && I[k + 2].op <= OP_GE && I[k + 3].op == OP_FILTER) {
      const int ct = p.ctype[I[k].a];
      const bool f = ct == F64 || ct == F32;
      // the comparison must be the column's own type against a constant of it
      if (I[k + 2].t != (f ? F64 : I[k + 2].t) || (f && I[k + 2].t != F64) ||
          (!f && I[k + 2].t == F64) || a->nst >= kChainMaxStages)
        return false;
      a->st[a->nst++] = ChainStage{1, I[k].a, 0, I[k + 2].op, I[k + 1].imm};
      k += 4;
    } else {
      break;
    }
  }
  while (k < I.size()) {
    if (k + 1 < I.size() && I[k].op == OP_COL && I[k + 1].op == OP_OUT) {
      const int ct = p.ctype[I[k].a];
      const int ot = p.out_t[I[k + 1].a];
      const int nat = ct == F32 ? F64 : ct;
      if (ct == BOOL || ot != nat || a->nout >= kPipeMaxOut || I[k + 1].a != a->nout)
        return false;
      a->outs[a->nout++] = ChainOut{outs[I[k + 1].a], I[k].a, ot};
      k += 2;
    } else {
      return false;
    }
  }
  return true;
}

// Host side -----------------------------------------------------------------
static uint64_t grid_for(pj_ctx *ctx, uint64_t n, int block) {
  uint64_t g = ceil_div(n, (uint64_t)block);
  uint64_t cap = (uint64_t)ctx->num_cus * 8;
  if (g > cap) g = cap;
  return g ? g : 1;
}

static int check_prog(const Prog &p) {
  if (p.n <= 0 || p.n > kMaxInstr) return fail(PJ_E_ARG, "expression program of %d instructions", p.n);
  if (p.depth <= 0 || p.depth > kMaxDepth)
    return fail(PJ_E_UNSUPPORTED, "expression needs %d stack slots (limit %d)", p.depth, kMaxDepth);
  for (int k = 0; k < p.n; ++k) {
    if (p.ins[k].op == OP_COL && (p.ins[k].a < 0 || p.ins[k].a >= p.ncols || !p.cols[p.ins[k].a]))
      return fail(PJ_E_ARG, "expression reads an unbound column");
    if (p.ins[k].op == OP_BLOOM && (p.ins[k].a < 0 || p.ins[k].a > 32 || !p.ins[k].imm))
      return fail(PJ_E_ARG, "expression bloom filter without words or with log2 size %d", p.ins[k].a);
  }
  return PJ_OK;
}

int eval(pj_ctx *ctx, const Prog &p, uint64_t n, void *out, int out_t) {
  PJ_TRY(check_prog(p));
  if (!n) return PJ_OK;
  size_t lds = (size_t)p.depth * kVmBlock * 8;
  PJ_LAUNCH(ctx, "rel_eval", k_vm_eval, grid_for(ctx, n, kVmBlock), kVmBlock, lds, p, n, out, out_t);
  return PJ_OK;
}

static int compact_mask(pj_ctx *ctx, const uint8_t *mask, uint64_t n, uint64_t tiles,
                        uint32_t *tile_cnt, uint32_t *out_idx, uint64_t *n_out) {
  Scratch off;
  PJ_TRY(off.alloc(ctx, (tiles + 1) * 8));
  PJ_LAUNCH(ctx, "rel_scan", k_scan_counts, 1, 1024, 0, tile_cnt, tiles,
            off.as<unsigned long long>(), (unsigned long long *)nullptr);
  PJ_LAUNCH(ctx, "rel_compact", k_sel_write, tiles, kVmBlock, 0, mask, n,
            off.as<const unsigned long long>(), out_idx);
  unsigned long long total = 0;
  PJ_TRY(d2h(ctx, &total, off.as<unsigned long long>() + tiles, 8));
  *n_out = total;
  return PJ_OK;
}

int select(pj_ctx *ctx, const Prog &pred, uint64_t n, uint32_t *out_idx, uint64_t *n_out) {
  PJ_TRY(check_prog(pred));
  *n_out = 0;
  if (!n) return PJ_OK;
  if (n > 0xffffffffULL) return fail(PJ_E_CAPACITY, "select over more than 2^32-1 rows");
  const uint64_t tiles = ceil_div(n, kSelTile);
  Scratch mask, cnt;
  PJ_TRY(mask.alloc(ctx, n));
  PJ_TRY(cnt.alloc(ctx, tiles * 4));
  size_t lds = (size_t)pred.depth * kVmBlock * 8;
  PJ_LAUNCH(ctx, "rel_select", k_sel_mask, tiles, kVmBlock, lds, pred, n, mask.as<uint8_t>(),
            cnt.as<uint32_t>());
  return compact_mask(ctx, mask.as<uint8_t>(), n, tiles, cnt.as<uint32_t>(), out_idx, n_out);
}

int gather(pj_ctx *ctx, const void *src, int esize, const uint32_t *idx, uint64_t n, void *dst) {
  if (!n) return PJ_OK;
  if (esize != 1 && esize != 4 && esize != 8) return fail(PJ_E_ARG, "gather of %d-byte elements", esize);
  PJ_LAUNCH(ctx, "rel_gather", k_gather, grid_for(ctx, n, 256), 256, 0, src, esize, idx, n, dst);
  return PJ_OK;
}

int iota(pj_ctx *ctx, uint32_t *d, uint64_t n) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "rel_iota", k_iota, grid_for(ctx, n, 256), 256, 0, d, n);
  return PJ_OK;
}

int reduce(pj_ctx *ctx, const void *col, int t, uint64_t n, int monoid, int64_t *out_i,
           double *out_f) {
  const bool f = t == F64 || t == F32;
  if (monoid == M_CNT) {
    if (out_i) *out_i = (int64_t)n;
    if (out_f) *out_f = (double)n;
    return PJ_OK;
  }
  if (f && (monoid == M_AND || monoid == M_OR))
    return fail(PJ_E_UNSUPPORTED, "and/or accumulator over a float expression");
  Scratch part, res;
  const int np = (int)std::min<uint64_t>(1024, grid_for(ctx, n ? n : 1, 256));
  PJ_TRY(part.alloc(ctx, (uint64_t)np * 8));
  PJ_TRY(res.alloc(ctx, 8));
  PJ_LAUNCH(ctx, "rel_reduce", k_reduce_part, np, 256, 0, col, t, n, monoid, part.as<int64_t>());
  PJ_LAUNCH(ctx, "rel_reduce_final", k_reduce_final, 1, 1024, 0, part.as<const int64_t>(), np, t,
            monoid, res.as<int64_t>());
  int64_t v = 0;
  PJ_TRY(d2h(ctx, &v, res.p, 8));
  if (f) {
    double d = __builtin_bit_cast(double, v);
    if (out_f) *out_f = d;
    if (out_i) *out_i = (int64_t)d;
  } else {
    if (out_i) *out_i = v;
    if (out_f) *out_f = (double)v;
  }
  return PJ_OK;
}

// MIN and MAX of an int32 column in one pass: per-workgroup partials
// part[2 b] = min, part[2 b + 1] = max (four loads in flight per thread),
// then one workgroup folds them into res[0], res[1].
__global__ __launch_bounds__(256) void k_minmax_i32(const int32_t *__restrict__ v, uint64_t n,
                                                    int64_t *__restrict__ part) {
  __shared__ int64_t smn[4], smx[4];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  const uint64_t stride = (uint64_t)gridDim.x * 1024;
  for (uint64_t b = (uint64_t)blockIdx.x * 1024 + threadIdx.x; b < n; b += stride) {
    int32_t x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = v[b + j * 256 < n ? b + j * 256 : b];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      mn = x[j] < mn ? x[j] : mn;
      mx = x[j] > mx ? x[j] : mx;
    }
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  if ((threadIdx.x & 63) == 0) {
    smn[threadIdx.x >> 6] = mn;
    smx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      mn = smn[w] < mn ? smn[w] : mn;
      mx = smx[w] > mx ? smx[w] : mx;
    }
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}
__global__ __launch_bounds__(256) void k_minmax_fold(const int64_t *__restrict__ part, int np,
                                                     int64_t *__restrict__ res) {
  __shared__ int64_t smn[4], smx[4];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (int i = threadIdx.x; i < np; i += 256) {
    mn = part[2 * i] < mn ? part[2 * i] : mn;
    mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  if ((threadIdx.x & 63) == 0) {
    smn[threadIdx.x >> 6] = mn;
    smx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      mn = smn[w] < mn ? smn[w] : mn;
      mx = smx[w] > mx ? smx[w] : mx;
    }
    res[0] = mn;
    res[1] = mx;
  }
}

int minmax_i32_device(pj_ctx *ctx, const int32_t *col, uint64_t n, Scratch &part, int64_t *d_res) {
  if (!n) return fail(PJ_E_ARG, "minmax of no rows");
  const int np = (int)std::min<uint64_t>(1024, std::max<uint64_t>(1, ceil_div(n, 8192)));
  PJ_TRY(part.alloc(ctx, (uint64_t)np * 16));
  PJ_LAUNCH(ctx, "rel_minmax", k_minmax_i32, np, 256, 0, col, n, part.as<int64_t>());
  PJ_LAUNCH(ctx, "rel_minmax_fold", k_minmax_fold, 1, 256, 0, part.as<const int64_t>(), np,
            d_res);
  return PJ_OK;
}

int d2h_many(pj_ctx *ctx, int n, void *const *dst, const void *const *src, const uint64_t *bytes) {
  uint64_t tot = 0;
  for (int i = 0; i < n; ++i) tot += (bytes[i] + 15) & ~15ULL;
  if (tot > kBounceBytes || !ctx->pinned_rb) {
    for (int i = 0; i < n; ++i) PJ_TRY(d2h(ctx, dst[i], src[i], bytes[i]));
    return PJ_OK;
  }
  std::lock_guard<std::mutex> g(ctx->rb_mu);
  char *rb = (char *)ctx->pinned_rb;
  uint64_t off = 0;
  for (int i = 0; i < n; ++i) {
    if (bytes[i])
      PJ_HIP(hipMemcpyAsync(rb + off, src[i], bytes[i], hipMemcpyDeviceToHost, ctx->stream));
    off += (bytes[i] + 15) & ~15ULL;
  }
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  off = 0;
  for (int i = 0; i < n; ++i) {
    if (bytes[i]) memcpy(dst[i], rb + off, bytes[i]);
    off += (bytes[i] + 15) & ~15ULL;
  }
  return PJ_OK;
}

int reduce_minmax(pj_ctx *ctx, const void *col, int t, uint64_t n, int64_t *mn, int64_t *mx) {
  if (t == F64 || t == F32) return fail(PJ_E_ARG, "reduce_minmax over an integer column only");
  if (!n) return fail(PJ_E_ARG, "reduce_minmax of no rows");
  if (t == I32) {  // one pass for both, one fold, one read-back
    Scratch part, res;
    PJ_TRY(res.alloc(ctx, 16));
    PJ_TRY(minmax_i32_device(ctx, (const int32_t *)col, n, part, res.as<int64_t>()));
    int64_t v[2] = {0, 0};
    PJ_TRY(d2h(ctx, v, res.p, 16));
    *mn = v[0];
    *mx = v[1];
    return PJ_OK;
  }
  Scratch part, res;
  const int np = (int)std::min<uint64_t>(1024, grid_for(ctx, n, 256));
  PJ_TRY(part.alloc(ctx, (uint64_t)np * 16));
  PJ_TRY(res.alloc(ctx, 16));
  // both reductions queued, then one read-back of the two results
  PJ_LAUNCH(ctx, "rel_reduce", k_reduce_part, np, 256, 0, col, t, n, (int)M_MIN, part.as<int64_t>());
  PJ_LAUNCH(ctx, "rel_reduce_final", k_reduce_final, 1, 1024, 0, part.as<const int64_t>(), np, t,
            (int)M_MIN, res.as<int64_t>());
  PJ_LAUNCH(ctx, "rel_reduce", k_reduce_part, np, 256, 0, col, t, n, (int)M_MAX,
            part.as<int64_t>() + np);
  PJ_LAUNCH(ctx, "rel_reduce_final", k_reduce_final, 1, 1024, 0, part.as<const int64_t>() + np, np,
            t, (int)M_MAX, res.as<int64_t>() + 1);
  int64_t v[2] = {0, 0};
  PJ_TRY(d2h(ctx, v, res.p, 16));
  *mn = v[0];
  *mx = v[1];
  return PJ_OK;
}

int groupby(pj_ctx *ctx, const GbKey *keys, int nk, const GbAgg *aggs, int na, uint64_t n,
            uint64_t max_groups, const GbOut &out, uint64_t *n_groups) {
  *n_groups = 0;
  if (nk < 1 || nk > kGbMaxKeys) return fail(PJ_E_UNSUPPORTED, "group-by over %d keys (1..%d)", nk, kGbMaxKeys);
  if (na < 0 || na > kGbMaxAggs) return fail(PJ_E_UNSUPPORTED, "group-by with %d aggregates (max %d)", na, kGbMaxAggs);
  if (n > 0xfffffffeULL) return fail(PJ_E_CAPACITY, "group-by input above 2^32-2 rows");
  if (!n) return PJ_OK;
  if (max_groups < 1) max_groups = 1;
  GbArgs a{};
  a.nk = nk;
  a.na = na;
  for (int k = 0; k < nk; ++k) {
    a.kc[k] = keys[k].col;
    a.kt[k] = keys[k].t;
  }
  for (int j = 0; j < na; ++j) {
    a.ac[j] = aggs[j].col;
    a.at[j] = aggs[j].t;
    a.am[j] = aggs[j].monoid;
    if (aggs[j].monoid != M_CNT && !aggs[j].col) return fail(PJ_E_ARG, "aggregate %d has no column", j);
  }
  uint64_t S = 2ULL << ilog2_ceil(std::min<uint64_t>(max_groups, n));
  if (S < 64) S = 64;
  // integer aggregates over few groups: workgroup-private LDS tables
  bool ints = true;
  for (int j = 0; j < na; ++j)
    ints &= aggs[j].monoid == M_CNT || !(aggs[j].t == F64 || aggs[j].t == F32);
  uint32_t LS = 0;  // local slots: a power of two, <= S, rows + list + accumulators in 64 KiB
  if (ints && n >= (1u << 16))
    for (uint32_t l = 4096; l >= 256 && !LS; l >>= 1)
      if (l <= S && (uint64_t)l * (kGbLdsSlotBytes + 8 * (uint64_t)na) <= 65536) LS = l;
  // With LDS tables, every workgroup merges each of its groups into the
  // global accumulators once: R copies of them (workgroup b merges into copy
  // b % R, the emit folds the copies) keep the same-address atomics of a
  // small group count to about 32 per address.  Copies are bounded to
  // kGbCopyBytes in all.
  const uint64_t grid = LS ? std::min<uint64_t>(ceil_div(n, kGbLdsRows), (uint64_t)ctx->num_cus * 8) : 0;
  uint32_t R = 1;
  if (LS) {
    const uint64_t per = std::max<uint64_t>(1, (uint64_t)na) * S * 8;
    R = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({ceil_div(grid, 32), kGbMaxCopies, kGbCopyBytes / per}));
  }
  Scratch slots, acc, flag;
  PJ_TRY(slots.alloc(ctx, S * 4));
  PJ_TRY(zero(ctx, slots.p, S * 4));
  PJ_TRY(acc.alloc(ctx, std::max<uint64_t>(1, (uint64_t)na) * R * S * 8));
  PJ_TRY(flag.alloc(ctx, 4));
  PJ_TRY(zero(ctx, flag.p, 4));
  for (int j = 0; j < na; ++j) {  // aggregate j: copies j R .. j R + R - 1 of S slots
    const bool f = aggs[j].monoid != M_CNT && (aggs[j].t == F64 || aggs[j].t == F32);
    int64_t ident;
    switch (aggs[j].monoid) {
      case M_MIN: ident = f ? __builtin_bit_cast(int64_t, __builtin_inf()) : INT64_MAX; break;
      case M_MAX: ident = f ? __builtin_bit_cast(int64_t, -__builtin_inf()) : INT64_MIN; break;
      case M_AND: ident = 1; break;
      default: ident = 0; break;
    }
    PJ_LAUNCH(ctx, "rel_fill", k_fill_i64, grid_for(ctx, R * S, 256), 256, 0,
              acc.as<int64_t>() + (uint64_t)j * R * S, ident, R * S);
  }
  if (LS) {
    PJ_LAUNCH(ctx, "rel_groupby_lds", k_gb_lds, grid, 256, (size_t)LS * (kGbLdsSlotBytes + 8 * na), a, n,
              slots.as<uint32_t>(), S - 1, acc.as<int64_t>(), R, flag.as<unsigned int>(), LS);
  } else {
    PJ_LAUNCH(ctx, "rel_groupby", k_gb_generic, grid_for(ctx, n, 256), 256, 0, a, n,
              slots.as<uint32_t>(), S - 1, acc.as<int64_t>(), flag.as<unsigned int>());
  }
  // the occupied slots, then the overflow flag and their count in one read-back
  const uint64_t tiles = ceil_div(S, kSelTile);
  Scratch mask, cnt, occ, off;
  PJ_TRY(mask.alloc(ctx, S));
  PJ_TRY(cnt.alloc(ctx, tiles * 4));
  PJ_TRY(occ.alloc(ctx, S * 4));
  PJ_TRY(off.alloc(ctx, (tiles + 1) * 8));
  PJ_LAUNCH(ctx, "rel_groupby_occ", k_nz_mask, tiles, kVmBlock, 0, slots.as<const uint32_t>(), S,
            mask.as<uint8_t>(), cnt.as<uint32_t>());
  PJ_LAUNCH(ctx, "rel_scan", k_scan_counts, 1, 1024, 0, cnt.as<const uint32_t>(), tiles,
            off.as<unsigned long long>(), (unsigned long long *)nullptr);
  PJ_LAUNCH(ctx, "rel_compact", k_sel_write, tiles, kVmBlock, 0, mask.as<const uint8_t>(), S,
            off.as<const unsigned long long>(), occ.as<uint32_t>());
  unsigned int of = 0;
  unsigned long long ngl = 0;
  {
    void *dst[2] = {&of, &ngl};
    const void *src[2] = {flag.p, off.as<unsigned long long>() + tiles};
    const uint64_t by[2] = {4, 8};
    PJ_TRY(d2h_many(ctx, 2, dst, src, by));
  }
  if (of) return fail(PJ_E_CAPACITY, "group-by: more groups than the table of %llu slots", (unsigned long long)S);
  const uint64_t ng = ngl;
  if (ng > max_groups)
    return fail(PJ_E_CAPACITY, "group-by produced %llu groups, maxInputSize %llu",
                (unsigned long long)ng, (unsigned long long)max_groups);
  if (ng)
    PJ_LAUNCH(ctx, "rel_groupby_emit", k_gb_emit, grid_for(ctx, ng, 256), 256, 0, a,
              slots.as<const uint32_t>(), occ.as<const uint32_t>(), ng, acc.as<const int64_t>(), S,
              R, out);
  PJ_TRY(sync(ctx));
  *n_groups = ng;
  return PJ_OK;
}

// Per-workgroup output regions of a pipeline and their compaction into the
// dense outputs (block order, so the output order is deterministic).
struct Segments {
  std::vector<std::unique_ptr<Scratch>> bufs;
  Scratch counts, offs;
  uint64_t grid = 0, region = 0;
  int init(pj_ctx *ctx, const std::vector<int32_t> &types, uint64_t g, uint64_t r) {
    grid = g;
    region = r;
    for (int32_t t : types) {
      bufs.emplace_back(new Scratch());
      PJ_TRY(bufs.back()->alloc(ctx, g * r * col_size(t)));
    }
    PJ_TRY(counts.alloc(ctx, g * 4));
    PJ_TRY(offs.alloc(ctx, (g + 2) * 8));  // offsets, the total, the largest count
    return PJ_OK;
  }
  int compact(pj_ctx *ctx, const std::vector<int32_t> &types, void *const *outs, uint64_t *n_out) {
    PJ_LAUNCH(ctx, "rel_scan", k_scan_counts, 1, 1024, 0, counts.as<const uint32_t>(), grid,
              offs.as<unsigned long long>(), offs.as<unsigned long long>() + grid + 1);
    // the total and the largest segment, read back together: the largest
    // segment sizes the copy grid (most segments hold a few hundred rows)
    unsigned long long tm[2] = {0, 0};
    PJ_TRY(d2h(ctx, tm, offs.as<unsigned long long>() + grid, 16));
    const unsigned long long total = tm[0];
    const uint32_t mx = (uint32_t)tm[1];
    if (total && !types.empty()) {
      if (types.size() > (size_t)kPipeMaxOut)
        return fail(PJ_E_UNSUPPORTED, "pipeline with more than %d outputs", kPipeMaxOut);
      SegCols sc{};
      for (size_t j = 0; j < types.size(); ++j) {
        sc.src[j] = bufs[j]->p;
        sc.dst[j] = outs[j];
        sc.esize[j] = (int32_t)col_size(types[j]);
      }
      const uint32_t gx = (uint32_t)std::min<uint64_t>(ceil_div(mx, 256), 16);
      hipLaunchKernelGGL(k_seg_compact, dim3(gx ? gx : 1, (uint32_t)grid, (uint32_t)types.size()),
                         dim3(256), 0, ctx->stream, sc, counts.as<const uint32_t>(),
                         offs.as<const unsigned long long>(), region);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return fail(PJ_E_HIP, "pipeline compaction: %s", hipGetErrorString(e));
    }
    PJ_TRY(sync(ctx));
    *n_out = total;
    return PJ_OK;
  }
};

int pipeline(pj_ctx *ctx, const PipeProg &p, uint64_t n, void *const *outs, uint64_t *n_out) {
  *n_out = 0;
  const int nlev = (int)p.tables.size() + 1;
  if (nlev > kPipeMaxLevels) return fail(PJ_E_UNSUPPORTED, "pipeline of more than %d joins", kPipeMaxLevels - 1);
  if ((int)p.out_t.size() > kPipeMaxOut) return fail(PJ_E_UNSUPPORTED, "pipeline with more than %d outputs", kPipeMaxOut);
  if (p.depth < 1 || p.depth > kMaxDepth || p.ins.empty()) return fail(PJ_E_ARG, "bad pipeline program");
  {
    int sp = 0;
    bool out_seen = false;
    for (auto &i : p.ins) {
      if (i.op == OP_COL && (i.a < 0 || i.a >= (int)p.cols.size() || p.clevel[i.a] < 0 || p.clevel[i.a] >= nlev))
        return fail(PJ_E_ARG, "pipeline column out of range");
      if (i.op == OP_PROBE && (i.a < 0 || i.a >= (int)p.tables.size())) return fail(PJ_E_ARG, "pipeline table out of range");
      if (i.op == OP_OUT && (i.a < 0 || i.a >= (int)p.out_t.size())) return fail(PJ_E_ARG, "pipeline output out of range");
      switch (i.op) {
        case OP_COL: case OP_CONST: ++sp; break;
        case OP_NEG: case OP_NOT: case OP_CAST: case OP_BLOOM: break;
        case OP_IF: sp -= 2; break;
        default: --sp; break;
      }
      if (i.op == OP_BLOOM && (i.a < 0 || i.a > 32 || !i.imm))
        return fail(PJ_E_ARG, "pipeline bloom filter without words or with log2 size %d", i.a);
      if (sp < 0 || sp > p.depth) return fail(PJ_E_ARG, "pipeline stack out of range");
      if ((i.op == OP_PROBE || i.op == OP_FILTER) && (sp != 0 || out_seen))
        return fail(PJ_E_ARG, "pipeline probes and filters need an empty stack, before the outputs");
      if (i.op == OP_OUT) out_seen = true;
    }
    if (sp != 0) return fail(PJ_E_ARG, "pipeline leaves values on the stack");
  }
  if (n >= (1ULL << 32)) return fail(PJ_E_CAPACITY, "pipeline source above 2^32-1 rows");
  if (!n) return PJ_OK;
  {
    ChainArgs ca{};
    std::vector<void *> sparse(p.out_t.size());
    if (as_chain(p, sparse.data(), &ca)) {
      for (size_t j = 0; j < p.tables.size(); ++j) ca.tables[j] = p.tables[j];
      const uint64_t tile = 256 * kChainRowsA;
      const uint64_t grid = std::min<uint64_t>(ceil_div(n, tile), (uint64_t)ctx->num_cus * 8);
      ca.n = n;
      ca.ntab = (int32_t)std::max<size_t>(1, p.tables.size());
      ca.region = ceil_div(ceil_div(n, tile), grid) * tile;
      Segments seg;
      PJ_TRY(seg.init(ctx, p.out_t, grid, ca.region));
      for (size_t j = 0; j < p.out_t.size(); ++j) ca.outs[j].out = seg.bufs[j]->p;
      ca.counts = seg.counts.as<uint32_t>();
      // the argument block carries a zeroed dummy key after ChainArgs
      Scratch dargs;
      PJ_TRY(dargs.alloc(ctx, sizeof(ChainArgs) + 256));
      const int32_t *dummy =
          reinterpret_cast<const int32_t *>(dargs.as<char>() + ((sizeof(ChainArgs) + 15) & ~15));
      for (int st = 0; st < kChainMaxStages; ++st) {
        const bool on = st < ca.nst && (ca.st[st].kind == 0 || ca.st[st].kind == 2) &&
                        ca.clevel[ca.st[st].col] == 0;
        ca.pf_on[st] = on ? 1 : 0;
        ca.pf[st] = on ? (const int32_t *)ca.cols[ca.st[st].col] : dummy;
      }
      std::vector<uint8_t> blob(sizeof(ChainArgs) + 256, 0);
      std::memcpy(blob.data(), &ca, sizeof(ChainArgs));
      PJ_TRY(h2d(ctx, dargs.p, blob.data(), blob.size()));
      const ChainArgs *da = dargs.as<const ChainArgs>();
      // ridl slots of every table for phase A's rows, and per wave a queue
      // of (row, stage-0 build row) pairs
      constexpr int RA = kChainRowsA, RB = kChainRowsB;
      const size_t lds = (size_t)ca.ntab * RA * 256 * 4 + 4 * (size_t)chain_queue_words(ca.nst) * 4;
      switch (ca.nst) {
        case 1: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<1, RA, RB>), grid, 256, lds, da); break;
        case 2: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<2, RA, RB>), grid, 256, lds, da); break;
        case 3: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<3, RA, RB>), grid, 256, lds, da); break;
        case 4: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<4, RA, RB>), grid, 256, lds, da); break;
        case 5: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<5, RA, RB>), grid, 256, lds, da); break;
        case 6: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<6, RA, RB>), grid, 256, lds, da); break;
        case 7: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<7, RA, RB>), grid, 256, lds, da); break;
        default: PJ_LAUNCH(ctx, "rel_pipeline_chain", (k_chain_q<8, RA, RB>), grid, 256, lds, da); break;
      }
      return seg.compact(ctx, p.out_t, outs, n_out);
    }
  }
  if (!n) return PJ_OK;
  // program, column tables and output tables in one device block
  const size_t ni = p.ins.size() * sizeof(Instr), nc = p.cols.size(), no = p.out_t.size();
  const uint64_t grid = std::min<uint64_t>(ceil_div(n, kPipeTile), (uint64_t)ctx->num_cus * 32);
  const uint64_t region = ceil_div(ceil_div(n, kPipeTile), grid) * kPipeTile;
  Segments seg;
  PJ_TRY(seg.init(ctx, p.out_t, grid, region));
  std::vector<void *> sp(no);
  for (size_t j = 0; j < no; ++j) sp[j] = seg.bufs[j]->p;
  const size_t bytes = ni + nc * 8 + nc * 4 + nc * 4 + no * 8 + no * 4 + 64;
  std::vector<char> h(bytes, 0);
  size_t off = 0;
  auto put = [&](const void *src, size_t len) {
    size_t at = off;
    if (len) memcpy(h.data() + off, src, len);
    off = (off + len + 7) & ~(size_t)7;
    return at;
  };
  const size_t o_ins = put(p.ins.data(), ni);
  const size_t o_cols = put(p.cols.data(), nc * 8);
  const size_t o_ct = put(p.ctype.data(), nc * 4);
  const size_t o_cl = put(p.clevel.data(), nc * 4);
  const size_t o_sp = put(sp.data(), no * 8);
  const size_t o_ot = put(p.out_t.data(), no * 4);
  Scratch dev;
  PJ_TRY(dev.alloc(ctx, off));
  PJ_TRY(h2d(ctx, dev.p, h.data(), off));
  PipeArgs a{};
  char *d = (char *)dev.p;
  a.ins = (const Instr *)(d + o_ins);
  a.cols = (const void *const *)(d + o_cols);
  a.ctype = (const int32_t *)(d + o_ct);
  a.clevel = (const int32_t *)(d + o_cl);
  a.outs = (void *const *)(d + o_sp);
  a.out_t = (const int32_t *)(d + o_ot);
  for (size_t j = 0; j < p.tables.size(); ++j) a.tables[j] = p.tables[j];
  a.n = n;
  a.counts = seg.counts.as<uint32_t>();
  a.region = region;
  a.nins = (int32_t)p.ins.size();
  a.nout = (int32_t)no;
  a.depth = p.depth;
  a.nlev = nlev;
  const size_t lds = (size_t)p.depth * kPipeTile * 8 + (size_t)nlev * kPipeTile * 4 + kPipeTile * 2;
  if (lds > 64 * 1024) return fail(PJ_E_UNSUPPORTED, "pipeline needs %zu bytes of LDS", lds);
  PJ_LAUNCH(ctx, "rel_pipeline", k_pipe, grid, kPipeWave, lds, a);
  return seg.compact(ctx, p.out_t, outs, n_out);
}

int hash_buckets(pj_ctx *ctx, const int32_t *k, uint64_t n, uint32_t buckets, uint32_t *out) {
  if (!buckets) return fail(PJ_E_ARG, "hash-rearrange with 0 buckets");
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "rel_hash_buckets", k_hash_buckets, grid_for(ctx, n, 256), 256, 0, k, n, buckets, out);
  return PJ_OK;
}

int hash_values(pj_ctx *ctx, const int32_t *k, uint64_t n, int64_t *out) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "rel_hash_values", k_hash_values, grid_for(ctx, n, 256), 256, 0, k, n, out);
  return PJ_OK;
}

}  // namespace rel
}  // namespace pj
