// Device sort of the plan operator `sort` (the reference's GPU sort,
// core/olap/lib/operators/gpu/gpu-sort.cpp with
// core/platform/lib/util/sort/gpu/sort-util.cu): a stable LSD sort of a row
// permutation.  For each sort expression, from the last to the first, every
// row's value (read through the current permutation) is turned into an
// unsigned key whose order is the column's order, and the permutation is
// stably radix-sorted by it (hipCUB's DeviceRadixSort on the context
// stream).  Stability makes the result lexicographic over the expressions,
// with ties kept in input order.
//
// Keys: 32-bit types (int32, bool, dictionary ranks) flip the sign bit into
// a 32-bit key; int64 flips it into a 64-bit key; FP64 (and float32, widened)
// takes its bit pattern, flipped whole when negative and with the sign bit
// set otherwise, after mapping -0.0 to +0.0 so that the two compare equal.
// DESC inverts the key; NONE directions are skipped.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "pj_internal.hpp"
#include "rel_api.hpp"

namespace pj {
namespace rel {

__global__ __launch_bounds__(256) void k_sort_key(const void *__restrict__ col, int t, int desc,
                                                  const int32_t *__restrict__ rank, uint32_t nrank,
                                                  const uint32_t *__restrict__ perm, uint64_t n,
                                                  unsigned long long *__restrict__ key) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const uint32_t r = perm[i];
    unsigned long long u;
    switch (t) {
      case I32: {
        int32_t v = ((const int32_t *)col)[r];
        if (rank) v = (uint32_t)v < nrank ? rank[v] : (int32_t)nrank;  // unknown codes last
        u = (uint32_t)v ^ 0x80000000u;
        if (desc) u = ~u & 0xffffffffULL;
        break;
      }
      case BOOL: {
        u = ((const uint8_t *)col)[r] != 0;
        if (desc) u = ~u & 0xffffffffULL;
        break;
      }
      case I64: {
        u = (unsigned long long)((const int64_t *)col)[r] ^ 0x8000000000000000ULL;
        if (desc) u = ~u;
        break;
      }
      default: {  // F64, F32
        double d = t == F32 ? (double)((const float *)col)[r] : ((const double *)col)[r];
        if (d == 0.0) d = 0.0;
        const unsigned long long b = __builtin_bit_cast(unsigned long long, d);
        u = (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
        if (desc) u = ~u;
        break;
      }
    }
    key[i] = u;
  }
}

__global__ __launch_bounds__(256) void k_sort_iota(uint32_t *__restrict__ p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) p[i] = (uint32_t)i;
}

int sort_perm(pj_ctx *ctx, const SortKey *keys, int nk, uint64_t n, uint32_t *perm) {
  if (nk < 0 || (nk && !keys) || (n && !perm)) return fail(PJ_E_ARG, "sort with invalid arguments");
  if (!n) return PJ_OK;
  if (n > (uint64_t)INT32_MAX) return fail(PJ_E_CAPACITY, "sort of more than 2^31-1 rows");
  for (int k = 0; k < nk; ++k) {
    if (!keys[k].col) return fail(PJ_E_ARG, "sort key %d without a column", k);
    if (keys[k].t != I32 && keys[k].t != I64 && keys[k].t != F64 && keys[k].t != F32 &&
        keys[k].t != BOOL)
      return fail(PJ_E_ARG, "sort key %d of unknown type %d", k, keys[k].t);
    if (keys[k].rank && (keys[k].t != I32 || keys[k].nrank > (uint32_t)INT32_MAX))
      return fail(PJ_E_ARG, "sort key %d: ranks need int32 codes", k);
  }
  uint64_t grid = ceil_div(n, (uint64_t)256);
  if (grid > (uint64_t)ctx->num_cus * 16) grid = (uint64_t)ctx->num_cus * 16;
  Scratch p1, k0, k1, tmp;
  PJ_TRY(p1.alloc(ctx, n * 4));
  PJ_TRY(k0.alloc(ctx, n * 8));
  PJ_TRY(k1.alloc(ctx, n * 8));
  PJ_LAUNCH(ctx, "sort_iota", k_sort_iota, grid, 256, 0, perm, n);
  hipcub::DoubleBuffer<uint32_t> vals(perm, p1.as<uint32_t>());
  size_t tmp_bytes = 0;
  for (int bits : {32, 64}) {
    hipcub::DoubleBuffer<unsigned long long> kb(k0.as<unsigned long long>(),
                                                 k1.as<unsigned long long>());
    size_t b = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b, kb, vals, (int)n, 0, bits,
                                           ctx->stream) != hipSuccess)
      return fail(PJ_E_HIP, "radix sort sizing failed");
    tmp_bytes = std::max(tmp_bytes, b);
  }
  PJ_TRY(tmp.alloc(ctx, tmp_bytes ? tmp_bytes : 1));
  for (int k = nk - 1; k >= 0; --k) {
    if (!keys[k].dir) continue;
    const int bits = (keys[k].t == I32 || keys[k].t == BOOL) ? 32 : 64;
    PJ_LAUNCH(ctx, "sort_key", k_sort_key, grid, 256, 0, keys[k].col, keys[k].t,
              keys[k].dir < 0 ? 1 : 0, keys[k].rank, keys[k].nrank, vals.Current(), n,
              k0.as<unsigned long long>());
    hipcub::DoubleBuffer<unsigned long long> kb(k0.as<unsigned long long>(),
                                                 k1.as<unsigned long long>());
    size_t bytes = tmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kb, vals, (int)n, 0, bits,
                                           ctx->stream) != hipSuccess)
      return fail(PJ_E_HIP, "radix sort failed");
  }
  if (vals.Current() != perm)
    PJ_HIP(hipMemcpyAsync(perm, vals.Current(), n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return sync(ctx);
}

}  // namespace rel
}  // namespace pj
