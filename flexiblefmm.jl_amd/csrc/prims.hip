// Device primitives on rocPRIM: the stable radix sort of the Morton keys, the
// exclusive scans and the compactions of the level loop.  rocPRIM is AMD's own
// tuned implementation of these primitives (decoupled look-back scans, onesweep
// radix sort) for CDNA; the tree, list, pass and evaluation kernels are this
// library's own.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "ffmm_internal.h"

namespace ffmm {

static void *scratch(Context &ctx, size_t bytes) {
  if (bytes > ctx.tmp.cap) ctx.tmp.resize(bytes + bytes / 4 + 4096);
  return ctx.tmp.p;
}

void sort_pairs_u64(Context &ctx, const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout,
                    int64_t n, int end_bit) {
  if (n <= 0) return;
  size_t bytes = 0;
  FFMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0, end_bit, ctx.stream));
  void *t = scratch(ctx, bytes);
  FFMM_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, (size_t)n, 0, end_bit, ctx.stream));
}

void sort_pairs_u32(Context &ctx, const uint32_t *kin, uint32_t *kout, const int32_t *vin, int32_t *vout,
                    int64_t n, int end_bit) {
  if (n <= 0) return;
  size_t bytes = 0;
  FFMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0, end_bit, ctx.stream));
  void *t = scratch(ctx, bytes);
  FFMM_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, (size_t)n, 0, end_bit, ctx.stream));
}

int64_t exclusive_scan_i32(Context &ctx, const int32_t *in, int32_t *out, int64_t n) {
  if (n <= 0) return 0;
  size_t bytes = 0;
  FFMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(),
                                   ctx.stream));
  void *t = scratch(ctx, bytes);
  FFMM_HIP(rocprim::exclusive_scan(t, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(), ctx.stream));
  int32_t *h = reinterpret_cast<int32_t *>(ctx.pinned);
  FFMM_HIP(hipMemcpyAsync(h, out + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx.stream));
  FFMM_HIP(hipMemcpyAsync(h + 1, in + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  return (int64_t)h[0] + (int64_t)h[1];
}

int64_t compact_flags_u8(Context &ctx, const uint8_t *flag, int32_t *idx, int64_t n) {
  if (n <= 0) return 0;
  rocprim::counting_iterator<int32_t> it(0);
  // the count lands after the primitive's own scratch
  size_t bytes = 0;
  FFMM_HIP(rocprim::select(nullptr, bytes, it, flag, idx, (int32_t *)nullptr, (size_t)n, ctx.stream));
  size_t off = (bytes + 255) & ~size_t(255);
  char *t = static_cast<char *>(scratch(ctx, off + 256));
  int32_t *dcount = reinterpret_cast<int32_t *>(t + off);
  FFMM_HIP(rocprim::select(t, bytes, it, flag, idx, dcount, (size_t)n, ctx.stream));
  int32_t *h = reinterpret_cast<int32_t *>(ctx.pinned);
  FFMM_HIP(hipMemcpyAsync(h, dcount, sizeof(int32_t), hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  return h[0];
}

}  // namespace ffmm
