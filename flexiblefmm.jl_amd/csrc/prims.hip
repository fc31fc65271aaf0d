// Device primitives on rocPRIM: the stable radix sort of the Morton keys, the
// exclusive scans and the compactions of the level loop.  rocPRIM is AMD's own
// tuned implementation of these primitives (decoupled look-back scans, onesweep
// radix sort) for CDNA; the tree, list, pass and evaluation kernels are this
// library's own.
#include <chrono>
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

// lane 0 of one wave: out[n-1] + in[n-1], the total of an exclusive scan, into pinned host memory
__global__ void k_scan_total(const int32_t *__restrict__ out, const int32_t *__restrict__ in, int64_t n,
                             int64_t *__restrict__ host) {
  if (threadIdx.x == 0) host[0] = (int64_t)out[n - 1] + (int64_t)in[n - 1];
}
__global__ void k_copy_count(const int32_t *__restrict__ dcount, int64_t *__restrict__ host) {
  if (threadIdx.x == 0) host[0] = dcount[0];
}
__global__ void k_set_value(int32_t *__restrict__ out, int32_t v) {
  if (threadIdx.x == 0) out[0] = v;
}

int64_t exclusive_scan_i32(Context &ctx, const int32_t *in, int32_t *out, int64_t n) {
  if (n <= 0) return 0;
  size_t bytes = 0;
  FFMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(),
                                   ctx.stream));
  void *t = scratch(ctx, bytes);
  FFMM_HIP(rocprim::exclusive_scan(t, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(), ctx.stream));
  // the total, stored by a kernel straight into the pinned readback area
  k_scan_total<<<1, 64, 0, ctx.stream>>>(out, in, n, ctx.pinned_dev);
  FFMM_HIP(hipGetLastError());
  ctx.sync();
  return ctx.pinned[0];
}

// lane 0 of one wave: the total of an exclusive scan into out[n] and into pinned slot host[0]
__global__ void k_scan_total_slot(int32_t *__restrict__ out, const int32_t *__restrict__ in, int64_t n,
                                  int64_t *__restrict__ host) {
  if (threadIdx.x == 0) {
    const int64_t t = n > 0 ? (int64_t)out[n - 1] + (int64_t)in[n - 1] : 0;
    out[n] = (int32_t)t;
    host[0] = t;
  }
}

// the exclusive scan of in[0, n) into out[0, n], out[n] = the total (32-bit), and the total
// (64-bit) into pinned readback slot `slot`: no synchronisation
void exclusive_scan_i32_async(Context &ctx, const int32_t *in, int32_t *out, int64_t n, int slot) {
  if (n > 0) {
    size_t bytes = 0;
    FFMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(),
                                     ctx.stream));
    void *t = scratch(ctx, bytes);
    FFMM_HIP(rocprim::exclusive_scan(t, bytes, in, out, int32_t(0), (size_t)n, rocprim::plus<int32_t>(),
                                     ctx.stream));
  }
  k_scan_total_slot<<<1, 64, 0, ctx.stream>>>(out, in, n, ctx.pinned_dev + slot);
  FFMM_HIP(hipGetLastError());
}

// The value a kernel on the stream stores into pinned slot `slot`, which the host set to -1
// before launching it: the host polls the slot (coherent host memory, written through by
// the kernel) instead of synchronising the stream, whose blocking wait returns some tens
// of microseconds after the GPU is done.  The kernel that stores the value runs after all
// earlier work of the stream.  Without the value after 2 s the stream is synchronised,
// which reports a failed launch.
static int64_t poll_slot(Context &ctx, int slot) {
  volatile int64_t *v = ctx.pinned + slot;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t k = 1;; k++) {
    const int64_t x = __atomic_load_n(v, __ATOMIC_ACQUIRE);
    if (x != -1) return x;
    if ((k & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
      ctx.sync();
      const int64_t y = __atomic_load_n(v, __ATOMIC_ACQUIRE);
      if (y == -1) fail(FFMM_ERR_DEVICE, "a count was not written back by the device");
      return y;
    }
  }
}
constexpr int POLL_SLOT = NPINNED - 1;   // the values that the host waits for

// Pinned slots read back one by one: arm_slots sets slots [first, first + n) to -1 before
// the kernels that store them are queued, and read_slot polls each slot until its value
// is there (the values read back this way are counts, offsets and keys, never -1).  Each
// value is awaited itself, so nothing rests on the order in which the device's stores to
// host memory from different kernels become visible.  Arm a slot only when every earlier
// store to it has been read back or the stream synchronised since.
void arm_slots(Context &ctx, int first, int n) {
  if (first < 0 || n < 0 || first + n > POLL_SLOT) fail(FFMM_ERR_STATE, "pinned slot out of range");
  for (int i = 0; i < n; i++) __atomic_store_n(ctx.pinned + first + i, (int64_t)-1, __ATOMIC_RELEASE);
}
int64_t read_slot(Context &ctx, int slot) { return poll_slot(ctx, slot); }

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
  __atomic_store_n(ctx.pinned + POLL_SLOT, (int64_t)-1, __ATOMIC_RELEASE);
  k_copy_count<<<1, 64, 0, ctx.stream>>>(dcount, ctx.pinned_dev + POLL_SLOT);
  FFMM_HIP(hipGetLastError());
  return poll_slot(ctx, POLL_SLOT);
}

__global__ void k_copy_count_slot(const int32_t *__restrict__ dcount, int64_t *__restrict__ host) {
  if (threadIdx.x == 0) host[0] = dcount[0];
}

// the same compaction with its count stored into pinned readback slot `slot` by the
// stream: no synchronisation, so that several can be read back at once
void compact_flags_u8_async(Context &ctx, const uint8_t *flag, int32_t *idx, int64_t n, int slot) {
  if (n <= 0) {
    set_pinned(ctx, slot, 0);
    return;
  }
  rocprim::counting_iterator<int32_t> it(0);
  size_t bytes = 0;
  FFMM_HIP(rocprim::select(nullptr, bytes, it, flag, idx, (int32_t *)nullptr, (size_t)n, ctx.stream));
  size_t off = (bytes + 255) & ~size_t(255);
  char *t = static_cast<char *>(scratch(ctx, off + 256));
  int32_t *dcount = reinterpret_cast<int32_t *>(t + off);
  FFMM_HIP(rocprim::select(t, bytes, it, flag, idx, dcount, (size_t)n, ctx.stream));
  k_copy_count_slot<<<1, 64, 0, ctx.stream>>>(dcount, ctx.pinned_dev + slot);
  FFMM_HIP(hipGetLastError());
}

__global__ void k_set_slot(int64_t *__restrict__ host, int64_t v) {
  if (threadIdx.x == 0) host[0] = v;
}
void set_pinned(Context &ctx, int slot, int64_t v) {
  k_set_slot<<<1, 64, 0, ctx.stream>>>(ctx.pinned_dev + slot, v);
  FFMM_HIP(hipGetLastError());
}

void set_i32(Context &ctx, int32_t *out, int32_t v) {
  k_set_value<<<1, 64, 0, ctx.stream>>>(out, v);
  FFMM_HIP(hipGetLastError());
}

}  // namespace ffmm
