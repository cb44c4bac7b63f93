// tk_sort.hip -- device-wide sort and scan (hipcub) used off the hot path:
// dump ordering (jellyfish dump output is emitted in ascending key order so
// it diffs without a host sort) and offset scans for read packing and
// selection.
#include <hipcub/hipcub.hpp>

#include "tk_internal.h"

namespace tk {

int sort_pairs_u64_u32(tk_ctx* ctx, uint64_t* d_keys, uint32_t* d_vals, uint64_t n,
                       int key_bits) {
  if (n < 2) return TK_OK;
  uint64_t* k2 = nullptr;
  uint32_t* v2 = nullptr;
  size_t kb = n * sizeof(uint64_t), vb = n * sizeof(uint32_t);
  int rc = dev_alloc((void**)&k2, kb);
  if (rc) return rc;
  rc = dev_alloc((void**)&v2, vb);
  if (rc) { dev_free(k2, kb); return rc; }
  hipcub::DoubleBuffer<uint64_t> dk(d_keys, k2);
  hipcub::DoubleBuffer<uint32_t> dv(d_vals, v2);
  // hipcub takes a 32-bit item count; larger dumps are sorted per key range
  // by the caller.
  if (n > 0x7FFFFFFFull) {
    dev_free(k2, kb);
    dev_free(v2, vb);
    return fail(TK_ERR_ARG, "sort of %llu items exceeds the 2^31 single-call limit",
                (unsigned long long)n);
  }
  size_t temp = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp, dk, dv, (int)n, 0,
                                                    key_bits, ctx->stream);
  if (e != hipSuccess) { dev_free(k2, kb); dev_free(v2, vb); return hip_fail(e, "SortPairs(size)"); }
  void* tp = nullptr;
  rc = ensure_scratch(ctx, 3, temp, &tp);
  if (rc) { dev_free(k2, kb); dev_free(v2, vb); return rc; }
  hipEvent_t a;
  timing_begin(ctx, &a);
  e = hipcub::DeviceRadixSort::SortPairs(tp, temp, dk, dv, (int)n, 0, key_bits, ctx->stream);
  timing_end(ctx, "sort_pairs", a);
  if (e != hipSuccess) { dev_free(k2, kb); dev_free(v2, vb); return hip_fail(e, "SortPairs"); }
  if (dk.Current() != d_keys) {
    e = hipMemcpyAsync(d_keys, dk.Current(), kb, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_vals, dv.Current(), vb, hipMemcpyDeviceToDevice, ctx->stream);
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  dev_free(k2, kb);
  dev_free(v2, vb);
  if (e != hipSuccess) return hip_fail(e, "sort copy-back");
  if (e2 != hipSuccess) return hip_fail(e2, "sort sync");
  return TK_OK;
}

template <typename T>
static int scan_impl(tk_ctx* ctx, const T* d_in, T* d_out, uint64_t n, const char* name) {
  if (n == 0) return TK_OK;
  if (n > 0x7FFFFFFFull) return fail(TK_ERR_ARG, "scan of %llu items too large", (unsigned long long)n);
  size_t temp = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, temp, d_in, d_out, (int)n, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "ExclusiveSum(size)");
  void* tp = nullptr;
  int rc = ensure_scratch(ctx, 2, temp, &tp);
  if (rc) return rc;
  hipEvent_t a;
  timing_begin(ctx, &a);
  e = hipcub::DeviceScan::ExclusiveSum(tp, temp, d_in, d_out, (int)n, ctx->stream);
  timing_end(ctx, name, a);
  if (e != hipSuccess) return hip_fail(e, "ExclusiveSum");
  return TK_OK;
}

// Exclusive sum into caller-provided temporary storage (allocates nothing).
int exclusive_scan_u32_in(tk_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* temp,
                          size_t temp_bytes) {
  if (n == 0) return TK_OK;
  if (n > 0x7FFFFFFFull) return fail(TK_ERR_ARG, "scan of %llu items too large", (unsigned long long)n);
  size_t need = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, d_in, d_out, (int)n, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "ExclusiveSum(size)");
  if (need > temp_bytes)
    return fail(TK_ERR_ARG, "scan needs %zu bytes of temporary storage, %zu reserved", need, temp_bytes);
  hipEvent_t a;
  timing_begin(ctx, &a);
  e = hipcub::DeviceScan::ExclusiveSum(temp, need, d_in, d_out, (int)n, ctx->stream);
  timing_end(ctx, "scan_u32", a);
  if (e != hipSuccess) return hip_fail(e, "ExclusiveSum");
  return TK_OK;
}

int exclusive_scan_u64(tk_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n) {
  return scan_impl<uint64_t>(ctx, d_in, d_out, n, "scan_u64");
}
int exclusive_scan_u32(tk_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n) {
  return scan_impl<uint32_t>(ctx, d_in, d_out, n, "scan_u32");
}

}  // namespace tk
