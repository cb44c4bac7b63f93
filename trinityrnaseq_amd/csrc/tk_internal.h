// tk_internal.h -- host-side objects behind the C ABI in include/trinity_gpu.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <string>
#include <vector>

#include "../../include/trinity_gpu.h"
#include "tk_device.h"

namespace tk {

// Thread-local error message; returns code for `return fail(...)`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define TK_HIP(call)                                 \
  do {                                               \
    hipError_t _e = (call);                          \
    if (_e != hipSuccess) return ::tk::hip_fail(_e, #call); \
  } while (0)

// Device allocations made by this library (bytes), for admission control.
extern std::atomic<uint64_t> g_dev_bytes;
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p, size_t bytes);
// Per-thread device block cache: between cache_begin and cache_end, blocks
// this thread frees (up to limit_bytes in total) are kept for its later
// allocations.  Only for a thread that runs all its work on one stream and
// frees every object it creates on that thread.
void cache_begin(size_t limit_bytes);
void cache_end();
// Device bytes the region-resident count of this thread may use for its
// word streams (0 = 40% of the free device memory).
void resident_limit(size_t bytes);

struct PendingEvent {
  std::string name;
  hipEvent_t a, b;
};

struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace tk

struct tk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  std::vector<tk::PendingEvent> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, uint64_t>> times;
  uint32_t* d_flags = nullptr;     // 64 words of small device counters / error bits
  uint32_t* d_tickets = nullptr;   // per-XCD position counters of the per-read passes (after d_flags)
  uint32_t* h_flags = nullptr;     // pinned mirror
  tk::Scratch scratch[8];          // reusable device scratch buffers (4-7: region-resident count)
  int num_cu = 0;
  void* stage[2] = {nullptr, nullptr};        // pinned staging for pageable uploads
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  uint64_t count_words = 0;  // merged words of the last region-resident count
};

struct tk_reads {
  int device = 0;
  uint64_t n_reads = 0;
  uint64_t n_bases = 0;
  uint64_t n_words = 0;
  uint32_t max_len = 0;
  uint32_t* d_len = nullptr;
  uint64_t* d_word_off = nullptr;  // n_reads + 1
  uint64_t* d_codes = nullptr;     // n_words + 1
  uint32_t* d_nmask = nullptr;     // n_words + 1
  std::map<int, uint64_t> windows; // k -> total windows
  std::map<int, double> est;       // 2k + canonical -> distinct estimate of all reads
  uint32_t* d_order = nullptr;     // locality order of the table passes (n_reads)
  bool order_done = false;         // d_order built, or found not worth building
  uint64_t bytes = 0;
};

struct tk_table {
  int device = 0;
  int k = 25;
  int canonical = 1;
  double max_load = 0.5;
  uint64_t nb = 0;                 // buckets (regions * 2^rbits)
  uint32_t rbits = 0;              // log2(buckets per region)
  uint64_t* d_slots = nullptr;
  uint64_t ovf_cap = 0;            // pairs
  uint64_t* d_ovf = nullptr;
  uint64_t ovf_bound = 0;          // upper bound on keys that may need a side entry
  uint64_t n_keys = 0;             // distinct keys claimed (exact, from device counter)
  bool finalized = true;           // an empty table is trivially final
  bool zero_pending = false;       // slots not yet cleared (cleared before first use)
  uint64_t bytes = 0;
};

struct tk_ctable {
  int device = 0;
  int k = 25;
  int canonical = 1;
  uint64_t nbuckets = 0;           // 4 slots of 16 bytes per bucket
  uint64_t* d_slots = nullptr;     // pairs (tagged key, comp)
  uint64_t n_keys = 0;
  uint64_t bytes = 0;
};

namespace tk {
TableView table_view(const tk_table* t);
// Clear a pending table and fold its side table into final counts (every
// entry point that reads slots calls this first).
int table_finalize(tk_ctx* ctx, tk_table* t);
// Sorting/scan helpers implemented with hipcub in tk_sort.hip.
int sort_pairs_u64_u32(tk_ctx* ctx, uint64_t* d_keys, uint32_t* d_vals, uint64_t n,
                       int key_bits);
int exclusive_scan_u64(tk_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n);
int exclusive_scan_u32(tk_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n);
int exclusive_scan_u32_in(tk_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* temp,
                          size_t temp_bytes);
int ensure_scratch(tk_ctx* ctx, int slot, size_t bytes, void** p);
// The pair selection on device stats (tk_select_pairs_d) with its working
// memory given: consumes and idx (n words each), the scan's temporary storage
// (NULL: the context's) and 16 flag words; select_check validates the
// arguments.
int select_check(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n, int max_cov,
                 int max_cv, uint32_t max_nwin, uint8_t* dsel);
int select_run(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n, int max_cov,
               int min_cov, uint8_t* dsel, uint32_t* consumes, uint32_t* idx, void* scan_temp,
               size_t scan_bytes, uint32_t* flags, uint64_t* n_selected);
// Launch bracket for per-kernel timing on ctx->stream.
void timing_begin(tk_ctx* ctx, hipEvent_t* a);
void timing_end(tk_ctx* ctx, const char* name, hipEvent_t a);
int collect_timing(tk_ctx* ctx);
}  // namespace tk
