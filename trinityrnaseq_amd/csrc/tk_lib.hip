// tk_lib.hip -- gfx950 implementation of the C ABI in include/trinity_gpu.h.
//
// Kernels (all integer work; no MFMA):
//   k_read_words / k_pack   ASCII reads -> 2-bit codes + non-ACGT mask
//   k_hll                   HyperLogLog registers of canonical k-mers (table sizing)
//   k_count                 jellyfish count: one thread per window, packed-slot table
//   k_stats                 fastaToKmerCoverageStats: one wavefront per read,
//                           LDS-staged window values, radix-select median
//   k_select_*              nbkc merge + normalize with a jump-ahead drand48
//   k_ctable_build / k_vote ReadsToTranscripts
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <unordered_map>

#include "tk_count_agg.h"
#include "tk_internal.h"
#include "tk_perread.h"

namespace tk {

static thread_local std::string g_err;
std::atomic<uint64_t> g_dev_bytes{0};

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? TK_ERR_OOM : TK_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

// Device blocks of one host thread that issues all its work on one stream.
// A freed block goes back to the thread's free list instead of hipFree
// (which waits for the whole device), and the next allocation of a similar
// size takes it: stream order makes the reuse safe without any wait.
struct BlockCache {
  std::multimap<size_t, void*> free_blocks;  // capacity -> block
  std::unordered_map<void*, size_t> owned;   // every block of this cache -> capacity
  size_t cached = 0;                         // bytes on the free list
  size_t limit = 0;
};
static thread_local BlockCache* t_cache = nullptr;

static size_t cache_round(size_t bytes) {
  const size_t g = bytes >= (64u << 20) ? (size_t)(2u << 20) : (size_t)4096;
  return (bytes + g - 1) / g * g;
}

void cache_begin(size_t limit_bytes) {
  if (!t_cache) t_cache = new BlockCache();
  t_cache->limit = limit_bytes;
}

// Device bytes the region-resident count may hold for its streams (0 = 40%
// of the free memory).  A server worker sets its own share.
static thread_local uint64_t t_resident_limit = 0;
void resident_limit(size_t bytes) { t_resident_limit = bytes; }

void cache_end() {
  BlockCache* c = t_cache;
  if (!c) return;
  t_cache = nullptr;
  for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
  // blocks still handed out are released by their own dev_free (plain hipFree)
  delete c;
}

// TK_DEVICE_ALLOC_LIMIT=n: an allocation that would take the library's
// device bytes in use past n fails with TK_ERR_OOM, as an exhausted device
// would (the out-of-memory paths are tested with it).
static uint64_t alloc_limit() {
  const char* e = getenv("TK_DEVICE_ALLOC_LIMIT");
  return e ? strtoull(e, nullptr, 10) : 0;
}

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  if (const uint64_t lim = alloc_limit()) {
    if (g_dev_bytes.load() + bytes > lim)
      return fail(TK_ERR_OOM, "device allocation of %zu bytes past the limit of %llu bytes (TK_DEVICE_ALLOC_LIMIT)",
                  bytes, (unsigned long long)lim);
  }
  size_t cap = bytes;
  if (BlockCache* c = t_cache) {
    auto it = c->free_blocks.lower_bound(bytes);
    if (it != c->free_blocks.end() && it->first <= 2 * bytes + (1u << 20)) {
      *p = it->second;
      c->cached -= it->first;
      c->free_blocks.erase(it);
      g_dev_bytes += bytes;
      return TK_OK;
    }
    cap = cache_round(bytes);
  }
  hipError_t e = hipMalloc(p, cap);
  if (e != hipSuccess && t_cache && !t_cache->free_blocks.empty()) {
    // give the cached blocks back and try once more
    (void)hipGetLastError();
    for (auto& kv : t_cache->free_blocks) {
      (void)hipFree(kv.second);
      t_cache->owned.erase(kv.second);
    }
    t_cache->free_blocks.clear();
    t_cache->cached = 0;
    e = hipMalloc(p, cap);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? TK_ERR_OOM : TK_ERR_HIP,
                "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  if (t_cache) t_cache->owned[*p] = cap;
  g_dev_bytes += bytes;
  return TK_OK;
}

void dev_free(void* p, size_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 256;
  g_dev_bytes -= bytes;
  if (BlockCache* c = t_cache) {
    auto it = c->owned.find(p);
    if (it != c->owned.end()) {
      const size_t cap = it->second;
      if (cap <= c->limit) {
        c->free_blocks.emplace(cap, p);
        c->cached += cap;
        // over the limit: release the largest free blocks
        while (c->cached > c->limit && !c->free_blocks.empty()) {
          auto last = std::prev(c->free_blocks.end());
          (void)hipFree(last->second);
          c->cached -= last->first;
          c->owned.erase(last->second);
          c->free_blocks.erase(last);
        }
        return;
      }
      c->owned.erase(it);
    }
  }
  (void)hipFree(p);
}

int ensure_scratch(tk_ctx* ctx, int slot, size_t bytes, void** p) {
  Scratch& s = ctx->scratch[slot];
  if (s.bytes < bytes) {
    if (s.p) {
      (void)hipStreamSynchronize(ctx->stream);
      dev_free(s.p, s.bytes);
      s.p = nullptr;
      s.bytes = 0;
    }
    size_t nb = std::max<size_t>(bytes, 1 << 20);
    int rc = dev_alloc(&s.p, nb);
    if (rc) return rc;
    s.bytes = nb;
  }
  *p = s.p;
  return TK_OK;
}

void timing_begin(tk_ctx* ctx, hipEvent_t* a) {
  *a = nullptr;
  if (!ctx->timing) return;
  if (ctx->event_pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ctx->event_pool.push_back(e);
  }
  *a = ctx->event_pool.back();
  ctx->event_pool.pop_back();
  (void)hipEventRecord(*a, ctx->stream);
}

void timing_end(tk_ctx* ctx, const char* name, hipEvent_t a) {
  if (!ctx->timing || !a) return;
  hipEvent_t b;
  if (ctx->event_pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  }
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.push_back(PendingEvent{name, a, b});
}

int collect_timing(tk_ctx* ctx) {
  if (ctx->pending.empty()) return TK_OK;
  TK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& slot = ctx->times[p.name];
      slot.first += ms;
      slot.second += 1;
    }
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
  return TK_OK;
}

#define TK_LAUNCH(ctx, name, grid, block, shmem, kernel, ...)            \
  do {                                                                   \
    hipEvent_t _a;                                                       \
    ::tk::timing_begin((ctx), &_a);                                      \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem), (ctx)->stream, \
                       __VA_ARGS__);                                     \
    ::tk::timing_end((ctx), (name), _a);                                 \
    hipError_t _le = hipGetLastError();                                  \
    if (_le != hipSuccess) return ::tk::hip_fail(_le, "launch " name);   \
  } while (0)

static inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static inline uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

static int set_device(int dev) {
  TK_HIP(hipSetDevice(dev));
  return TK_OK;
}

static ReadsView reads_view(const tk_reads* r) {
  ReadsView v;
  v.codes = r->d_codes;
  v.nmask = r->d_nmask;
  v.word_off = r->d_word_off;
  v.len = r->d_len;
  v.n_reads = r->n_reads;
  v.order = r->d_order;
  return v;
}

TableView table_view(const tk_table* t) {
  TableView v;
  v.slots = t->d_slots;
  v.boff = 0;
  v.nb = t->nb;
  v.rbits = t->rbits;
  v.rmask = (1ull << t->rbits) - 1;
  v.occ = 1ull << (2 * t->k);
  v.kmo = (v.occ << 1) - 1;
  v.cshift = 2 * t->k + 1;
  v.cbits = 63 - 2 * t->k;
  v.cmax = (1ull << v.cbits) - 1;
  v.ovf = t->d_ovf;
  v.omask = t->ovf_cap - 1;
  v.max_probe = (uint32_t)std::min<uint64_t>(1ull << t->rbits, 4096);
  return v;
}

// =====================================================================
// Read packing
// =====================================================================
// Lengths and word counts of every read; one atomicMax per block for the
// longest read.  Grid-stride (the grid is sized to the device, not to n).
__global__ __launch_bounds__(256) void k_read_words(const uint64_t* __restrict__ off, uint64_t n,
                                                    uint32_t* __restrict__ len,
                                                    uint64_t* __restrict__ nwords,
                                                    uint32_t* __restrict__ maxlen) {
  __shared__ uint32_t s_max[4];
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < n) {
      const uint64_t l64 = off[i + 1] - off[i];
      const uint32_t L = l64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)l64;
      len[i] = L;
      nwords[i] = (l64 + 31) / 32;
      m = L > m ? L : m;
    } else {
      nwords[i] = 0;
    }
  }
  m = wave_max_u32(m);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = s_max[i] > m ? s_max[i] : m;
    if (m) atomicMax(maxlen, m);
  }
}

__device__ __forceinline__ uint32_t base_code(uint8_t ch, uint32_t& bad) {
  // A/a=0 C/c=1 G/g=2 T/t=3 ; everything else invalid
  const uint8_t u = ch & 0xDF;  // upper-case letters
  uint32_t c = 0;
  bad = 0;
  if (u == 'A') c = 0;
  else if (u == 'C') c = 1;
  else if (u == 'G') c = 2;
  else if (u == 'T') c = 3;
  else bad = 1;
  if (ch < 'A') bad = 1;  // digits/punctuation whose low bits alias letters
  return c;
}

// blk_read[b] = read holding word 256*b (thread per read; reads are
// ordered, so every 256-word block start lies in exactly one read).
__global__ void k_block_starts(const uint64_t* __restrict__ word_off, uint64_t n,
                               uint64_t* __restrict__ blk_read) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t a = word_off[r], e = word_off[r + 1];
    for (uint64_t w = (a + 255) & ~255ull; w < e; w += 256) blk_read[w >> 8] = r;
  }
}

// Bytes of x that are zero get 0x80, the others 0 (exact per byte).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t v) {
  const uint32_t t = (v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
  return ~(t | v) & 0x80808080u;
}

// Four ASCII bases (byte 0 first) -> 8 code bits (first base in the top two;
// 0 for an invalid base) and 4 mask bits (first base in bit 3; set for
// anything but A/C/G/T/a/c/g/t).
// Codes: ((ch >> 1) ^ (ch >> 2)) & 3 maps A/a 0, C/c 1, G/g 2, T/t 3.
__device__ __forceinline__ void pack4(uint32_t x, uint32_t& code8, uint32_t& mask4) {
  const uint32_t u = x & 0xDFDFDFDFu;
  const uint32_t ok = zero_bytes(u ^ 0x41414141u) | zero_bytes(u ^ 0x43434343u) |
                      zero_bytes(u ^ 0x47474747u) | zero_bytes(u ^ 0x54545454u);
  const uint32_t y = ((x >> 1) ^ (x >> 2)) & ((ok >> 7) * 3u);  // invalid bases code 0
  code8 = ((y & 3u) << 6) | (((y >> 8) & 3u) << 4) | (((y >> 16) & 3u) << 2) | ((y >> 24) & 3u);
  const uint32_t bad = ~ok & 0x80808080u;
  mask4 = ((bad >> 4) & 8u) | ((bad >> 13) & 4u) | ((bad >> 22) & 2u) | (bad >> 31);
}

// One thread per output word: 32 bases -> 64-bit code word + 32-bit mask.
// The thread reads its bases straight from the ASCII with aligned 4-byte
// loads (only those holding at least one of its bytes, so nothing outside
// the buffer is touched) and converts four at a time with SWAR byte tests.
// The word offsets of the block's reads are staged in LDS so each thread
// finds its read with a short binary search.
constexpr int kPackReadsLds = 512;

__global__ __launch_bounds__(256) void k_pack(const char* __restrict__ seq,
                                              const uint64_t* __restrict__ off,
                                              const uint64_t* __restrict__ word_off,
                                              const uint32_t* __restrict__ len,
                                              const uint64_t* __restrict__ blk_read,
                                              uint64_t n_reads, uint64_t n_words, uint64_t n_bytes,
                                              uint64_t* __restrict__ codes,
                                              uint32_t* __restrict__ nmask) {
  __shared__ uint64_t s_woff[kPackReadsLds + 1];
  const uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x;
  if (w0 >= n_words) return;
  const uint64_t wl = std::min<uint64_t>(w0 + blockDim.x, n_words) - 1;
  // first read of this block and of the next bound the reads of words w0..wl
  const uint64_t rf = blk_read[blockIdx.x];
  const uint64_t rnext = (wl + 1 < n_words) ? blk_read[blockIdx.x + 1] : n_reads - 1;
  uint64_t rl = rnext;
  while (rl > rf && word_off[rl] > wl) rl--;  // exact read of word wl (skips empty reads)
  const bool lds = rl - rf + 2 <= kPackReadsLds + 1;
  if (lds)
    for (uint64_t i = threadIdx.x; i < rl - rf + 2; i += blockDim.x) s_woff[i] = word_off[rf + i];
  __syncthreads();
  const uint64_t w = w0 + threadIdx.x;
  if (w > wl) return;
  uint64_t lo = rf, hi = rl + 1;  // largest r in [rf, rl] with word_off[r] <= w
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint64_t wm = lds ? s_woff[mid - rf] : word_off[mid];
    if (wm <= w) lo = mid; else hi = mid;
  }
  const uint64_t r = lo;
  const uint64_t j = w - (lds ? s_woff[r - rf] : word_off[r]);
  const uint32_t L = len[r];
  const uint64_t b0 = j * 32;
  const uint32_t nb = (uint32_t)std::min<uint64_t>(32, L - b0);
  const uintptr_t p = reinterpret_cast<uintptr_t>(seq) + off[r] + b0;
  const uintptr_t pa = p & ~(uintptr_t)3;
  const uint32_t sh = (uint32_t)(p & 3);
  const uintptr_t pend = p + nb;  // <= seq + n_bytes
  uint32_t wv[9];
#pragma unroll
  for (int i = 0; i < 9; i++)
    wv[i] = (pa + 4 * i < pend) ? *reinterpret_cast<const uint32_t*>(pa + 4 * i) : 0u;
  uint64_t code = 0;
  uint32_t mask = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t d = __builtin_amdgcn_alignbyte(wv[i + 1], wv[i], sh);
    const int keep = (int)nb - 4 * i;  // bytes of this dword inside the read
    if (keep < 4) d = keep <= 0 ? 0u : (d & ((1u << (8 * keep)) - 1u));
    uint32_t c8, m4;
    pack4(d, c8, m4);
    code |= (uint64_t)c8 << (56 - 8 * i);
    mask |= m4 << (28 - 4 * i);
  }
  codes[w] = code;
  nmask[w] = mask;
  (void)n_bytes;
}

__global__ void k_windows(const uint32_t* __restrict__ len, uint64_t n, int k,
                          unsigned long long* __restrict__ total) {
  uint64_t s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t L = len[i];
    s += L >= (uint32_t)k ? L - k + 1 : 0;
  }
  block_sum_add(s, total);
}

static int reads_build(tk_ctx* ctx, const char* d_seq, const uint64_t* d_off, uint64_t n,
                       uint64_t n_bases, tk_reads** out) {
  tk_reads* r = new tk_reads();
  r->device = ctx->device;
  r->n_reads = n;
  r->n_bases = n_bases;
  int rc = TK_OK;
  uint64_t* d_nw = nullptr;
  size_t nwb = (n + 1) * sizeof(uint64_t);
  do {
    if ((rc = dev_alloc((void**)&r->d_len, std::max<uint64_t>(n, 1) * 4))) break;
    if ((rc = dev_alloc((void**)&r->d_word_off, nwb))) break;
    if ((rc = dev_alloc((void**)&d_nw, nwb))) break;
    hipError_t e = hipMemsetAsync(ctx->d_flags, 0, 64 * 4, ctx->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "memset flags"); break; }
    {
      uint64_t grid = std::min<uint64_t>(ceil_div(n + 1, 256), (uint64_t)std::max(ctx->num_cu, 1) * 8);
      hipEvent_t a;
      timing_begin(ctx, &a);
      hipLaunchKernelGGL(k_read_words, dim3(grid), dim3(256), 0, ctx->stream, d_off, n,
                         r->d_len, d_nw, ctx->d_flags);
      timing_end(ctx, "read_words", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_read_words"); break; }
    }
    if ((rc = exclusive_scan_u64(ctx, d_nw, r->d_word_off, n + 1))) break;
    uint64_t nwords = 0;
    e = hipMemcpyAsync(&nwords, r->d_word_off + n, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->h_flags, ctx->d_flags, 64 * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "read words"); break; }
    r->max_len = ctx->h_flags[0];
    r->n_words = nwords;
    if ((rc = dev_alloc((void**)&r->d_codes, (nwords + 1) * 8))) break;
    if ((rc = dev_alloc((void**)&r->d_nmask, (nwords + 1) * 4))) break;
    e = hipMemsetAsync(r->d_codes + nwords, 0, 8, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(r->d_nmask + nwords, 0xFF, 4, ctx->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "pad"); break; }
    if (nwords) {
      const uint64_t nblk = ceil_div(nwords, 256);
      void* bp = nullptr;
      if ((rc = ensure_scratch(ctx, 1, nblk * 8, &bp))) break;
      uint64_t* blk_read = (uint64_t*)bp;
      hipEvent_t a;
      timing_begin(ctx, &a);
      hipLaunchKernelGGL(k_block_starts, dim3(std::min<uint64_t>(ceil_div(n, 256), 65536)), dim3(256), 0,
                         ctx->stream, r->d_word_off, n, blk_read);
      hipLaunchKernelGGL(k_pack, dim3(nblk), dim3(256), 0, ctx->stream,
                         d_seq, d_off, r->d_word_off, r->d_len, blk_read, n, nwords, n_bases,
                         r->d_codes, r->d_nmask);
      timing_end(ctx, "pack", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_pack"); break; }
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "pack sync"); break; }
  } while (0);
  dev_free(d_nw, nwb);
  r->bytes = std::max<uint64_t>(n, 1) * 4 + nwb + (r->n_words + 1) * 12;
  if (rc) {
    tk_reads_free(r);
    return rc;
  }
  *out = r;
  return TK_OK;
}


__global__ __launch_bounds__(256) void k_minimizer(ReadsView rv, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ idx) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rv.n_reads;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t L = rv.len[r];
    const uint64_t wb = rv.word_off[r];
    MinKey mk;
    for (uint32_t w = 0; 32 * w < L; w++) {
      const uint64_t code = rv.codes[wb + w];
      const uint32_t nb = L - 32 * w < 32u ? L - 32 * w : 32u;
      for (uint32_t j = 0; j < nb; j++) mk.push((uint32_t)(code >> (62 - 2 * j)) & 3u, 32 * w + j);
    }
    keys[r] = mk.key(L);
    idx[r] = (uint32_t)r;
  }
}

// Which table passes take reads in the locality order: TK_ORDER=0 none,
// TK_ORDER=stats the scoring pass only, TK_ORDER=1 (default) count and
// scoring.
enum { kOrderNone = 0, kOrderStats = 1, kOrderAll = 2 };
static int order_mode() {
  const char* e = getenv("TK_ORDER");
  if (!e) return kOrderAll;
  if (e[0] == '0') return kOrderNone;
  if (e[0] == 's') return kOrderStats;
  return kOrderAll;
}

// Read sets below this many reads stay in input order (TK_ORDER_MIN_READS).
static uint64_t order_min_reads() {
  const char* e = getenv("TK_ORDER_MIN_READS");
  return e ? strtoull(e, nullptr, 10) : (1ull << 16);
}

// Build the locality order of r once (read set large enough to be out of
// cache); results never depend on it, only the memory access pattern does.
static bool order_wanted(const tk_reads* r) {
  return !r->order_done && order_mode() != kOrderNone && r->n_reads >= order_min_reads() &&
         r->n_reads <= 0x7FFFFFFFull;
}

// Sort the per-read keys into r's locality order (takes ownership of both
// buffers: keys is freed, order becomes r->d_order).
static int order_finish(tk_ctx* ctx, tk_reads* r, uint64_t* keys, uint32_t* order) {
  const uint64_t n = r->n_reads;
  int rc = sort_pairs_u64_u32(ctx, keys, order, n, kOrderKeyBits);
  dev_free(keys, n * 8);
  if (rc) { dev_free(order, n * 4); return rc; }
  r->d_order = order;
  r->bytes += n * 4;
  r->order_done = true;
  return TK_OK;
}

static int reads_order(tk_ctx* ctx, const tk_reads* rc_) {
  tk_reads* r = const_cast<tk_reads*>(rc_);
  if (r->order_done) return TK_OK;
  if (!order_wanted(r)) {
    r->order_done = true;
    return TK_OK;
  }
  const uint64_t n = r->n_reads;
  uint64_t* keys = nullptr;
  uint32_t* order = nullptr;
  int rc = dev_alloc((void**)&keys, n * 8);
  if (rc) return rc;
  if ((rc = dev_alloc((void**)&order, n * 4))) { dev_free(keys, n * 8); return rc; }
  const uint64_t grid = std::min<uint64_t>(ceil_div(n, 256), (uint64_t)std::max(ctx->num_cu, 1) * 16);
  hipEvent_t a;
  timing_begin(ctx, &a);
  hipLaunchKernelGGL(k_minimizer, dim3(grid), dim3(256), 0, ctx->stream, reads_view(r), keys, order);
  timing_end(ctx, "minimizer", a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    dev_free(keys, n * 8);
    dev_free(order, n * 4);
    return hip_fail(e, "k_minimizer");
  }
  return order_finish(ctx, r, keys, order);
}

static int reads_windows(tk_ctx* ctx, const tk_reads* rc_, int k, uint64_t* total) {
  tk_reads* r = const_cast<tk_reads*>(rc_);
  auto it = r->windows.find(k);
  if (it != r->windows.end()) { *total = it->second; return TK_OK; }
  unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->d_flags + 8);
  TK_HIP(hipMemsetAsync(d, 0, 8, ctx->stream));
  if (r->n_reads) {
    uint64_t grid = std::min<uint64_t>(ceil_div(r->n_reads, 256), 4096);
    TK_LAUNCH(ctx, "windows", grid, 256, 0, k_windows, r->d_len, r->n_reads, k, d);
  }
  uint64_t h = 0;
  TK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  TK_HIP(hipStreamSynchronize(ctx->stream));
  r->windows[k] = h;
  *total = h;
  return TK_OK;
}

// =====================================================================
// HyperLogLog sizing pass
// =====================================================================
constexpr int kHllBits = 12;
constexpr int kHllRegs = 1 << kHllBits;

// HyperLogLog registers of the canonical k-mers.  Windows whose hash has its
// top `sbits` bits zero (a 2^-sbits sample chosen by key, so every
// occurrence of a sampled key is sampled) update the registers; the estimate
// is scaled by 2^sbits.  (Sampling by a multiplicative hash of the key
// instead was no faster and biased the estimate by +5.5%.)  Thread per read with rolling forward and
// reverse-complement words (a window ends at base i once `run` valid bases
// end there), so each base is read once instead of once per window;
// block-private registers in LDS merged with one atomicMax per register.
// When `keys`/`idx` are given, the same pass writes each read's locality
// sort key exactly as k_minimizer does.
__global__ __launch_bounds__(256) void k_hll(ReadsView rv, uint64_t r_begin, uint64_t r_end, int k,
                                                  int canonical, uint32_t sbits, uint32_t* __restrict__ regs,
                                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  __shared__ uint32_t s_reg[kHllRegs];
  for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x) s_reg[i] = 0;
  __syncthreads();
  const uint64_t kmask = (1ull << (2 * k)) - 1;  // k <= 31
  const int rsh = 2 * k - 2;
  for (uint64_t r = r_begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < r_end;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t L = rv.len[r];
    const uint64_t wb = rv.word_off[r];
    uint64_t fwd = 0, rc = 0;
    uint32_t run = 0;
    MinKey mk;
    for (uint32_t w = 0; 32 * w < L; w++) {
      const uint64_t code = rv.codes[wb + w];
      const uint32_t msk = rv.nmask[wb + w];
      const uint32_t nb = L - 32 * w < 32u ? L - 32 * w : 32u;
      for (uint32_t j = 0; j < nb; j++) {
        const uint32_t b = (uint32_t)(code >> (62 - 2 * j)) & 3u;
        fwd = ((fwd << 2) | b) & kmask;
        rc = (rc >> 2) | ((uint64_t)(3u - b) << rsh);
        run = ((msk >> (31 - j)) & 1u) ? 0u : run + 1;
        if (run >= (uint32_t)k) {
          const uint64_t key = (canonical && rc < fwd) ? rc : fwd;
          const uint64_t h0 = mix64(key ^ 0x5bd1e9955bd1e995ULL);
          if (!(sbits && (h0 >> (64 - sbits)))) {
            const uint64_t h = h0 << sbits;
            const uint32_t ix = (uint32_t)(h >> (64 - kHllBits));
            const uint64_t rest = (h << kHllBits) | (1ull << (kHllBits - 1 + sbits));
            atomicMax(&s_reg[ix], (uint32_t)__clzll(rest) + 1);
          }
        }
        if (keys) mk.push(b, 32 * w + j);
      }
    }
    if (keys) {
      keys[r] = mk.key(L);
      idx[r] = (uint32_t)r;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x)
    if (s_reg[i]) atomicMax(&regs[i], s_reg[i]);
}

// =====================================================================
// Counting
// =====================================================================

// One thread per k-window of a block of 64 consecutive reads.  flags[0] gets
// error bits, flags[2..3] (u64) the number of newly claimed slots.  Four
// lanes of 16 bytes probe each key's bucket.
__global__ __launch_bounds__(256) void k_count(ReadsView rv, uint64_t r_begin,
                                               uint64_t r_end, int k, int canonical,
                                               TableView t, uint32_t* __restrict__ flags) {
  __shared__ uint32_t s_off[kReadsPerBlock + 1];
  __shared__ uint64_t s_wb[kReadsPerBlock];
  __shared__ uint32_t s_new;
  const uint64_t r0 = r_begin + xcd_block(blockIdx.x, gridDim.x) * kReadsPerBlock;
  if (r0 >= r_end) return;
  const int nr = (int)std::min<uint64_t>(kReadsPerBlock, r_end - r0);
  const int tid = threadIdx.x;
  if (tid < kReadsPerBlock) {
    uint32_t nw = 0;
    if (tid < nr) {
      const uint64_t rr = read_at(rv, r0 + tid);
      const uint32_t L = rv.len[rr];
      nw = L >= (uint32_t)k ? L - k + 1 : 0;
      s_wb[tid] = rv.word_off[rr];
    }
    uint32_t x = nw;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o);
      if (tid >= o) x += y;
    }
    s_off[tid + 1] = x;
    if (tid == 0) { s_off[0] = 0; s_new = 0; }
  }
  __syncthreads();
  const uint32_t total = s_off[nr];
  uint32_t mynew = 0;
  // block-uniform trip count: every lane of a wave takes part in the
  // cooperative rounds, lanes past the end with want = false
  for (uint32_t w0 = 0; w0 < total; w0 += blockDim.x) {
    const uint32_t w = w0 + tid;
    uint64_t fwd = 0;
    bool valid = false;
    if (w < total) {
      int lo = 0, hi = nr;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_off[mid] <= w) lo = mid; else hi = mid;
      }
      window_at(rv, s_wb[lo], w - s_off[lo], k, fwd, valid);
    }
    const uint64_t key = valid ? canon_key(fwd, k, canonical) : 0;
    bool lost = false;
    mynew += coop4_add(t, key, valid, tid & 63, &flags[0], lost);
    if (lost) atomicOr(&flags[0], 1u);
  }
  mynew = wave_sum_u32(mynew);
  if ((tid & 63) == 0 && mynew) atomicAdd(&s_new, mynew);
  __syncthreads();
  if (tid == 0 && s_new)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)s_new);
}

// The block-merging count pass, k_count_agg: tk_count_agg.h.

// ---------------------------------------------------------------------
// Region-resident counting (reads in the locality order, k <= 25).  The
// block merge above emits its merged words (tagged key | occurrences <<
// (2k+1)) to a stream; two counting-sort passes group the stream by table
// region (coarse bins of 2^fs regions, then single regions); then one
// workgroup per region loads the region's buckets into LDS, adds the
// region's words there with LDS atomics and writes the region back.  Each
// table line is read and written once per pass instead of taking one
// memory-side atomic per merged word, so the pass is bound by streaming
// HBM traffic, not by the device's atomic rate.
// ---------------------------------------------------------------------
constexpr uint32_t kResRegionBitsMax = 10;  // 1024 buckets = 64 KB of LDS
constexpr uint64_t kResSlice = 65536;       // words per block of a sort pass
constexpr uint32_t kResMaxBins = 2048;      // bins per sort pass

// A slice of a sort pass: words [begin, end) of the input, whose bin b
// counts into hist[hbase + b * m + j] (m slices in the slice's group).
struct ResSlice {
  uint64_t begin, end, hbase;
  uint32_t m, j;
};

__device__ __forceinline__ uint32_t res_region(const TableView& t, uint64_t w) {
  return (uint32_t)((bucket0(t, w & (t.occ - 1)) - t.boff) >> t.rbits);
}

// Bin of every nonzero word: (region >> shift) & mask, bins < nbins.
__global__ __launch_bounds__(256) void k_res_hist(const uint64_t* __restrict__ in,
                                                  const ResSlice* __restrict__ sl, TableView t,
                                                  uint32_t shift, uint32_t mask, uint32_t nbins,
                                                  uint64_t* __restrict__ hist) {
  __shared__ uint32_t s_cnt[kResMaxBins];
  const ResSlice d = sl[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
  for (uint64_t i0 = d.begin; i0 < d.end; i0 += 4 * blockDim.x) {
    uint64_t w[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint64_t i = i0 + u * blockDim.x + threadIdx.x;
      w[u] = i < d.end ? in[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (w[u]) atomicAdd(&s_cnt[(res_region(t, w[u]) >> shift) & mask], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) hist[d.hbase + (uint64_t)i * d.m + d.j] = s_cnt[i];
}

// Scatter every nonzero word to its bin's range: out[base[bin entry] + rank].
// Each tile of kResTile words is counting-sorted by bin in LDS first, so
// the stores leave in runs of consecutive addresses (one run per bin and
// tile) instead of one scattered 8-B store per word.
// A tile of 8,192 words keeps no per-word bin in LDS (the write loop hashes
// each word again), so that two blocks still fit one CU.
static size_t res_scatter_lds(uint32_t nbins, uint32_t tile) {
  return (size_t)tile * (tile > 4096 ? 8 : 10) + (size_t)nbins * 16;
}

// Exclusive scan of s_cnt[0, nbins) into s_loc (nbins <= 2048, blockDim
// NT); returns the total.
template <uint32_t NT>
__device__ __forceinline__ uint32_t res_bin_scan(const uint32_t* s_cnt, uint32_t* s_loc, uint32_t nbins,
                                                 uint32_t* s_w) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t E = (nbins + NT - 1) / NT, b0 = tid * E;
  uint32_t sum = 0;
  for (uint32_t e = 0; e < E; e++)
    if (b0 + e < nbins) sum += s_cnt[b0 + e];
  uint32_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wv] = x;
  __syncthreads();
  uint32_t pre = x - sum;
  for (int i = 0; i < wv; i++) pre += s_w[i];
  for (uint32_t e = 0; e < E; e++)
    if (b0 + e < nbins) {
      s_loc[b0 + e] = pre;
      pre += s_cnt[b0 + e];
    }
  uint32_t tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < NT / 64; i++) tot += s_w[i];
  return tot;
}

// kThreads threads per block; tiles above 4,096 words recompute each word's
// bin in the write loop instead of keeping it in LDS.
template <uint32_t kResTile, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_res_scatter(const uint64_t* __restrict__ in,
                                                          const ResSlice* __restrict__ sl, TableView t,
                                                          uint32_t shift, uint32_t mask, uint32_t nbins,
                                                          const uint64_t* __restrict__ base,
                                                          uint64_t* __restrict__ out) {
  constexpr bool kKeepBin = kResTile <= 4096;
  extern __shared__ uint64_t s_res[];
  uint64_t* s_tile = s_res;                                              // kResTile words
  uint64_t* s_pos = s_res + kResTile;                                    // nbins: next output index
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_pos + nbins);          // nbins: tile counts
  uint32_t* s_loc = s_cnt + nbins;                                       // nbins: tile offsets
  uint16_t* s_tbin = reinterpret_cast<uint16_t*>(s_loc + nbins);         // kResTile bins (kKeepBin)
  __shared__ uint32_t s_w[kThreads / 64];
  constexpr int kPer = kResTile / kThreads;
  const int tid = threadIdx.x;
  const ResSlice d = sl[blockIdx.x];
  for (uint32_t i = tid; i < nbins; i += kThreads) s_pos[i] = base[d.hbase + (uint64_t)i * d.m + d.j];
  for (uint64_t t0 = d.begin; t0 < d.end; t0 += kResTile) {
    const uint32_t nt = (uint32_t)std::min<uint64_t>(kResTile, d.end - t0);
    for (uint32_t i = tid; i < nbins; i += kThreads) s_cnt[i] = 0;
    __syncthreads();
    uint64_t w[kPer];
    uint32_t bin[kPer], rank[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const uint32_t i = u * kThreads + tid;
      w[u] = i < nt ? in[t0 + i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      bin[u] = 0;
      rank[u] = 0;
      if (w[u]) {
        bin[u] = (res_region(t, w[u]) >> shift) & mask;
        rank[u] = atomicAdd(&s_cnt[bin[u]], 1u);
      }
    }
    __syncthreads();
    const uint32_t nv = res_bin_scan<kThreads>(s_cnt, s_loc, nbins, s_w);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; u++)
      if (w[u]) {
        const uint32_t p = s_loc[bin[u]] + rank[u];
        s_tile[p] = w[u];
        if (kKeepBin) s_tbin[p] = (uint16_t)bin[u];
      }
    __syncthreads();
    for (uint32_t i = tid; i < nv; i += kThreads) {
      const uint64_t x = s_tile[i];
      const uint32_t b = kKeepBin ? (uint32_t)s_tbin[i] : (res_region(t, x) >> shift) & mask;
      out[s_pos[b] + (i - s_loc[b])] = x;
    }
    __syncthreads();
    for (uint32_t i = tid; i < nbins; i += kThreads) s_pos[i] += s_cnt[i];
  }
}

// First entry of every coarse bin of the scanned coarse histogram (bin c at
// c * ns1) and the total, into out[0, C].
__global__ void k_res_coarse_starts(const uint64_t* __restrict__ hs, uint64_t ns1, uint32_t C,
                                    uint64_t* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= C) out[c] = hs[(uint64_t)c * ns1];
}

// First word of every region (region g = coarse g >> fs, fine g & (2^fs-1))
// from the scanned fine histogram; rstart[R] = total.
__global__ void k_res_bounds(const uint64_t* __restrict__ base2, const uint64_t* __restrict__ hb,
                             const uint32_t* __restrict__ m, uint32_t fs, uint32_t R, uint64_t total,
                             uint64_t* __restrict__ rstart) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g > R) return;
  if (g == R) {
    rstart[R] = total;
    return;
  }
  const uint32_t c = (uint32_t)(g >> fs), f = (uint32_t)g & ((1u << fs) - 1);
  rstart[g] = base2[hb[c] + (uint64_t)f * m[c]];
}

// One workgroup per region: its buckets in LDS (zero when the table is
// known empty) with an 8-bit tag per slot beside them, its words added there
// with the global slot protocol (claim an empty slot with CAS, add to the
// slot holding the key; occupied slots stay a prefix of their bucket) through
// the tags, so that a probe reads 8 bytes of tags and only the slots whose
// tag matches instead of the whole 64-B bucket; carries to the side table;
// the region's slots written back.  A word that finds no room in the region
// is marked in place (occupied bit cleared) and flags[6] set; it is added
// after the table grows.
template <bool FRESH>
__global__ __launch_bounds__(1024) void k_res_count(uint64_t* __restrict__ words,
                                                    const uint64_t* __restrict__ rstart, TableView t,
                                                    uint32_t* __restrict__ flags, int zero_empty) {
  extern __shared__ uint64_t s_reg[];  // 8 << rbits slots, then 8 << rbits tag bytes
  __shared__ uint32_t s_new;
  constexpr int kPre = 4;  // words loaded ahead of their LDS inserts
  const uint64_t a = rstart[blockIdx.x], e = rstart[blockIdx.x + 1];
  const uint32_t nslots = 8u << t.rbits, n16 = nslots / 2;
  ulonglong2* g = reinterpret_cast<ulonglong2*>(t.slots + ((t.boff + ((uint64_t)blockIdx.x << t.rbits)) << 3));
  if (a == e) {
    // a region with no words of an uncleared table is written as empty
    if (zero_empty)
      for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) g[i] = make_ulonglong2(0ull, 0ull);
    return;
  }
  uint8_t* s_tag = reinterpret_cast<uint8_t*>(s_reg + nslots);
  ulonglong2* l = reinterpret_cast<ulonglong2*>(s_reg);
  for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) {
    if (FRESH) {
      l[i] = make_ulonglong2(0ull, 0ull);
      reinterpret_cast<uint16_t*>(s_tag)[i] = 0;
    } else {
      const ulonglong2 v = g[i];
      l[i] = v;
      const uint32_t t0 = v.x ? slot_tag(mix64(v.x & (t.occ - 1))) : 0u;
      const uint32_t t1 = v.y ? slot_tag(mix64(v.y & (t.occ - 1))) : 0u;
      reinterpret_cast<uint16_t*>(s_tag)[i] = (uint16_t)(t0 | (t1 << 8));
    }
  }
  if (threadIdx.x == 0) s_new = 0;
  __syncthreads();
  const uint32_t rmask = (uint32_t)t.rmask;
  uint32_t mynew = 0;
  bool spill = false;
  for (uint64_t i0 = a; i0 < e; i0 += (uint64_t)kPre * blockDim.x) {
    uint64_t wp[kPre];
#pragma unroll
    for (int u = 0; u < kPre; u++) {
      const uint64_t i = i0 + (uint64_t)u * blockDim.x + threadIdx.x;
      wp[u] = i < e ? words[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPre; u++) {
      const uint64_t w = wp[u];
      if (!w) continue;
      const uint64_t tagged = w & t.kmo;
      const uint64_t c = w >> t.cshift;
      const uint64_t h = mix64(w & (t.occ - 1));
      const uint32_t b = (uint32_t)__umul64hi(h, t.nb) & rmask;
      uint64_t old = 0;
      // every bucket of the region, each rescanned at most 8 times
      const int res = lds_tag_add(s_reg, s_tag, b, rmask + 1, w, slot_tag(h), t.kmo, t.cshift, 9 * (rmask + 1), old);
      const bool carry = res == 1 && (old >> t.cshift) + c > t.cmax;
      if (res == 2) mynew++;
      if (carry) ovf_add(t, tagged, t.cmax + 1, &flags[0]);
      if (!res) {
        words[i0 + (uint64_t)u * blockDim.x + threadIdx.x] = w & ~t.occ;
        spill = true;
      }
    }
  }
  if (spill) atomicOr(&flags[6], 1u);
  mynew = wave_sum_u32(mynew);
  if ((threadIdx.x & 63) == 0 && mynew) atomicAdd(&s_new, mynew);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) g[i] = l[i];
  if (threadIdx.x == 0 && s_new)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)s_new);
}

// Add the words that k_res_count marked (occupied bit clear) to a grown
// table; words that still find no room stay marked (flags[6]).
__global__ void k_res_retry(uint64_t* __restrict__ words, uint64_t n, TableView t,
                            uint32_t* __restrict__ flags) {
  uint32_t mynew = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t w = words[i];
    if (!w || (w & t.occ)) continue;
    const uint32_t r = table_add(t, w & (t.occ - 1), w >> t.cshift, &flags[0]);
    if (r == 2u) {
      atomicOr(&flags[6], 1u);
    } else {
      mynew += r;
      words[i] = w | t.occ;
    }
  }
  mynew = wave_sum_u32(mynew);
  if ((threadIdx.x & 63) == 0 && mynew)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)mynew);
}

// ---------------------------------------------------------------------
// Region-partitioned counting (tables of several Infinity-Cache-sized
// regions).  Pass 1 groups the canonical keys of a read chunk by region
// (per-block histogram, one exclusive scan, scatter to exact positions);
// pass 2 streams the region-major key array through the cooperative insert,
// so the few regions being updated at any time stay cache-resident.
// ---------------------------------------------------------------------
constexpr int kPartReads = 256;
constexpr uint32_t kMaxRegions = 2048;

// Window prefix of the block's reads [r0, r0+nr), nr <= 256, blockDim 256.
__device__ __forceinline__ uint32_t block_windows_256(const ReadsView& rv, uint64_t r0, int nr,
                                                      int k, uint32_t* s_off, uint64_t* s_wb,
                                                      uint32_t* s_wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t nw = 0;
  if (tid < nr) {
    const uint32_t L = rv.len[r0 + tid];
    nw = L >= (uint32_t)k ? L - k + 1 : 0;
    s_wb[tid] = rv.word_off[r0 + tid];
  }
  uint32_t x = nw;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[wv] = x;
  __syncthreads();
  uint32_t add = 0;
  for (int i = 0; i < wv; i++) add += s_wsum[i];
  s_off[tid + 1] = x + add;
  if (tid == 0) s_off[0] = 0;
  __syncthreads();
  return s_off[nr];
}

__device__ __forceinline__ int read_of_window(const uint32_t* s_off, int nr, uint32_t w) {
  int lo = 0, hi = nr;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_off[mid] <= w) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_part_hist(ReadsView rv, uint64_t r_begin, uint64_t r_end,
                                                   int k, int canonical, TableView t, uint32_t nparts,
                                                   uint32_t nblocks, uint64_t* __restrict__ hist) {
  __shared__ uint32_t s_off[kPartReads + 1];
  __shared__ uint64_t s_wb[kPartReads];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_cnt[kMaxRegions];
  const uint64_t r0 = r_begin + (uint64_t)blockIdx.x * kPartReads;
  const int nr = (int)std::min<uint64_t>(kPartReads, r_end - r0);
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s_cnt[i] = 0;
  const uint32_t total = block_windows_256(rv, r0, nr, k, s_off, s_wb, s_wsum);
  for (uint32_t w = threadIdx.x; w < total; w += blockDim.x) {
    const int j = read_of_window(s_off, nr, w);
    uint64_t fwd;
    bool valid;
    window_at(rv, s_wb[j], w - s_off[j], k, fwd, valid);
    if (valid) atomicAdd(&s_cnt[bucket0(t, canon_key(fwd, k, canonical)) >> t.rbits], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x)
    hist[(uint64_t)i * nblocks + blockIdx.x] = s_cnt[i];
}

__global__ __launch_bounds__(256) void k_part_scatter(ReadsView rv, uint64_t r_begin, uint64_t r_end,
                                                      int k, int canonical, TableView t,
                                                      uint32_t nparts, uint32_t nblocks,
                                                      const uint64_t* __restrict__ base,
                                                      uint64_t* __restrict__ keys_out) {
  __shared__ uint32_t s_off[kPartReads + 1];
  __shared__ uint64_t s_wb[kPartReads];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_rank[kMaxRegions];
  __shared__ uint64_t s_base[kMaxRegions];
  const uint64_t r0 = r_begin + (uint64_t)blockIdx.x * kPartReads;
  const int nr = (int)std::min<uint64_t>(kPartReads, r_end - r0);
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
    s_rank[i] = 0;
    s_base[i] = base[(uint64_t)i * nblocks + blockIdx.x];
  }
  const uint32_t total = block_windows_256(rv, r0, nr, k, s_off, s_wb, s_wsum);
  for (uint32_t w = threadIdx.x; w < total; w += blockDim.x) {
    const int j = read_of_window(s_off, nr, w);
    uint64_t fwd;
    bool valid;
    window_at(rv, s_wb[j], w - s_off[j], k, fwd, valid);
    if (valid) {
      const uint64_t key = canon_key(fwd, k, canonical);
      const uint32_t p = (uint32_t)(bucket0(t, key) >> t.rbits);
      keys_out[s_base[p] + atomicAdd(&s_rank[p], 1u)] = key;
    }
  }
}

// Pass 2: 4096 keys per block, cooperative insert with a short probe limit;
// keys whose region is crowded past it go to the spill list, or, once that is
// full, are marked in place with bit 63; both are finished after the table
// grows.  flags[0] error bits, flags[2..3] new slots, flags[4..5] spill
// cursor, flags[6] marks present.
constexpr uint32_t kPartKeysPerBlock = 4096;

__global__ __launch_bounds__(256) void k_part_count(uint64_t* __restrict__ keys, uint64_t n,
                                                    TableView t, uint64_t* __restrict__ spill,
                                                    uint64_t spill_cap, uint32_t* __restrict__ flags) {
  const uint64_t a = (uint64_t)blockIdx.x * kPartKeysPerBlock;
  const uint64_t e = std::min<uint64_t>(n, a + kPartKeysPerBlock);
  const int lane = threadIdx.x & 63;
  uint32_t mynew = 0;
  uint32_t spill_err = 0;
  for (uint64_t i0 = a; i0 < e; i0 += blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    const bool want = i < e;
    const uint64_t key = want ? __builtin_nontemporal_load(keys + i) : 0;
    bool lost = false;
    mynew += coop_add64(t, key, want, lane, &flags[0], lost);
    if (lost) {
      const unsigned long long at =
          atomicAdd(reinterpret_cast<unsigned long long*>(&flags[4]), 1ull);
      if (at < spill_cap) {
        spill[at] = key;
      } else {
        keys[i] = key | (1ull << 63);  // left for the retry pass
        spill_err = 1;
      }
    }
  }
  if (spill_err) atomicOr(&flags[6], 1u);
  mynew = wave_sum_u32(mynew);
  if (lane == 0 && mynew)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)mynew);
}

// Move every slot of an old table into a larger one (slot values unchanged).
__global__ void k_rehash(const uint64_t* __restrict__ old, uint64_t n_slots, TableView t,
                         uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = old[i];
    if (!s) continue;
    uint64_t b = bucket0(t, s & (t.occ - 1));
    bool done = false;
    for (uint32_t p = 0; p < t.max_probe && !done; p++) {
      uint64_t* line = t.slots + (b << 3);
      for (int j = 0; j < 8; j++) {
        if (line[j] == 0 &&
            atomicCAS((unsigned long long*)&line[j], 0ULL, (unsigned long long)s) == 0ULL) {
          done = true;
          break;
        }
      }
      b = bucket_next(t, b);
    }
    if (!done) atomicOr(&flags[0], 1u);
  }
}

__global__ void k_ovf_rehash(const uint64_t* __restrict__ old, uint64_t n_pairs, TableView t,
                             uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = old[2 * i];
    if (!key) continue;
    ovf_add(t, key, old[2 * i + 1], &flags[0]);
  }
}

__device__ inline uint64_t* table_find_slot(const TableView& t, uint64_t tagged) {
  uint64_t b = bucket0(t, tagged & (t.occ - 1));
  for (uint32_t p = 0; p < t.max_probe; p++) {
    uint64_t* line = t.slots + (b << 3);
    for (int j = 0; j < 8; j++) {
      const uint64_t s = line[j];
      if ((s & t.kmo) == tagged) return &line[j];
      if (s == 0) return nullptr;
    }
    b = bucket_next(t, b);
  }
  return nullptr;
}

// Fold side-table carries into full counts (finalise) or back (reopen).
__global__ void k_ovf_fold(TableView t, uint64_t n_pairs, int finalize,
                           uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t tagged = t.ovf[2 * i];
    if (!tagged) continue;
    uint64_t* s = table_find_slot(t, tagged);
    if (!s) { atomicOr(&flags[0], 4u); continue; }
    if (finalize) {
      const uint64_t total = t.ovf[2 * i + 1] + (*s >> t.cshift);
      t.ovf[2 * i + 1] = total;
      *s = total ? ((*s & t.kmo) | (t.cmax << t.cshift)) : (*s & t.kmo);
    } else {
      const uint64_t total = t.ovf[2 * i + 1];
      const uint64_t r = total & t.cmax;
      t.ovf[2 * i + 1] = total - r;
      *s = (*s & t.kmo) | (r << t.cshift);
    }
  }
}

static int table_alloc_slots(tk_ctx* ctx, uint64_t nbuckets, uint64_t** p, bool zero = true) {
  int rc = dev_alloc((void**)p, nbuckets * 64);
  if (rc) return rc;
  if (zero) TK_HIP(hipMemsetAsync(*p, 0, nbuckets * 64, ctx->stream));
  return TK_OK;
}

// Clear the slots of a table created without clearing them.  Every entry
// point that reads or writes slots calls this (through table_set_final)
// first, except the region-resident count of an empty table, which writes
// every region itself.
static int table_zero(tk_ctx* ctx, tk_table* t) {
  if (!t->zero_pending) return TK_OK;
  TK_HIP(hipMemsetAsync(t->d_slots, 0, t->nb * 64, ctx->stream));
  t->zero_pending = false;
  return TK_OK;
}

static int table_ensure_ovf(tk_ctx* ctx, tk_table* t, uint64_t bound) {
  uint64_t need = next_pow2(std::max<uint64_t>(4096, 2 * bound + 64));
  if (t->d_ovf && t->ovf_cap >= need) return TK_OK;
  uint64_t* nov = nullptr;
  int rc = dev_alloc((void**)&nov, need * 16);
  if (rc) return rc;
  TK_HIP(hipMemsetAsync(nov, 0, need * 16, ctx->stream));
  uint64_t* old = t->d_ovf;
  uint64_t oldcap = t->ovf_cap;
  t->d_ovf = nov;
  t->ovf_cap = need;
  if (old) {
    TK_HIP(hipMemsetAsync(ctx->d_flags, 0, 4, ctx->stream));
    TableView v = table_view(t);
    TK_LAUNCH(ctx, "ovf_rehash", std::min<uint64_t>(ceil_div(oldcap, 256), 8192), 256, 0,
              k_ovf_rehash, old, oldcap, v, ctx->d_flags);
    TK_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(old, oldcap * 16);
  }
  t->bytes = t->nb * 64 + t->ovf_cap * 16;
  return TK_OK;
}

// Region size for large tables (bytes); a region is the unit counted while
// resident in the Infinity Cache.
// Bytes per table region: 64 KB by default, the size one workgroup of the
// region-resident count holds in LDS.  TK_REGION_KB or TK_REGION_MB
// overrides it (read when a table is laid out).
static uint64_t region_bytes() {
  if (const char* e = getenv("TK_REGION_MB"))
    return next_pow2(std::max<uint64_t>(strtoull(e, nullptr, 10), 1)) << 20;
  if (const char* e = getenv("TK_REGION_KB"))
    return next_pow2(std::max<uint64_t>(strtoull(e, nullptr, 10), 1)) << 10;
  return 64ull << 10;
}

// Bucket layout holding `keys` at load max_load: nb = regions * 2^rbits.
static void table_layout(uint64_t keys, double max_load, uint64_t* nb, uint32_t* rbits) {
  const uint64_t need = std::max<uint64_t>(ceil_div((uint64_t)std::ceil((double)keys / max_load), 8), 8);
  const uint64_t rb_max = region_bytes() / 64;
  uint64_t rb = std::min<uint64_t>(next_pow2(need), rb_max);
  uint32_t bits = 0;
  while ((1ull << bits) < rb) bits++;
  *rbits = bits;
  *nb = ceil_div(need, rb) * rb;
}

// Rehash into a table that holds max(min_keys, 2 x current) keys at max_load.
static int table_grow(tk_ctx* ctx, tk_table* t, uint64_t min_keys) {
  uint64_t nb = 0;
  uint32_t rbits = 0;
  const uint64_t doubled = (uint64_t)(2.0 * t->max_load * (double)(t->nb * 8));
  table_layout(std::max<uint64_t>(min_keys, doubled), t->max_load, &nb, &rbits);
  uint64_t* ns = nullptr;
  int rc = table_alloc_slots(ctx, nb, &ns);
  if (rc) return rc;
  uint64_t* old = t->d_slots;
  uint64_t oldnb = t->nb;
  t->d_slots = ns;
  t->nb = nb;
  t->rbits = rbits;
  TK_HIP(hipMemsetAsync(ctx->d_flags, 0, 4, ctx->stream));
  TableView v = table_view(t);
  TK_LAUNCH(ctx, "rehash", std::min<uint64_t>(ceil_div(oldnb * 8, 256), 65536), 256, 0,
            k_rehash, old, oldnb * 8, v, ctx->d_flags);
  TK_HIP(hipMemcpyAsync(ctx->h_flags, ctx->d_flags, 4, hipMemcpyDeviceToHost, ctx->stream));
  TK_HIP(hipStreamSynchronize(ctx->stream));
  dev_free(old, oldnb * 64);
  t->bytes = t->nb * 64 + t->ovf_cap * 16;
  if (ctx->h_flags[0]) return fail(TK_ERR_FULL, "rehash into %llu buckets failed",
                                   (unsigned long long)nb);
  (void)oldnb;
  return TK_OK;
}

static int table_set_final(tk_ctx* ctx, tk_table* t, bool final_) {
  if (int rc = table_zero(ctx, t)) return rc;
  if (t->finalized == final_) return TK_OK;
  TK_HIP(hipMemsetAsync(ctx->d_flags, 0, 4, ctx->stream));
  TableView v = table_view(t);
  TK_LAUNCH(ctx, "ovf_fold", std::min<uint64_t>(ceil_div(t->ovf_cap, 256), 8192), 256, 0,
            k_ovf_fold, v, t->ovf_cap, final_ ? 1 : 0, ctx->d_flags);
  TK_HIP(hipMemcpyAsync(ctx->h_flags, ctx->d_flags, 4, hipMemcpyDeviceToHost, ctx->stream));
  TK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->h_flags[0]) return fail(TK_ERR_STATE, "side-table key missing from main table");
  t->finalized = final_;
  return TK_OK;
}

int table_finalize(tk_ctx* ctx, tk_table* t) { return table_set_final(ctx, t, true); }

// Thread per (key,count) pair.
__global__ void k_load(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ counts,
                       uint64_t n, TableView t, uint32_t* __restrict__ flags) {
  uint32_t mynew = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = counts[i];
    const uint64_t key = keys[i];
    if (c == 0) continue;
    if (key & ~(t.occ - 1)) { atomicOr(&flags[0], 8u); continue; }
    const uint32_t r = table_add(t, key, c, &flags[0]);
    if (r == 2u) atomicOr(&flags[0], 1u);
    else mynew += r;
  }
  mynew = wave_sum_u32(mynew);
  if ((threadIdx.x & 63) == 0 && mynew)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)mynew);
}

// =====================================================================
// Dump / histo / prune / lookup over a finalised table
// =====================================================================
__global__ void k_dump_count(TableView t, uint64_t n_slots, uint64_t lo, uint64_t hi,
                             unsigned long long* __restrict__ out) {
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = t.slots[i];
    if (!s || (s >> t.cshift) == 0) continue;
    uint64_t v = slot_count(t, s);
    v = v > 0xFFFFFFFFull ? 0xFFFFFFFFull : v;  // counts print saturated at 2^32-1
    c += (v >= lo && v <= hi);
  }
  block_sum_add(c, out);
}

// Keys with count in [lo, hi] and key in [klo, khi] appended at *cursor.
__global__ void k_dump(TableView t, uint64_t n_slots, uint64_t lo, uint64_t hi, uint64_t klo, uint64_t khi,
                       uint64_t* __restrict__ keys, uint32_t* __restrict__ counts,
                       uint64_t cap, unsigned long long* __restrict__ cursor) {
  const int lane = threadIdx.x & 63;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n_slots;
       base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool take = false;
    uint64_t key = 0, v = 0;
    if (i < n_slots) {
      const uint64_t s = t.slots[i];
      if (s && (s >> t.cshift) != 0) {
        v = slot_count(t, s);
        v = v > 0xFFFFFFFFull ? 0xFFFFFFFFull : v;
        key = s & (t.occ - 1);
        take = v >= lo && v <= hi && key >= klo && key <= khi;
      }
    }
    const uint64_t m = __ballot(take);
    if (!m) continue;
    unsigned long long start = 0;
    if (lane == 0) start = atomicAdd(cursor, (unsigned long long)__popcll(m));
    start = __shfl(start, 0);
    if (take) {
      const uint64_t pos = start + __popcll(m & ((1ull << lane) - 1));
      if (pos < cap) {
        keys[pos] = key;
        counts[pos] = v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
      }
    }
  }
}

constexpr int kHistoLds = 4096;

// Keys with count in [lo, hi] and key in [klo, khi] per key range:
// bin = (key - klo) >> shift (< kHistoLds bins), so that a dump of more keys
// than one sort call takes can be cut into key ranges that are sorted and
// downloaded one after another.
__global__ __launch_bounds__(256) void k_dump_bins(TableView t, uint64_t n_slots, uint64_t lo, uint64_t hi,
                                                   uint64_t klo, uint64_t khi, uint32_t shift,
                                                   unsigned long long* __restrict__ bins) {
  __shared__ uint32_t s_bins[kHistoLds];
  for (int i = threadIdx.x; i < kHistoLds; i += blockDim.x) s_bins[i] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = t.slots[i];
    if (!s || (s >> t.cshift) == 0) continue;
    uint64_t v = slot_count(t, s);
    v = v > 0xFFFFFFFFull ? 0xFFFFFFFFull : v;
    const uint64_t key = s & (t.occ - 1);
    if (v >= lo && v <= hi && key >= klo && key <= khi) atomicAdd(&s_bins[(key - klo) >> shift], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHistoLds; i += blockDim.x)
    if (s_bins[i]) atomicAdd(&bins[i], (unsigned long long)s_bins[i]);
}

__global__ __launch_bounds__(256) void k_histo(TableView t, uint64_t n_slots, uint64_t base,
                                               uint64_t ceil_, uint64_t inc, uint64_t nb,
                                               unsigned long long* __restrict__ bins) {
  __shared__ uint32_t s_bins[kHistoLds];
  const bool lds = nb <= kHistoLds;
  if (lds) {
    for (int i = threadIdx.x; i < kHistoLds; i += blockDim.x) s_bins[i] = 0;
    __syncthreads();
  }
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = t.slots[i];
    if (!s || (s >> t.cshift) == 0) continue;
    const uint64_t v = slot_count(t, s);
    uint64_t b;
    if (v < base) b = 0;
    else if (v >= ceil_) b = nb - 1;
    else b = (v - base) / inc;
    if (lds) atomicAdd(&s_bins[b], 1u);
    else atomicAdd(&bins[b], 1ull);
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < (int)nb; i += blockDim.x)
      if (s_bins[i]) atomicAdd(&bins[i], (unsigned long long)s_bins[i]);
  }
}

__device__ __forceinline__ void composition(uint64_t key, int k, int& a, int& c, int& g) {
  const uint64_t M = (k == 32) ? 0x5555555555555555ULL
                               : (0x5555555555555555ULL & ((1ull << (2 * k)) - 1));
  const uint64_t lo = key & M, hi = (key >> 1) & M;
  const int nt = __popcll(hi & lo);
  g = __popcll(hi & ~lo & M);
  c = __popcll(~hi & lo & M);
  a = k - nt - g - c;
}

__global__ void k_prune(TableView t, uint64_t n_slots, int k, uint64_t min_count,
                        const uint8_t* __restrict__ keep, unsigned long long* __restrict__ kept) {
  uint64_t nk = 0;
  const int K1 = k + 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = t.slots[i];
    if (!s) continue;
    const uint64_t f = s >> t.cshift;
    if (f == 0) continue;
    const uint64_t v = slot_count(t, s);
    bool ok = v >= min_count;
    if (ok && keep) {
      int a, c, g;
      composition(s & (t.occ - 1), k, a, c, g);
      ok = keep[(a * K1 + c) * K1 + g] != 0;
    }
    if (ok) {
      nk++;
    } else {
      if (f == t.cmax) {
        uint64_t* o = ovf_find(t, s & t.kmo);
        if (o) *o = 0;
      }
      t.slots[i] = s & t.kmo;
    }
  }
  block_sum_add(nk, kept);
}

__global__ void k_lookup(TableView t, const uint64_t* __restrict__ keys, uint64_t n,
                         uint32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[i];
    uint64_t v = (key & ~(t.occ - 1)) ? 0 : table_get(t, key);
    out[i] = v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
  }
}

// =====================================================================
// Per-read coverage statistics: one wavefront per read
// =====================================================================
template <bool kShort, bool kTicket>
__device__ __forceinline__ void stats_pass(ReadsView rv, int k, int canonical, TableView t, StatsParams sp,
                                           tk_read_stat* __restrict__ out, uint32_t slab, uint32_t* tickets,
                                           uint32_t grab) {
  extern __shared__ uint32_t s_dyn[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t* vals = s_dyn + (uint64_t)wv * slab;
  uint64_t* tmp = reinterpret_cast<uint64_t*>(s_dyn + wpb * slab) + wv * kLookupScratch;
  // XCD x (= blockIdx % 8, the dispatch order; checked on the device with
  // the XCC_ID register) takes positions [x*n/8, (x+1)*n/8) of the read
  // order (gridDim % 8 == 0) and its waves stride through that range
  // together, so the reads in flight on one XCD are neighbours in the order
  // and share its L2.  Measured slower: each wave scoring a run of
  // consecutive reads by itself (fewer simultaneous misses on one locus, but
  // too few reads in flight once the runs fit L2), and a shared work counter
  // per XCD (same-address atomics serialize at the memory side).
  const uint64_t per_xcd = (rv.n_reads + 7) >> 3;
  const uint64_t x = blockIdx.x & 7;
  const uint64_t lo = x * per_xcd;
  const uint64_t hi = lo + per_xcd < rv.n_reads ? lo + per_xcd : rv.n_reads;
  if (lo >= hi) return;
  xcd_positions<kTicket>(kTicket ? tickets + x * kTicketStride : nullptr, hi - lo, grab, (blockIdx.x >> 3) * wpb + wv,
                         (uint64_t)(gridDim.x >> 3) * wpb, lane, [&](uint64_t j) {
                           const uint64_t r = read_at(rv, lo + j);
                           stats_read<kShort>(rv, r, k, canonical, t, sp, vals, tmp, lane, out + r);
                         });
}

// The scoring pass for any read lengths, and for read sets whose reads all
// have at most 128 windows (no long-read path compiled: 77 instead of 86
// VGPRs, 6 waves per SIMD instead of 5).
template <bool kTicket>
__global__ __launch_bounds__(256) void k_stats(ReadsView rv, int k, int canonical, TableView t,
                                               StatsParams sp, tk_read_stat* __restrict__ out,
                                               uint32_t slab, uint32_t* tickets, uint32_t grab) {
  stats_pass<false, kTicket>(rv, k, canonical, t, sp, out, slab, tickets, grab);
}
template <bool kTicket>
__global__ __launch_bounds__(256) void k_stats_short(ReadsView rv, int k, int canonical, TableView t,
                                                     StatsParams sp, tk_read_stat* __restrict__ out,
                                                     uint32_t slab, uint32_t* tickets, uint32_t grab) {
  stats_pass<true, kTicket>(rv, k, canonical, t, sp, out, slab, tickets, grab);
}

// k_stats_short over ticket runs, software-pipelined over each wave's reads:
// while read j is scored, read j+1's code words (spread over lanes 0-5),
// read j+2's length and first word and read j+3's index in the order are
// loaded, so a read's table lookups wait on one memory round trip instead of
// four dependent ones (order -> length / offset -> codes -> buckets).
__global__ __launch_bounds__(256) void k_stats_pipe(ReadsView rv, int k, int canonical, TableView t,
                                                    StatsParams sp, tk_read_stat* __restrict__ out,
                                                    uint32_t* tickets, uint32_t grab) {
  extern __shared__ uint64_t s_pipe[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  uint64_t* tmp = s_pipe + wv * kLookupScratch;
  const uint64_t per_xcd = (rv.n_reads + 7) >> 3;
  const uint64_t x = blockIdx.x & 7;
  const uint64_t lo = x * per_xcd;
  const uint64_t hi = lo + per_xcd < rv.n_reads ? lo + per_xcd : rv.n_reads;
  if (lo >= hi) return;
  const uint32_t n = (uint32_t)(hi - lo);
  TicketStream ts(tickets + x * kTicketStride, n, grab, lane);
  // stage a: index known; b: length and first word known; c: words known
  bool va = false, vb = false, vc = false;
  uint32_t ra = 0, rb = 0, rc = 0, Lb = 0, Lc = 0, mwc = 0;
  uint64_t wbb = 0, cwc = 0;
  for (;;) {
    const uint32_t p = ts.next();
    const bool vn = p < n;
    uint32_t rn = 0;
    if (vn) rn = rv.order ? rv.order[lo + p] : (uint32_t)(lo + p);
    uint32_t La = 0;
    uint64_t wba = 0;
    if (va) {
      La = rv.len[ra];
      wba = rv.word_off[ra];
    }
    uint64_t cwb = 0;
    uint32_t mwb = 0;
    if (vb && lane < 6 && (uint32_t)lane <= ((Lb + 31) >> 5)) {
      cwb = rv.codes[wbb + lane];
      mwb = rv.nmask[wbb + lane];
    }
    if (vc) {
      const uint32_t nwin = Lc >= (uint32_t)k ? Lc - k + 1 : 0;
      uint32_t nn, med;
      uint64_t sum, sqlo, sqhi;
      stats_short_core(nwin, k, canonical, t, sp, tmp, lane,
                       [&](uint32_t pos, uint64_t& f, bool& v) { window_lanes(cwc, mwc, pos, k, f, v); }, nn, med,
                       sum, sqlo, sqhi);
      write_stat(lane, med, nn, sum, sqlo, sqhi, out + rc);
      wave_sync();
    }
    vc = vb;
    rc = rb;
    Lc = Lb;
    cwc = cwb;
    mwc = mwb;
    vb = va;
    rb = ra;
    Lb = La;
    wbb = wba;
    va = vn;
    ra = rn;
    if (!va && !vb && !vc) break;
  }
}

// k_stats for ordered reads of at most 128 windows each: every block takes
// kCacheChunk consecutive reads of the order (four at a time, one per wave)
// and keeps the counts it looked up in an LDS cache of CACHE slots shared by
// its waves (CountCache), so a locus' k-mers are fetched from the table about
// once per block instead of once per read.  Blocks map to XCDs as xcd_block.
constexpr uint32_t kCacheChunk = 64;

template <int CACHE>
__global__ __launch_bounds__(256) void k_stats_cached(ReadsView rv, int k, int canonical, TableView t,
                                                      StatsParams sp, tk_read_stat* __restrict__ out) {
  extern __shared__ uint64_t s_lds[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint32_t wpb = blockDim.x >> 6;
  uint64_t* tmp = s_lds + CACHE + (uint32_t)wv * kLookupScratch;
  CountCache cc{s_lds, reinterpret_cast<uint32_t*>(s_lds + CACHE + wpb * kLookupScratch), CACHE - 1};
  for (int i = threadIdx.x; i < CACHE; i += blockDim.x) {
    cc.keys[i] = 0;
    cc.vals[i] = kNotReady;
  }
  __syncthreads();
  const uint64_t i0 = xcd_block(blockIdx.x, gridDim.x) * kCacheChunk;
  for (uint32_t j = wv; j < kCacheChunk; j += wpb) {
    const uint64_t i = i0 + j;
    if (i >= rv.n_reads) break;
    const uint64_t r = read_at(rv, i);
    stats_read_cached(rv, r, k, canonical, t, sp, cc, tmp, lane, out + r);
  }
}

// =====================================================================
// Pair selection (nbkc_merge_left_right_stats.pl + nbkc_normalize.pl)
// =====================================================================
struct Lcg48Jump {
  uint64_t a[48];
  uint64_t c[48];
};

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// consumes[i] = 1 when pair i reaches nbkc_normalize.pl's rand() draw
// (:91-111).  The merged median is exact ((l+r)/2 of integers, printed with
// "%.1f").  The zero-mean test is exact when every read with windows has a
// mean of at least 1 (sum >= nwin, as with missing k-mers counted as 1): the
// merged mean is then 0 exactly when neither read has a window.  err gets
// bit 0 for a read whose mean is below 1 (its printed mean could round the
// merged mean to "0.0"), bit 1 for a drawn pair whose median is 0 (the
// script divides max_cov by it and dies).
// out[i] = stats[idx[i]]: the rows of one selection's units in their order
// (the merged-row order of a pair selection).  Indices are checked on the
// host; one out of range gives an empty row.
__global__ void k_gather_stats(const tk_read_stat* __restrict__ stats, uint64_t n_reads,
                               const uint64_t* __restrict__ idx, uint64_t n, tk_read_stat* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = idx[i];
    out[i] = j < n_reads ? stats[j] : tk_read_stat{0, 0, 0, 0, 0};
  }
}

__global__ void k_select_flags(const tk_read_stat* __restrict__ L, const tk_read_stat* __restrict__ R,
                               uint64_t n, int min_cov, uint32_t* __restrict__ consumes,
                               uint32_t* __restrict__ err) {
  uint32_t e = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const tk_read_stat l = L[i], r = R[i];
    const double med = ((double)l.median + (double)r.median) / 2.0;
    bool use = !(med < (double)min_cov);
    if (l.sum + r.sum == 0) use = false;  // merged mean "0.0" <= 0
    if ((l.nwin && l.sum < l.nwin) || (r.nwin && r.sum < r.nwin)) e |= 1u;
    if (use && med == 0.0) e |= 2u;
    consumes[i] = use ? 1u : 0u;
  }
  if (e) atomicOr(err, e);
}

// drand48 state after `steps` < 2^NB draws from x: x <- A_j x + C_j for each
// set bit j.  The loop runs over all NB bits so that every jump.a[j] /
// jump.c[j] index is a constant (one scalar load each, shared by the wave).
template <int NB>
__device__ __forceinline__ uint64_t lcg48_jump(const Lcg48Jump& jump, uint64_t x, uint64_t steps) {
  const uint64_t M = (1ull << 48) - 1;
#pragma unroll
  for (int j = 0; j < NB; j++)
    if ((steps >> j) & 1) x = (jump.a[j] * x + jump.c[j]) & M;
  return x;
}

// One thread per pair.  The ranks `idx` of the rand() draws are ascending, so
// each block jumps once to its first rank and every thread jumps the short
// distance from there.
__global__ __launch_bounds__(256) void k_select(const tk_read_stat* __restrict__ L,
                                                const tk_read_stat* __restrict__ R, uint64_t n,
                                                const uint32_t* __restrict__ consumes,
                                                const uint32_t* __restrict__ idx, Lcg48Jump jump,
                                                uint64_t state0, int max_cov, uint8_t* __restrict__ sel,
                                                unsigned long long* __restrict__ nsel) {
  __shared__ uint64_t s_base;
  __shared__ uint32_t s_idx0;
  uint64_t cnt = 0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += (uint64_t)gridDim.x * blockDim.x) {
    __syncthreads();
    if (threadIdx.x == 0) {
      s_idx0 = idx[i0];
      s_base = lcg48_jump<32>(jump, state0, idx[i0]);  // state before draw idx[i0] (u32 rank)
    }
    __syncthreads();
    const uint64_t i = i0 + threadIdx.x;
    if (i < n) {
      uint8_t s = 0;
      if (consumes[i]) {
        // ranks within a block of 256 pairs differ by at most 255: at most 256 draws
        const uint64_t x = lcg48_jump<9>(jump, s_base, (uint64_t)(idx[i] - s_idx0) + 1);
        const double r = (double)x * 0x1p-48;
        const double med = ((double)L[i].median + (double)R[i].median) / 2.0;
        s = (r <= (double)max_cov / med) ? 1 : 0;
      }
      sel[i] = s;
      cnt += s;
    }
  }
  block_sum_add(cnt, nsel);
}

// =====================================================================
// ReadsToTranscripts
// =====================================================================
__global__ __launch_bounds__(256) void k_ctable_build(ReadsView rv, const uint32_t* __restrict__ comp,
                                                      int k, int canonical, uint64_t* slots,
                                                      uint64_t bmask, uint32_t* __restrict__ flags) {
  __shared__ uint32_t s_off[kReadsPerBlock + 1];
  __shared__ uint64_t s_wb[kReadsPerBlock];
  __shared__ uint32_t s_comp[kReadsPerBlock];
  __shared__ uint32_t s_new;
  const uint64_t r0 = (uint64_t)blockIdx.x * kReadsPerBlock;
  if (r0 >= rv.n_reads) return;
  const int nr = (int)std::min<uint64_t>(kReadsPerBlock, rv.n_reads - r0);
  const int tid = threadIdx.x;
  if (tid < kReadsPerBlock) {
    uint32_t nw = 0;
    if (tid < nr) {
      const uint32_t L = rv.len[r0 + tid];
      nw = L >= (uint32_t)k ? L - k + 1 : 0;
      s_wb[tid] = rv.word_off[r0 + tid];
      s_comp[tid] = comp[r0 + tid];
    }
    uint32_t x = nw;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o);
      if (tid >= o) x += y;
    }
    s_off[tid + 1] = x;
    if (tid == 0) { s_off[0] = 0; s_new = 0; }
  }
  __syncthreads();
  const uint64_t occ = 1ull << (2 * k);
  const uint32_t total = s_off[nr];
  uint32_t mynew = 0;
  for (uint32_t w = tid; w < total; w += blockDim.x) {
    int lo = 0, hi = nr;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (s_off[mid] <= w) lo = mid; else hi = mid;
    }
    uint64_t fwd;
    bool valid;
    window_at(rv, s_wb[lo], w - s_off[lo], k, fwd, valid);
    if (valid) ctable_add(slots, 0, bmask, occ, canon_key(fwd, k, canonical), s_comp[lo], mynew, &flags[0]);
  }
  mynew = wave_sum_u32(mynew);
  if ((tid & 63) == 0 && mynew) atomicAdd(&s_new, mynew);
  __syncthreads();
  if (tid == 0 && s_new)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)s_new);
}

template <bool kTicket>
__global__ __launch_bounds__(256) void k_vote(ReadsView rv, int k, int canonical,
                                              const uint64_t* __restrict__ slots, uint64_t bmask,
                                              tk_r2t_vote* __restrict__ out, uint32_t slab, uint32_t* tickets,
                                              uint32_t grab) {
  extern __shared__ uint32_t s_dyn[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t* hits = s_dyn + (uint64_t)wv * slab;
  uint32_t* tmp = s_dyn + wpb * slab + wv * 64;
  // reads in the locality order, one contiguous range of it per XCD (as k_stats)
  const uint64_t per_xcd = (rv.n_reads + 7) >> 3;
  const uint64_t x = blockIdx.x & 7;
  const uint64_t lo = x * per_xcd;
  const uint64_t hi = lo + per_xcd < rv.n_reads ? lo + per_xcd : rv.n_reads;
  if (lo >= hi) return;
  xcd_positions<kTicket>(kTicket ? tickets + x * kTicketStride : nullptr, hi - lo, grab, (blockIdx.x >> 3) * wpb + wv,
                         (uint64_t)(gridDim.x >> 3) * wpb, lane, [&](uint64_t j) {
                           const uint64_t r = read_at(rv, lo + j);
                           vote_read(rv, r, k, canonical, slots, 0, bmask, hits, tmp, lane, out + r);
                         });
}

// k_vote over ticket runs for reads of at most 128 windows (unpacked
// component table), software-pipelined over each wave's reads as
// k_stats_pipe: read j+1's code words, read j+2's length and first word and
// read j+3's index load while read j votes (64 windows per lookup round).
__global__ __launch_bounds__(256) void k_vote_pipe(ReadsView rv, int k, int canonical,
                                                   const uint64_t* __restrict__ slots, uint64_t bmask,
                                                   tk_r2t_vote* __restrict__ out, uint32_t* tickets, uint32_t grab) {
  __shared__ uint32_t s_vote[4][192];  // per wave: 128 hits, 64 lookup scratch
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  uint32_t* hits = s_vote[wv];
  uint32_t* tmp = s_vote[wv] + 128;
  const uint64_t occ = 1ull << (2 * k);
  const uint64_t per_xcd = (rv.n_reads + 7) >> 3;
  const uint64_t x = blockIdx.x & 7;
  const uint64_t lo = x * per_xcd;
  const uint64_t hi = lo + per_xcd < rv.n_reads ? lo + per_xcd : rv.n_reads;
  if (lo >= hi) return;
  const uint32_t n = (uint32_t)(hi - lo);
  TicketStream ts(tickets + x * kTicketStride, n, grab, lane);
  bool va = false, vb = false, vc = false;
  uint32_t ra = 0, rb = 0, rc = 0, Lb = 0, Lc = 0, mwc = 0;
  uint64_t wbb = 0, cwc = 0;
  for (;;) {
    const uint32_t p = ts.next();
    const bool vn = p < n;
    uint32_t rn = 0;
    if (vn) rn = rv.order ? rv.order[lo + p] : (uint32_t)(lo + p);
    uint32_t La = 0;
    uint64_t wba = 0;
    if (va) {
      La = rv.len[ra];
      wba = rv.word_off[ra];
    }
    uint64_t cwb = 0;
    uint32_t mwb = 0;
    if (vb && lane < 6 && (uint32_t)lane <= ((Lb + 31) >> 5)) {
      cwb = rv.codes[wbb + lane];
      mwb = rv.nmask[wbb + lane];
    }
    if (vc) {
      const uint32_t nwin = Lc >= (uint32_t)k ? Lc - k + 1 : 0;
      uint32_t nh = 0;
      for (uint32_t base = 0; base < nwin; base += 64) {
        const uint32_t pos = base + lane;
        uint64_t f = 0;
        bool valid = false;
        window_lanes(cwc, mwc, pos, k, f, valid);
        valid = valid && pos < nwin;
        const uint64_t key = valid ? canon_key(f, k, canonical) : 0;
        const uint32_t cmp = coop_cget64(slots, 0, bmask, occ, key, valid, tmp, lane);
        const bool has = cmp != 0xFFFFFFFFu;
        const uint64_t m = __ballot(has);
        if (has) hits[nh + __popcll(m & ((1ull << lane) - 1))] = cmp;
        nh += (uint32_t)__popcll(m);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      uint32_t best = 0xFFFFFFFFu, bestn = 0;
      uint64_t prev = 0;
      bool first = true;
      while (nh) {
        uint32_t cand = 0xFFFFFFFFu;
        for (uint32_t i = lane; i < nh; i += 64) {
          const uint32_t v = hits[i];
          if ((first || v > (uint32_t)prev) && v < cand) cand = v;
        }
        cand = wave_min_u32(cand);
        if (cand == 0xFFFFFFFFu) break;
        uint32_t cnt = 0;
        for (uint32_t i = lane; i < nh; i += 64) cnt += hits[i] == cand;
        cnt = wave_sum_u32(cnt);
        if (cnt > bestn) {
          bestn = cnt;
          best = cand;
        }
        prev = cand;
        first = false;
      }
      if (lane == 0) {
        tk_r2t_vote v;
        v.comp = best;
        v.hits = bestn;
        v.nwin = nwin;
        out[rc] = v;
      }
      wave_sync();
    }
    vc = vb;
    rc = rb;
    Lc = Lb;
    cwc = cwb;
    mwc = mwb;
    vb = va;
    rb = ra;
    Lb = La;
    wbb = wba;
    va = vn;
    ra = rn;
    if (!va && !vb && !vc) break;
  }
}

// Grid of the per-read table passes (k_stats, k_vote): at most TK_READ_BLOCKS_PER_CU
// (default 128, the fastest of 4..128 and unbounded; 0 = one block per wpb
// reads) blocks per CU.
static int per_read_grid(tk_ctx* ctx, uint64_t n_reads, int wpb) {
  const char* e = getenv("TK_READ_BLOCKS_PER_CU");
  const uint64_t per_cu = e ? strtoull(e, nullptr, 10) : 128;
  const uint64_t want = ceil_div(n_reads, (uint64_t)wpb);
  const uint64_t cap = per_cu ? (uint64_t)std::max(ctx->num_cu, 1) * per_cu : (uint64_t)0x7FFFFFF0;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(want, cap));
}

// Context flag block: 64 words of small counters / error bits, then the
// per-XCD ticket counters of the per-read passes.
constexpr size_t kFlagBytes = (64 + kTicketWords) * 4;

// Scoring reads of at most 128 windows with the software-pipelined kernel
// (TK_STATS_PIPE, default 1; 0 = k_stats_short, one read at a time).
static bool stats_pipe() {
  const char* e = getenv("TK_STATS_PIPE");
  return !e || e[0] != '0';
}

// Voting reads of at most 128 windows with the software-pipelined kernel
// (TK_VOTE_PIPE, default 1; 0 = k_vote, one read at a time).
static bool vote_pipe() {
  const char* e = getenv("TK_VOTE_PIPE");
  return !e || e[0] != '0';
}

// Reads per ticket grab in the per-read table passes (TK_READ_TICKET, default
// 4; 0 = each wave takes a fixed stride of the XCD's range instead).
static uint32_t read_ticket_grab() {
  const char* e = getenv("TK_READ_TICKET");
  if (!e) return 4;
  const unsigned long v = strtoul(e, nullptr, 10);
  return v > 1024 ? 1024u : (uint32_t)v;
}

static int grid_for_waves(tk_ctx* ctx, uint64_t n_items, int waves_per_block) {
  uint64_t want = ceil_div(n_items, (uint64_t)waves_per_block);
  uint64_t cap = (uint64_t)std::max(ctx->num_cu, 1) * 32;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(want, cap));
}

// LDS per wave for the per-read kernels: a value slab of `slab` u32 plus a
// 64 x u64 lookup scratch; at most 64 KiB per block.  Returns waves per block.
static int slab_layout(uint32_t max_nwin, uint32_t* slab, int* wpb) {
  uint32_t s = std::max<uint32_t>(64, (max_nwin + 63) / 64 * 64);
  if ((uint64_t)s * 4 + 8 * kLookupScratch > 65536) return -1;
  int w = 4;
  while (w > 1 && ((uint64_t)s * 4 + 8 * kLookupScratch) * w > 65536) w >>= 1;
  *slab = s;
  *wpb = w;
  return 0;
}

}  // namespace tk

using namespace tk;

// =====================================================================
// C ABI
// =====================================================================
namespace tk {
// Device scratch of the allocation-free pair selection
// (tk_select_pairs_d_noalloc): 256 bytes of flags, the consumes and idx words
// of up to kNoAllocPairs pairs, and the scan's temporary storage.  Each GPU
// the library runs on holds one copy until the process ends.  Host code never
// uses the symbol's name as a device address: tk_init takes the device
// address with hipGetSymbolAddress once per device (g_scratch_ptr).
constexpr uint64_t kNoAllocPairs = 1ull << 24;
constexpr size_t kNoAllocScanBytes = 4u << 20;
constexpr size_t kNoAllocBytes = 256 + 2 * kNoAllocPairs * 4 + kNoAllocScanBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];  // per device, set once with that device current

static int scratch_setup(int device) {
  if (device < 0 || device >= 64) return fail(TK_ERR_ARG, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return TK_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = p;
  return TK_OK;
}

static void* scratch_ptr(int device) {
  if (device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace tk

extern "C" {

const char* tk_last_error(void) { return tk::g_err.c_str(); }
int tk_abi_version(void) { return TK_ABI_VERSION; }

int tk_init(int device, tk_ctx** out) {
  if (!out) return fail(TK_ERR_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(TK_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(TK_ERR_ARG, "device %d out of range [0,%d)", device, n);
  TK_HIP(hipSetDevice(device));
  if (int rc0 = scratch_setup(device)) return rc0;
  tk_ctx* c = new tk_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
  int rc = dev_alloc((void**)&c->d_flags, kFlagBytes);
  if (rc) { (void)hipStreamDestroy(c->stream); delete c; return rc; }
  c->d_tickets = c->d_flags + 64;
  e = hipHostMalloc((void**)&c->h_flags, 64 * 4, hipHostMallocDefault);
  if (e != hipSuccess) { dev_free(c->d_flags, kFlagBytes); (void)hipStreamDestroy(c->stream); delete c; return hip_fail(e, "hipHostMalloc"); }
  *out = c;
  return TK_OK;
}

int tk_ctx_free(tk_ctx* c) {
  if (!c) return TK_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (auto& s : c->scratch) dev_free(s.p, s.bytes);
  for (int i = 0; i < 2; i++) {
    if (c->stage[i]) (void)hipHostFree(c->stage[i]);
    if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
  }
  dev_free(c->d_flags, kFlagBytes);
  (void)hipHostFree(c->h_flags);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return TK_OK;
}

int tk_sync(tk_ctx* c) {
  if (!c) return fail(TK_ERR_ARG, "ctx is NULL");
  TK_HIP(hipStreamSynchronize(c->stream));
  return collect_timing(c);
}

int tk_count_words(tk_ctx* c, uint64_t* words) {
  if (!c || !words) return fail(TK_ERR_ARG, "NULL argument");
  *words = c->count_words;
  return TK_OK;
}

int tk_timing_enable(tk_ctx* c, int on) {
  if (!c) return fail(TK_ERR_ARG, "ctx is NULL");
  c->timing = on != 0;
  return TK_OK;
}

int tk_timing_reset(tk_ctx* c) {
  if (!c) return fail(TK_ERR_ARG, "ctx is NULL");
  int rc = collect_timing(c);
  c->times.clear();
  return rc;
}

int tk_timing_get(tk_ctx* c, const char* kernel, double* total_ms, uint64_t* launches) {
  if (!c || !kernel) return fail(TK_ERR_ARG, "NULL argument");
  int rc = collect_timing(c);
  if (rc) return rc;
  auto it = c->times.find(kernel);
  if (total_ms) *total_ms = it == c->times.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c->times.end() ? 0 : it->second.second;
  return TK_OK;
}

int tk_device_bytes_in_use(uint64_t* bytes) {
  if (!bytes) return fail(TK_ERR_ARG, "NULL argument");
  *bytes = g_dev_bytes.load();
  return TK_OK;
}

int tk_thread_cache(uint64_t limit_bytes) {
  if (limit_bytes)
    cache_begin((size_t)limit_bytes);
  else
    cache_end();
  return TK_OK;
}

// ---------------- plain buffers ----------------
static std::mutex g_buf_mu;
static std::map<void*, uint64_t> g_bufs;

int tk_host_alloc(uint64_t bytes, void** h) {
  if (!h) return fail(TK_ERR_ARG, "NULL argument");
  *h = nullptr;
  hipError_t e = hipHostMalloc(h, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    *h = nullptr;
    return hip_fail(e, "hipHostMalloc");
  }
  return TK_OK;
}

int tk_host_free(void* h) {
  if (!h) return TK_OK;
  TK_HIP(hipHostFree(h));
  return TK_OK;
}

int tk_buffer_alloc(tk_ctx* c, uint64_t bytes, void** d) {
  if (!c || !d) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = dev_alloc(d, bytes);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_buf_mu);
  g_bufs[*d] = bytes;
  return TK_OK;
}

int tk_buffer_free(tk_ctx* c, void* d) {
  if (!c) return fail(TK_ERR_ARG, "NULL argument");
  if (!d) return TK_OK;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> g(g_buf_mu);
    auto it = g_bufs.find(d);
    if (it == g_bufs.end()) return fail(TK_ERR_ARG, "pointer not from tk_buffer_alloc");
    bytes = it->second;
    g_bufs.erase(it);
  }
  if (set_device(c->device)) return TK_ERR_HIP;
  TK_HIP(hipStreamSynchronize(c->stream));
  dev_free(d, bytes);
  return TK_OK;
}

static int upload_h2d(tk_ctx* c, void* dst, const void* src, uint64_t bytes);

int tk_memcpy_h2d(tk_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = upload_h2d(c, dst, src, bytes);
  if (rc) return rc;
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

int tk_memcpy_d2h(tk_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  if (bytes) TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

int tk_memcpy_d2d(tk_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  hipEvent_t a;
  timing_begin(c, &a);
  if (bytes) TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  timing_end(c, "memcpy_d2d", a);
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

// ---------------- reads ----------------
constexpr uint64_t kStageBytes = 16ull << 20;

// Host-to-device copy on the context's stream.  A page-locked source moves
// by one DMA.  A pageable source is staged through the context's two pinned
// buffers, chunk by chunk: the host copy of one chunk overlaps the DMA of
// the other, and threads with contexts of their own stage in parallel.
static int upload_h2d(tk_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!bytes) return TK_OK;
  hipPointerAttribute_t attr;
  bool pinned = hipPointerGetAttributes(&attr, src) == hipSuccess && attr.type == hipMemoryTypeHost;
  (void)hipGetLastError();
  if (pinned || bytes <= kStageBytes) {
    TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return TK_OK;
  }
  for (int i = 0; i < 2; i++) {
    if (!c->stage[i]) TK_HIP(hipHostMalloc(&c->stage[i], kStageBytes, hipHostMallocDefault));
    if (!c->stage_ev[i]) TK_HIP(hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
  }
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  for (uint64_t off = 0, i = 0; off < bytes; off += kStageBytes, i ^= 1) {
    const uint64_t n = std::min<uint64_t>(kStageBytes, bytes - off);
    TK_HIP(hipEventSynchronize(c->stage_ev[i]));  // the DMA that last read this buffer is done
    memcpy(c->stage[i], s + off, n);
    TK_HIP(hipMemcpyAsync(d + off, c->stage[i], n, hipMemcpyHostToDevice, c->stream));
    TK_HIP(hipEventRecord(c->stage_ev[i], c->stream));
  }
  return TK_OK;
}

int tk_reads_from_host(tk_ctx* c, const char* seq, const uint64_t* offsets, uint64_t n,
                       tk_reads** out) {
  if (!c || !offsets || !out || (!seq && n && offsets[n] > offsets[0]))
    return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (offsets[0] != 0) return fail(TK_ERR_ARG, "offsets[0] must be 0");
  for (uint64_t i = 0; i < n; i++)
    if (offsets[i + 1] < offsets[i]) return fail(TK_ERR_ARG, "offsets not monotone at %llu", (unsigned long long)i);
  if (set_device(c->device)) return TK_ERR_HIP;
  const uint64_t nb = offsets[n];
  char* d_seq = nullptr;
  uint64_t* d_off = nullptr;
  int rc = dev_alloc((void**)&d_seq, nb + 64);
  if (rc) return rc;
  rc = dev_alloc((void**)&d_off, (n + 1) * 8);
  if (rc) { dev_free(d_seq, nb + 64); return rc; }
  rc = upload_h2d(c, d_seq, seq, nb);
  if (!rc) rc = upload_h2d(c, d_off, offsets, (n + 1) * 8);
  if (!rc) rc = reads_build(c, d_seq, d_off, n, nb, out);
  (void)hipStreamSynchronize(c->stream);
  dev_free(d_seq, nb + 64);
  dev_free(d_off, (n + 1) * 8);
  return rc;
}

int tk_reads_from_device_d(tk_ctx* c, const char* d_seq, const uint64_t* d_offsets, uint64_t n,
                           tk_reads** out) {
  if (!c || !d_offsets || !out) return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (set_device(c->device)) return TK_ERR_HIP;
  uint64_t nb = 0;
  TK_HIP(hipMemcpyAsync(&nb, d_offsets + n, 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  return reads_build(c, d_seq, d_offsets, n, nb, out);
}

int tk_reads_info(const tk_reads* r, uint64_t* n_reads, uint64_t* n_bases, uint32_t* max_len) {
  if (!r) return fail(TK_ERR_ARG, "reads is NULL");
  if (n_reads) *n_reads = r->n_reads;
  if (n_bases) *n_bases = r->n_bases;
  if (max_len) *max_len = r->max_len;
  return TK_OK;
}

int tk_reads_windows(const tk_reads* r, int k, uint64_t* total) {
  if (!r || !total) return fail(TK_ERR_ARG, "NULL argument");
  if (k < 1 || k > 31) return fail(TK_ERR_ARG, "k=%d outside [1,31]", k);
  uint64_t t = 0;
  // lengths are on the device; a throwaway context is cheap compared to the pass
  tk_ctx* c = nullptr;
  int rc = tk_init(r->device, &c);
  if (rc) return rc;
  rc = reads_windows(c, r, k, &t);
  tk_ctx_free(c);
  *total = t;
  return rc;
}

int tk_reads_free(tk_reads* r) {
  if (!r) return TK_OK;
  (void)hipSetDevice(r->device);
  dev_free(r->d_len, std::max<uint64_t>(r->n_reads, 1) * 4);
  dev_free(r->d_word_off, (r->n_reads + 1) * 8);
  if (r->d_codes) dev_free(r->d_codes, (r->n_words + 1) * 8);
  if (r->d_nmask) dev_free(r->d_nmask, (r->n_words + 1) * 4);
  if (r->d_order) dev_free(r->d_order, r->n_reads * 4);
  delete r;
  return TK_OK;
}

// ---------------- table ----------------
static int hll_estimate(tk_ctx* c, const tk_reads* r, uint64_t rb, uint64_t re, int k, int canonical,
                        uint64_t windows, double* est);

int tk_estimate_distinct(tk_ctx* c, const tk_reads* r, int k, int canonical, double* est) {
  if (!c || !r || !est) return fail(TK_ERR_ARG, "NULL argument");
  if (k < 1 || k > 31) return fail(TK_ERR_ARG, "k=%d outside [1,31]", k);
  if (set_device(c->device)) return TK_ERR_HIP;
  uint64_t total = 0;
  int rc = reads_windows(c, r, k, &total);
  if (rc) return rc;
  rc = hll_estimate(c, r, 0, r->n_reads, k, canonical, total, est);
  if (!rc) const_cast<tk_reads*>(r)->est[2 * k + (canonical ? 1 : 0)] = *est;
  return rc;
}

int tk_table_create(tk_ctx* c, int k, int canonical, uint64_t expected, double max_load,
                    tk_table** out) {
  if (!c || !out) return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (k < 1 || k > 31) return fail(TK_ERR_ARG, "k=%d outside [1,31]", k);
  if (!(max_load > 0.0 && max_load <= 0.9)) return fail(TK_ERR_ARG, "max_load must be in (0,0.9]");
  if (set_device(c->device)) return TK_ERR_HIP;
  tk_table* t = new tk_table();
  t->device = c->device;
  t->k = k;
  t->canonical = canonical ? 1 : 0;
  t->max_load = max_load;
  table_layout(std::max<uint64_t>(expected, 64), max_load, &t->nb, &t->rbits);
  // the slots are cleared on first use (the region-resident count of an
  // empty table writes every region instead)
  int rc = table_alloc_slots(c, t->nb, &t->d_slots, false);
  t->zero_pending = true;
  if (!rc) rc = table_ensure_ovf(c, t, 0);
  if (!rc) { hipError_t e = hipStreamSynchronize(c->stream); if (e != hipSuccess) rc = hip_fail(e, "table init"); }
  if (rc) { tk_table_free(t); return rc; }
  t->bytes = t->nb * 64 + t->ovf_cap * 16;
  *out = t;
  return TK_OK;
}

int tk_table_info(const tk_table* t, int* k, int* canonical, uint64_t* capacity, uint64_t* bytes) {
  if (!t) return fail(TK_ERR_ARG, "table is NULL");
  if (k) *k = t->k;
  if (canonical) *canonical = t->canonical;
  if (capacity) *capacity = t->nb * 8;
  if (bytes) *bytes = t->bytes;
  return TK_OK;
}

// HyperLogLog estimate of the distinct keys of reads [rb, re).  Large inputs
// (>= 64M windows) update the registers from a 1/16 key sample.
static int hll_estimate(tk_ctx* c, const tk_reads* r, uint64_t rb, uint64_t re, int k, int canonical,
                        uint64_t windows, double* est) {
  void* p = nullptr;
  int rc = ensure_scratch(c, 0, kHllRegs * 4, &p);
  if (rc) return rc;
  uint32_t* regs = (uint32_t*)p;
  const uint32_t sbits = windows >= (64ull << 20) ? 4 : 0;
  TK_HIP(hipMemsetAsync(regs, 0, kHllRegs * 4, c->stream));
  // a whole-read-set estimate also yields the locality sort keys
  tk_reads* rw = const_cast<tk_reads*>(r);
  uint64_t* keys = nullptr;
  uint32_t* order = nullptr;
  if (rb == 0 && re == r->n_reads && order_wanted(r)) {
    if ((rc = dev_alloc((void**)&keys, r->n_reads * 8))) return rc;
    if ((rc = dev_alloc((void**)&order, r->n_reads * 4))) { dev_free(keys, r->n_reads * 8); return rc; }
  }
  if (re > rb) {
    // blocks per CU (TK_HLL_BLOCKS_PER_CU, 64 default): each block ends with one
    // atomicMax per register into the global registers
    const char* hb = getenv("TK_HLL_BLOCKS_PER_CU");
    const uint64_t per_cu = hb && atoi(hb) > 0 ? (uint64_t)atoi(hb) : 64;
    const uint64_t grid = std::min<uint64_t>(ceil_div(re - rb, 256), (uint64_t)std::max(c->num_cu, 1) * per_cu);
    hipEvent_t a;
    timing_begin(c, &a);
    hipLaunchKernelGGL(k_hll, dim3(grid), dim3(256), 0, c->stream, reads_view(r), rb, re, k, canonical,
                       sbits, regs, keys, order);
    timing_end(c, "hll", a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      if (keys) { dev_free(keys, r->n_reads * 8); dev_free(order, r->n_reads * 4); }
      return hip_fail(e, "k_hll");
    }
  }
  if (keys && (rc = order_finish(c, rw, keys, order))) return rc;
  std::vector<uint32_t> h(kHllRegs);
  TK_HIP(hipMemcpyAsync(h.data(), regs, kHllRegs * 4, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  const double m = kHllRegs;
  double sum = 0;
  int zeros = 0;
  for (uint32_t v : h) {
    sum += std::ldexp(1.0, -(int)v);
    zeros += v == 0;
  }
  const double alpha = 0.7213 / (1.0 + 1.079 / m);
  double e = alpha * m * m / sum;
  if (e <= 2.5 * m && zeros) e = m * std::log(m / zeros);
  *est = std::ldexp(e, (int)sbits);
  return TK_OK;
}

// TK_PARTITION=1 counts large tables in region-partitioned passes (off by
// default: measured slower than the direct count on the 20M-pair workload).
static bool partition_enabled() {
  const char* e = getenv("TK_PARTITION");
  return e && e[0] == '1';
}

__global__ void k_add_keys(const uint64_t* __restrict__ keys, uint64_t n, TableView t,
                           uint32_t* __restrict__ flags) {
  uint32_t mynew = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = table_add(t, keys[i], 1, &flags[0]);
    if (r == 2u) atomicOr(&flags[0], 1u);
    else mynew += r;
  }
  mynew = wave_sum_u32(mynew);
  if ((threadIdx.x & 63) == 0 && mynew)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)mynew);
}

// Insert keys marked with bit 63 (clearing the mark on success).
__global__ void k_retry_marked(uint64_t* __restrict__ keys, uint64_t n, TableView t,
                               uint32_t* __restrict__ flags) {
  uint32_t mynew = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = keys[i];
    if (!(v >> 63)) continue;
    const uint32_t r = table_add(t, v & ~(1ull << 63), 1, &flags[0]);
    if (r == 2u) {
      atomicOr(&flags[6], 1u);
    } else {
      mynew += r;
      keys[i] = v & ~(1ull << 63);
    }
  }
  mynew = wave_sum_u32(mynew);
  if ((threadIdx.x & 63) == 0 && mynew)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)mynew);
}

// Count reads [rb, rb+nreads) through the region-partitioned passes.
static int count_partitioned(tk_ctx* c, tk_table* t, const tk_reads* r, uint64_t rb, uint64_t nreads,
                             uint64_t windows_bound) {
  int rc0 = TK_OK;
  // When the chunk could overfill the table, size it from an estimate of the
  // chunk's distinct keys first, so that the regions keys are grouped by are
  // the final ones.
  if ((double)(t->n_keys + windows_bound) > 0.9 * (double)(t->nb * 8)) {
    double est = 0;
    if ((rc0 = hll_estimate(c, r, rb, rb + nreads, t->k, t->canonical, windows_bound, &est))) return rc0;
    const double want = (double)t->n_keys + 1.05 * est + 4096;
    if (want > t->max_load * (double)(t->nb * 8))
      if ((rc0 = table_grow(c, t, (uint64_t)want))) return rc0;
  }
  const uint32_t nparts = (uint32_t)(t->nb >> t->rbits);
  if (nparts > kMaxRegions) return fail(TK_ERR_ARG, "table has %u regions (limit %u)", nparts, kMaxRegions);
  const uint64_t nblocks = ceil_div(nreads, kPartReads);
  if (nblocks * nparts > 0x7FFFFFFFull) return fail(TK_ERR_ARG, "partition histogram too large");
  const uint64_t nh = nblocks * nparts;
  uint64_t *hist = nullptr, *base = nullptr, *keys = nullptr;
  int rc = dev_alloc((void**)&hist, (nh + 1) * 8);
  if (rc) return rc;
  if ((rc = dev_alloc((void**)&base, (nh + 1) * 8))) { dev_free(hist, (nh + 1) * 8); return rc; }
  TableView v = table_view(t);
  uint64_t nkeys = 0;
  do {
    hipError_t e = hipMemsetAsync(hist + nh, 0, 8, c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "memset"); break; }
    {
      hipEvent_t a;
      timing_begin(c, &a);
      hipLaunchKernelGGL(k_part_hist, dim3(nblocks), dim3(256), 0, c->stream, reads_view(r), rb,
                         rb + nreads, t->k, t->canonical, v, nparts, (uint32_t)nblocks, hist);
      timing_end(c, "part_hist", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_part_hist"); break; }
    }
    if ((rc = exclusive_scan_u64(c, hist, base, nh + 1))) break;
    e = hipMemcpyAsync(&nkeys, base + nh, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "part total"); break; }
    if (!nkeys) break;
    if ((rc = dev_alloc((void**)&keys, nkeys * 8))) break;
    {
      hipEvent_t a;
      timing_begin(c, &a);
      hipLaunchKernelGGL(k_part_scatter, dim3(nblocks), dim3(256), 0, c->stream, reads_view(r), rb,
                         rb + nreads, t->k, t->canonical, v, nparts, (uint32_t)nblocks, base, keys);
      timing_end(c, "part_scatter", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_part_scatter"); break; }
    }
  } while (0);
  (void)hipStreamSynchronize(c->stream);
  dev_free(hist, (nh + 1) * 8);
  dev_free(base, (nh + 1) * 8);
  if (rc || !nkeys) {
    if (keys) dev_free(keys, nkeys * 8);
    return rc;
  }
  // pass 2
  const uint64_t spill_cap = std::max<uint64_t>(1 << 20, std::min<uint64_t>(nkeys / 16, 64ull << 20));
  void* sp = nullptr;
  if ((rc = ensure_scratch(c, 1, spill_cap * 8, &sp))) { dev_free(keys, nkeys * 8); return rc; }
  TableView v2 = v;
  v2.max_probe = std::min<uint32_t>(v.max_probe, 32);
  hipError_t e = hipMemsetAsync(c->d_flags, 0, 32, c->stream);
  if (e == hipSuccess) {
    hipEvent_t a;
    timing_begin(c, &a);
    hipLaunchKernelGGL(k_part_count, dim3(ceil_div(nkeys, kPartKeysPerBlock)), dim3(256), 0, c->stream,
                       keys, nkeys, v2, (uint64_t*)sp, spill_cap, c->d_flags);
    timing_end(c, "part_count", a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(c->h_flags, c->d_flags, 32, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { dev_free(keys, nkeys * 8); return hip_fail(e, "k_part_count"); }
  uint64_t nnew = 0, nspill = 0;
  memcpy(&nnew, &c->h_flags[2], 8);
  memcpy(&nspill, &c->h_flags[4], 8);
  bool marked = c->h_flags[6] != 0;
  t->n_keys += nnew;
  if (c->h_flags[0]) { dev_free(keys, nkeys * 8); return fail(TK_ERR_FULL, "count: table error flags=%u", c->h_flags[0]); }
  // Crowded regions: grow, add the spilled occurrences, then retry marked
  // keys (growing again while any still finds no room).
  for (int round = 0; (nspill || marked) && !rc; round++) {
    if (round > 8) { rc = fail(TK_ERR_FULL, "count: keys still find no room after growing"); break; }
    const uint64_t pending = std::min(nspill, spill_cap) + (marked ? nkeys : 0);
    if ((rc = table_grow(c, t, t->n_keys + std::min<uint64_t>(pending, nkeys)))) break;
    e = hipMemsetAsync(c->d_flags, 0, 32, c->stream);
    TableView v3 = table_view(t);
    if (e == hipSuccess && nspill) {
      const uint64_t ns = std::min(nspill, spill_cap);
      hipLaunchKernelGGL(k_add_keys, dim3(std::min<uint64_t>(ceil_div(ns, 256), 4096)), dim3(256), 0,
                         c->stream, (const uint64_t*)sp, ns, v3, c->d_flags);
      e = hipGetLastError();
    }
    if (e == hipSuccess && marked) {
      hipLaunchKernelGGL(k_retry_marked, dim3(std::min<uint64_t>(ceil_div(nkeys, 256), 65536)), dim3(256), 0,
                         c->stream, keys, nkeys, v3, c->d_flags);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_flags, c->d_flags, 32, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "spill retry"); break; }
    if (c->h_flags[0] & 1u) { rc = fail(TK_ERR_FULL, "count: spilled keys found no room"); break; }
    memcpy(&nnew, &c->h_flags[2], 8);
    t->n_keys += nnew;
    nspill = 0;
    marked = c->h_flags[6] != 0;
  }
  dev_free(keys, nkeys * 8);
  return rc;
}

// LDS words of the block-level aggregation of the ordered count that adds its
// merged words with table atomics: 4096 (32 KB, windows merged 2048 at a
// time, four blocks per CU, which the kernel's registers allow anyway).  A
// 64 KB table merged more per pass but halved the blocks per CU and measured
// slower.
constexpr int kAggSlots = 4096;
// LDS words of the merge table of the region-resident count: 3840 (30 KB,
// five blocks per CU).  3,328 slots (more flushes) and 8,192 (two blocks per
// CU) measured slower.
constexpr int kResSlots = 3840;

// Whether a count of `windows` windows takes the region-resident passes
// (reads in the locality order, k <= 25, tables of 64 KB regions).  By
// default only counts of >= 2^32 windows do: their word streams are tens of
// GB held by the context, and on smaller jobs the gain does not pay for the
// memory (a context that holds them leaves less room for the per-job tables
// of a sample stream, whose allocations then stall).  TK_RESIDENT=0 never,
// TK_RESIDENT=1 always (when the streams fit).
static bool resident_enabled(uint64_t windows) {
  const char* e = getenv("TK_RESIDENT");
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  return windows >= (1ull << 32);
}

// The kernel attributes below are set once per process, by whichever thread
// gets here first (a function-local static is initialised under the
// language's own once-only guard, so server workers may race to it).
static bool resident_usable(tk_ctx* c, const tk_table* t) {
  if (t->rbits > kResRegionBitsMax || (t->nb >> t->rbits) > (uint64_t)kResMaxBins * kResMaxBins) return false;
  // 64 KB or more of dynamic LDS per workgroup needs the attribute
  static const bool ok = [] {
    const bool r =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_count<true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 72 << kResRegionBitsMax) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_count<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 72 << kResRegionBitsMax) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_scatter<8192, 512>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)res_scatter_lds(kResMaxBins, 8192)) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_scatter<4096, 256>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)res_scatter_lds(kResMaxBins, 4096)) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_scatter<2048, 256>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)res_scatter_lds(kResMaxBins, 2048)) == hipSuccess;
    (void)hipGetLastError();
    return r;
  }();
  (void)c;
  return ok;
}

// Words per LDS tile of the region sort's scatter: TK_RES_TILE=8192
// (default; 512 threads per block, two blocks per CU), 16384 (1,024
// threads, one block per CU, when its LDS can be granted: slower, its
// phases no longer overlap another block's), 4096 or 2048 (256 threads).  Larger tiles leave longer runs per bin: 8,192 words over
// 512 bins are 128 B per run on average.
static bool res_tile_16k_ok() {
  static const bool ok = [] {
    const bool r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_scatter<16384, 1024>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)res_scatter_lds(kResMaxBins, 16384)) == hipSuccess;
    (void)hipGetLastError();
    return r;
  }();
  return ok;
}
static uint32_t res_tile() {
  const char* e = getenv("TK_RES_TILE");
  const int v = e ? atoi(e) : 0;
  if (v == 16384 && res_tile_16k_ok()) return 16384;
  return v == 2048 || v == 4096 ? (uint32_t)v : 8192;
}

// Scratch slots of the region-resident count (kept by the context, so the
// multi-GB word streams are allocated once, not per count call).
enum { kSlotResA = 4, kSlotResB = 5, kSlotResHist = 6, kSlotResMisc = 7 };

// One counting-sort pass of the region-resident count: the nonzero words of
// `in` (slice table dsl, copied from sl) to `out`, grouped by bin
// (region >> shift) & mask.  h holds nhist + 1 counts, hs their scan.
static int res_sort_pass(tk_ctx* c, const std::vector<ResSlice>& sl, ResSlice* dsl, uint64_t nhist,
                         uint64_t* h, uint64_t* hs, const uint64_t* in, uint64_t* out, const TableView& v,
                         uint32_t shift, uint32_t mask, uint32_t nbins) {
  TK_HIP(hipMemcpyAsync(dsl, sl.data(), sl.size() * sizeof(ResSlice), hipMemcpyHostToDevice, c->stream));
  TK_HIP(hipMemsetAsync(h + nhist, 0, 8, c->stream));
  TK_LAUNCH(c, "count_hist", sl.size(), 256, 0, k_res_hist, in, dsl, v, shift, mask, nbins, h);
  int rc = exclusive_scan_u64(c, h, hs, nhist + 1);
  if (rc) return rc;
  uint32_t tile = res_tile();
  if (tile == 16384 && res_scatter_lds(nbins, 16384) + 1024 > 160 * 1024) tile = 8192;  // LDS of one CU
  if (tile == 2048)
    TK_LAUNCH(c, "count_scatter", sl.size(), 256, res_scatter_lds(nbins, 2048), (k_res_scatter<2048, 256>), in,
              dsl, v, shift, mask, nbins, hs, out);
  else if (tile == 8192)
    TK_LAUNCH(c, "count_scatter", sl.size(), 512, res_scatter_lds(nbins, 8192), (k_res_scatter<8192, 512>), in,
              dsl, v, shift, mask, nbins, hs, out);
  else if (tile == 16384)
    TK_LAUNCH(c, "count_scatter", sl.size(), 1024, res_scatter_lds(nbins, 16384), (k_res_scatter<16384, 1024>),
              in, dsl, v, shift, mask, nbins, hs, out);
  else
    TK_LAUNCH(c, "count_scatter", sl.size(), 256, res_scatter_lds(nbins, 4096), (k_res_scatter<4096, 256>), in,
              dsl, v, shift, mask, nbins, hs, out);
  // the slice table is host memory read by the copy above
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

// Windows per thread between the merge's block barriers (TK_MERGE_WPR = 2,
// the default, or 1; the 3,840-slot merge of k >= 16 only).  Two halve the
// barriers and shared-counter atomics of a latency-bound loop (36.4 vs
// 40.2 ms on 5.04 G windows) for 1.4% more merged words (earlier flushes).
static int merge_wpr() {
  const char* e = getenv("TK_MERGE_WPR");
  return e && atoi(e) == 1 ? 1 : 2;
}

// Words a merge block reserves in the stream at a time, for a pass over
// `windows` windows with `mgrid` blocks.
static uint64_t res_echunk(uint64_t windows, uint64_t mgrid) {
  return std::min<uint64_t>(65536, std::max<uint64_t>(16384, next_pow2(windows / (std::max<uint64_t>(mgrid, 1) * 4))));
}

// Smallest streams a chunked region-resident count uses (2^26 words, 512 MB
// each); below that the table atomics count.
constexpr uint64_t kResMinStreamWords = 1ull << 26;

// Stream capacity of a region-resident count of `windows` windows: the
// words the merge can reserve (merged words <= windows, plus < one flush
// (at most `slots` words) given up per reserved chunk of `echunk`, plus each
// block's last chunk), so no merged word ever misses the stream.  False when that does not fit the
// device memory the count may use: the table atomics then count instead
// (the merge of a very large or error-rich read set leaves too many words).
static bool resident_plan(tk_ctx* c, uint64_t windows, uint64_t* cap, uint64_t* echunk, uint64_t* mgrid,
                          uint64_t* cap_max, uint64_t* cap_chunked = nullptr) {
  const uint64_t slots = (uint64_t)kResSlots;
  static const int pc = [] {  // resident merge blocks per CU (same for every device of the process)
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (k_count_agg<kResSlots, true, true, 2>),
                                                                      256, kResSlots * 8);
    if (e != hipSuccess || nb < 1) nb = 1;
    (void)hipGetLastError();
    return nb;
  }();
  *mgrid = (uint64_t)std::max(c->num_cu, 1) * pc;
  *echunk = res_echunk(windows, *mgrid);
  *cap = windows + windows / (*echunk / slots - 1) + *mgrid * (*echunk + slots);
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  // this context's own resident scratch counts as available, and so do the
  // thread's idle cached blocks (dev_alloc hands them back when a stream
  // does not fit beside them)
  uint64_t held = t_cache ? t_cache->cached : 0;
  for (int i = kSlotResA; i <= kSlotResMisc; i++) held += c->scratch[i].bytes;
  const uint64_t budget = t_resident_limit ? t_resident_limit : (uint64_t)(0.4 * (double)(fr + held));
  *cap_max = budget / 17;  // two 8-B streams + histograms
  // a count in read chunks takes 60%: each chunk's region pass reads and
  // writes the whole table, so fewer, longer chunks pay off
  if (cap_chunked)
    *cap_chunked = (t_resident_limit ? t_resident_limit : (uint64_t)(0.6 * (double)(fr + held))) / 17;
  return *cap <= *cap_max;
}
constexpr int kNotResident = -1000;

// Count reads [rb, re) of the locality order through the region-resident
// passes: merge + emit, coarse sort, fine sort, per-region LDS count, in one
// pass whose streams hold every word the merge can emit (resident_plan);
// kNotResident, with the table unchanged, when they do not fit.
static int count_resident(tk_ctx* c, tk_table* t, const tk_reads* r, const ReadsView& rview, uint64_t rb,
                          uint64_t re, uint64_t windows, uint32_t max_nwin) {
  int rc = TK_OK;
  // the table's final regions must be known before words are grouped by them
  if ((double)(t->n_keys + windows) > 0.9 * (double)(t->nb * 8)) {
    double est = 0;
    auto it = r->est.find(2 * t->k + t->canonical);  // the caller's own estimate of these reads
    if (rb == 0 && re == r->n_reads && it != r->est.end())
      est = it->second;
    else if ((rc = hll_estimate(c, r, rb, re, t->k, t->canonical, windows, &est)))
      return rc;
    const double want = (double)t->n_keys + 1.05 * est + 4096;
    if (want > t->max_load * (double)(t->nb * 8)) {
      if ((rc = table_zero(c, t))) return rc;  // the rehash reads the old slots
      if ((rc = table_grow(c, t, (uint64_t)want))) return rc;
    }
    if (!resident_usable(c, t)) return fail(TK_ERR_ARG, "count: grown table has %llu-bucket regions",
                                            (unsigned long long)(1ull << t->rbits));
  }
  uint64_t cap = 0, echunk = 0, mgrid = 0, cap_max = 0, cap_chunked = 0;
  const bool whole = resident_plan(c, windows, &cap, &echunk, &mgrid, &cap_max, &cap_chunked);
  // TK_RES_CHUNK_READS=n: read chunks of n reads even when one pass fits
  // (tests), TK_RES_STREAM_WORDS=n: streams of at most n words.
  const char* ecr = getenv("TK_RES_CHUNK_READS");
  const char* esw = getenv("TK_RES_STREAM_WORDS");
  const uint64_t forced_reads = ecr ? strtoull(ecr, nullptr, 10) : 0;
  const uint64_t forced_words = esw ? strtoull(esw, nullptr, 10) : 0;
  const uint64_t n_all = re - rb;
  const double win_per_read = (double)windows / (double)std::max<uint64_t>(1, n_all);
  uint64_t chunk = n_all;
  double ratio = 2.5;  // windows per stream word, until a chunk has measured it
  const bool chunked = !whole || forced_reads || forced_words;
  if (chunked) {
    // The streams of every window do not fit (or a test asks for chunks):
    // count the reads in chunks of the locality order whose merged words
    // should fit streams of cap words, each chunk's region pass reading and
    // writing the table once.  A flush past the end of a chunk's stream is
    // added to the table with atomics, so a chunk that merges worse than
    // guessed costs time, not results.
    if (!whole) cap = cap_chunked;
    if (forced_words) cap = std::min(cap, forced_words);
    if (!forced_words && !forced_reads && cap < kResMinStreamWords) return kNotResident;
    chunk = forced_reads ? forced_reads : (uint64_t)(0.9 * ratio * (double)cap / std::max(win_per_read, 1.0));
    chunk = std::min<uint64_t>(n_all, std::max<uint64_t>(chunk, kReadsPerBlock));
    echunk = res_echunk((uint64_t)((double)chunk * win_per_read), mgrid);
    // A flush past a stream goes straight to the table, which must then be
    // clear: while it is not (a new table), the merge drops such a flush and
    // says so, and the chunk is merged again on the cleared table.  The first
    // chunk's region pass writes every region, so a chunk that fits its
    // stream costs no clearing pass.
  }
  void *pa = nullptr, *pb = nullptr;
  if ((rc = ensure_scratch(c, kSlotResA, cap * 8, &pa))) return rc;
  if ((rc = ensure_scratch(c, kSlotResB, cap * 8, &pb))) return rc;
  uint64_t* A = (uint64_t*)pa;
  uint64_t* B = (uint64_t*)pb;
  unsigned long long* cursor = reinterpret_cast<unsigned long long*>(c->d_flags + 8);
  for (uint64_t q = rb, qe = rb; q < re; q = qe) {
    qe = std::min<uint64_t>(re, q + chunk);
    const uint64_t keys_before = t->n_keys;
    TableView v = table_view(t);
    TK_HIP(hipMemsetAsync(c->d_flags, 0, 64, c->stream));
    // persistent: as many blocks per CU as its registers and LDS allow
    const uint64_t grid = std::max<uint64_t>(8, std::min<uint64_t>((ceil_div(qe - q, kReadsPerBlock) + 7) / 8 * 8, mgrid));
    const int table_ok = t->zero_pending ? 0 : 1;
    if (t->k >= 16 && merge_wpr() == 2)
      TK_LAUNCH(c, "count_merge", grid, 256, kResSlots * 8, (k_count_agg<kResSlots, true, true, 2>), rview, q, qe,
                t->k, t->canonical, v, c->d_flags, A, cap, echunk, cursor, table_ok);
    else if (t->k >= 16)
      TK_LAUNCH(c, "count_merge", grid, 256, kResSlots * 8, (k_count_agg<kResSlots, true, true>), rview, q, qe,
                t->k, t->canonical, v, c->d_flags, A, cap, echunk, cursor, table_ok);
    else
      TK_LAUNCH(c, "count_merge", grid, 256, kResSlots * 8, (k_count_agg<kResSlots, true, false>), rview, q, qe,
                t->k, t->canonical, v, c->d_flags, A, cap, echunk, cursor, table_ok);
    TK_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 64, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    if (c->h_flags[0]) return fail(TK_ERR_FULL, "count: table probe limit reached (flags=%u)", c->h_flags[0]);
    if (c->h_flags[6]) {  // a flush past the stream was dropped: clear the table, merge this chunk again
      if (!t->zero_pending) return fail(TK_ERR_STATE, "count: merge dropped words on a cleared table");
      if ((rc = table_zero(c, t))) return rc;
      qe = q;
      continue;
    }
    uint64_t nnew = 0, n = 0;
    memcpy(&nnew, &c->h_flags[2], 8);
    memcpy(&n, &c->h_flags[8], 8);
    if (t->zero_pending && nnew)  // the plan's stream holds every word the merge can emit
      return fail(TK_ERR_STATE, "count: merged words missed the stream of an uncleared table");
    t->n_keys += nnew;
    if (chunked && !forced_reads && qe < re && q == rb) {
      // the rest in equal chunks, from the first chunk's windows per stream
      // word (reserved words, so an underestimate of the merge)
      const double wq = (double)(qe - q) * win_per_read;
      ratio = std::max(1.0, std::min(64.0, wq / (double)std::max<uint64_t>(1, std::min(n, cap))));
      const double rest_words = (double)(re - qe) * win_per_read / ratio;
      const uint64_t parts = std::max<uint64_t>(1, (uint64_t)std::ceil(rest_words / (0.92 * (double)cap)));
      chunk = std::max<uint64_t>(ceil_div(re - qe, parts), kReadsPerBlock);
      echunk = res_echunk((uint64_t)((double)chunk * win_per_read), mgrid);
    }
    n = std::min(n, cap);
    if (!n) {
      if ((rc = table_zero(c, t))) return rc;
      continue;
    }
    // coarse bins of F = 2^fs regions, then single regions
    const uint32_t R = (uint32_t)(t->nb >> t->rbits);
    uint32_t fs = 0;
    while ((1ull << (2 * fs)) < R) fs++;
    const uint32_t F = 1u << fs, C = (uint32_t)ceil_div(R, F);
    const uint64_t ns1 = ceil_div(n, kResSlice);
    const uint64_t ns2_max = ns1 + C;  // sum over coarse bins of ceil(n_c / slice)
    const uint64_t nh1 = (uint64_t)C * ns1, nh_max = std::max<uint64_t>(nh1, (uint64_t)F * ns2_max);
    void *hp = nullptr, *mp = nullptr;
    const uint64_t sl_bytes = ns2_max * sizeof(ResSlice), hb_off = sl_bytes, m_off = hb_off + (uint64_t)C * 8,
                   rs_off = (m_off + (uint64_t)C * 4 + 7) / 8 * 8, cs_off = rs_off + ((uint64_t)R + 1) * 8,
                   misc_bytes = cs_off + ((uint64_t)C + 1) * 8;
    if ((rc = ensure_scratch(c, kSlotResHist, (nh_max + 1) * 16, &hp))) return rc;
    if ((rc = ensure_scratch(c, kSlotResMisc, misc_bytes, &mp))) return rc;
    uint64_t* h = (uint64_t*)hp;
    uint64_t* hs = h + nh_max + 1;
    ResSlice* dsl = (ResSlice*)mp;
    uint64_t* dhb = (uint64_t*)((char*)mp + hb_off);
    uint32_t* dm = (uint32_t*)((char*)mp + m_off);
    uint64_t* drs = (uint64_t*)((char*)mp + rs_off);
    uint64_t* dcs = (uint64_t*)((char*)mp + cs_off);
    std::vector<ResSlice> sl;
    sl.reserve(ns2_max);
    for (uint64_t j = 0; j < ns1; j++)
      sl.push_back({j * kResSlice, std::min<uint64_t>(n, (j + 1) * kResSlice), 0, (uint32_t)ns1, (uint32_t)j});
    if ((rc = res_sort_pass(c, sl, dsl, nh1, h, hs, A, B, v, fs, 0xFFFFFFFFu, C))) return rc;
    std::vector<uint64_t> cs(C + 1);
    TK_LAUNCH(c, "count_bounds", ceil_div((uint64_t)C + 1, 256), 256, 0, k_res_coarse_starts, hs, ns1, C, dcs);
    TK_HIP(hipMemcpyAsync(cs.data(), dcs, ((size_t)C + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    const uint64_t total = cs[C];
    c->count_words += total;
    std::vector<uint64_t> hb(C);
    std::vector<uint32_t> m(C);
    sl.clear();
    uint64_t nh2 = 0;
    for (uint32_t cc = 0; cc < C; cc++) {
      const uint64_t ncc = cs[cc + 1] - cs[cc];
      m[cc] = (uint32_t)std::max<uint64_t>(1, ceil_div(ncc, kResSlice));
      hb[cc] = nh2;
      for (uint32_t j = 0; j < m[cc]; j++)
        sl.push_back({cs[cc] + j * kResSlice, std::min<uint64_t>(cs[cc + 1], cs[cc] + (j + 1) * kResSlice), nh2,
                      m[cc], j});
      nh2 += (uint64_t)F * m[cc];
    }
    if (sl.size() > ns2_max) return fail(TK_ERR_HIP, "count: %zu fine slices (bound %llu)", sl.size(),
                                         (unsigned long long)ns2_max);
    TK_HIP(hipMemcpyAsync(dhb, hb.data(), (size_t)C * 8, hipMemcpyHostToDevice, c->stream));
    TK_HIP(hipMemcpyAsync(dm, m.data(), (size_t)C * 4, hipMemcpyHostToDevice, c->stream));
    if ((rc = res_sort_pass(c, sl, dsl, nh2, h, hs, B, A, v, 0, F - 1, F))) return rc;
    TK_LAUNCH(c, "count_bounds", ceil_div((uint64_t)R + 1, 256), 256, 0, k_res_bounds, hs, dhb, dm, fs, R, total,
              drs);
    TK_HIP(hipMemsetAsync(c->d_flags, 0, 64, c->stream));
    const size_t lds = (size_t)72 << t->rbits;  // slots + tags
    if (keys_before == 0 && nnew == 0)
      TK_LAUNCH(c, "count_regions", R, 1024, lds, k_res_count<true>, A, drs, v, c->d_flags,
                t->zero_pending ? 1 : 0);
    else
      TK_LAUNCH(c, "count_regions", R, 1024, lds, k_res_count<false>, A, drs, v, c->d_flags, 0);
    t->zero_pending = false;  // every region written (fresh) or the table was clear
    TK_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 64, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    if (c->h_flags[0]) return fail(TK_ERR_FULL, "count: table error (flags=%u)", c->h_flags[0]);
    memcpy(&nnew, &c->h_flags[2], 8);
    t->n_keys += nnew;
    // words of crowded regions: grow, then add them (growing again while any
    // still finds no room)
    for (int round = 0; c->h_flags[6]; round++) {
      if (round > 8) return fail(TK_ERR_FULL, "count: keys still find no room after growing");
      if ((rc = table_grow(c, t, t->n_keys + total))) return rc;
      TK_HIP(hipMemsetAsync(c->d_flags, 0, 64, c->stream));
      TableView v3 = table_view(t);
      TK_LAUNCH(c, "count_retry", std::min<uint64_t>(ceil_div(total, 256), 65536), 256, 0, k_res_retry, A, total,
                v3, c->d_flags);
      TK_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 64, hipMemcpyDeviceToHost, c->stream));
      TK_HIP(hipStreamSynchronize(c->stream));
      if (c->h_flags[0]) return fail(TK_ERR_FULL, "count: table error (flags=%u)", c->h_flags[0]);
      memcpy(&nnew, &c->h_flags[2], 8);
      t->n_keys += nnew;
    }
    if ((double)t->n_keys > t->max_load * (double)(t->nb * 8)) {
      if ((rc = table_grow(c, t, t->n_keys))) return rc;
      if (qe < re && !resident_usable(c, t))
        return fail(TK_ERR_ARG, "count: grown table has %llu-bucket regions", (unsigned long long)(1ull << t->rbits));
    }
  }
  return TK_OK;
}

int tk_count(tk_ctx* c, tk_table* t, const tk_reads* r) {
  if (!c || !t || !r) return fail(TK_ERR_ARG, "NULL argument");
  if (t->device != c->device || r->device != c->device) return fail(TK_ERR_ARG, "device mismatch");
  if (set_device(c->device)) return TK_ERR_HIP;
  c->count_words = 0;
  if (!r->n_reads) return TK_OK;
  // an uncleared (new, empty) table stays so here: the region-resident count
  // writes every region, any other path clears it first
  const bool unzeroed = t->zero_pending;
  t->zero_pending = false;
  int rc = table_set_final(c, t, false);
  t->zero_pending = unzeroed;
  if (rc) return rc;
  uint64_t total = 0;
  if ((rc = reads_windows(c, r, t->k, &total))) return rc;
  const uint32_t max_nwin = r->max_len >= (uint32_t)t->k ? r->max_len - t->k + 1 : 0;
  const uint64_t cbits = 63 - 2 * t->k;
  if ((rc = table_ensure_ovf(c, t, t->ovf_bound + (total >> cbits) + 1))) return rc;
  t->ovf_bound += (total >> cbits) + 1;
  ReadsView rview = reads_view(r);
  if (order_mode() == kOrderAll) {
    if ((rc = reads_order(c, r))) return rc;
    rview.order = r->d_order;
  } else {
    rview.order = nullptr;
  }
  if (rview.order && 63 - 2 * t->k >= 13 && resident_enabled(total) && max_nwin &&
      !partition_enabled() && resident_usable(c, t))
  {
    rc = count_resident(c, t, r, rview, 0, r->n_reads, total, max_nwin);
    if (rc != kNotResident) return rc;
  }
  if ((rc = table_zero(c, t))) return rc;
  uint64_t rb = 0;
  while (rb < r->n_reads) {
    const uint32_t nparts = (uint32_t)(t->nb >> t->rbits);
    if (partition_enabled() && nparts >= 2 && nparts <= kMaxRegions && max_nwin) {
      size_t fr = 0, tot = 0;
      TK_HIP(hipMemGetInfo(&fr, &tot));
      const uint64_t budget = std::min<uint64_t>((uint64_t)(0.45 * (double)fr), 96ull << 30);
      uint64_t nreads = std::max<uint64_t>(kPartReads, budget / 8 / max_nwin);
      nreads = std::min<uint64_t>(nreads, r->n_reads - rb);
      const uint64_t bound = (rb == 0 && nreads == r->n_reads) ? total : nreads * max_nwin;
      if ((rc = count_partitioned(c, t, r, rb, nreads, bound))) return rc;
      rb += nreads;
      if ((double)t->n_keys > t->max_load * (double)(t->nb * 8)) {
        if ((rc = table_grow(c, t, t->n_keys))) return rc;
      }
      continue;
    }
    // Read chunk whose windows cannot push the table past 0.9 load even if
    // every window were a new key.
    double room = 0.9 * (double)(t->nb * 8) - (double)t->n_keys;
    uint64_t nreads = max_nwin ? (uint64_t)(room / max_nwin) : r->n_reads;
    if (nreads < 64 && rb + nreads < r->n_reads) {
      if ((rc = table_grow(c, t, 2 * t->n_keys))) return rc;
      continue;
    }
    nreads = std::min<uint64_t>(nreads, r->n_reads - rb);
    nreads = std::min<uint64_t>(nreads, (uint64_t)0x7FFFFFF0 * kReadsPerBlock);
    TK_HIP(hipMemsetAsync(c->d_flags, 0, 16, c->stream));
    const uint64_t grid = (ceil_div(nreads, kReadsPerBlock) + 7) / 8 * 8;
    TableView v = table_view(t);
    // reads in the locality order: merge each block's windows in LDS first
    // (two windows per thread per round for k >= 16); else a window per thread
    const bool agg = rview.order && 63 - 2 * t->k >= 13;
    if (agg && t->k >= 16)
      TK_LAUNCH(c, "count", grid, 256, kAggSlots * 8, (k_count_agg<kAggSlots, false, true, 2>), rview, rb,
                rb + nreads, t->k, t->canonical, v, c->d_flags, nullptr, 0, 0, nullptr, 1);
    else if (agg)
      TK_LAUNCH(c, "count", grid, 256, kAggSlots * 8, (k_count_agg<kAggSlots, false, false>), rview, rb,
                rb + nreads, t->k, t->canonical, v, c->d_flags, nullptr, 0, 0, nullptr, 1);
    else
      TK_LAUNCH(c, "count", grid, 256, 0, k_count, rview, rb, rb + nreads, t->k, t->canonical, v, c->d_flags);
    TK_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 16, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    if (c->h_flags[0]) return fail(TK_ERR_FULL, "count: table probe limit reached (flags=%u)", c->h_flags[0]);
    uint64_t nnew = 0;
    memcpy(&nnew, &c->h_flags[2], 8);
    t->n_keys += nnew;
    rb += nreads;
    if ((double)t->n_keys > t->max_load * (double)(t->nb * 8)) {
      if ((rc = table_grow(c, t, t->n_keys))) return rc;
    }
  }
  return TK_OK;
}

int tk_table_load(tk_ctx* c, tk_table* t, const uint64_t* keys, const uint32_t* counts, uint64_t n) {
  if (!c || !t || (n && (!keys || !counts))) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  if (!n) return TK_OK;
  int rc = table_set_final(c, t, false);
  if (rc) return rc;
  // Keys that can need a side entry: at most one per loaded pair, and at
  // most one per 2^cbits of added count.
  const uint64_t cbits = 63 - 2 * t->k;
  unsigned __int128 tot = 0;
  for (uint64_t i = 0; i < n; i++) tot += counts[i];
  const unsigned __int128 carries = (tot >> cbits) + 1;
  uint64_t bound = carries < n ? (uint64_t)carries : n;
  if ((rc = table_ensure_ovf(c, t, t->ovf_bound + bound))) return rc;
  t->ovf_bound += bound;
  const uint64_t chunk_max = 1ull << 26;
  uint64_t* dk = nullptr;
  uint32_t* dc = nullptr;
  uint64_t cm = std::min<uint64_t>(n, chunk_max);
  if ((rc = dev_alloc((void**)&dk, cm * 8))) return rc;
  if ((rc = dev_alloc((void**)&dc, cm * 4))) { dev_free(dk, cm * 8); return rc; }
  uint64_t i = 0;
  while (i < n && !rc) {
    double room = 0.9 * (double)(t->nb * 8) - (double)t->n_keys;
    uint64_t m = std::min<uint64_t>(std::min<uint64_t>(cm, n - i), room > 0 ? (uint64_t)room : 0);
    if (m < std::min<uint64_t>(cm, n - i) / 2 || m == 0) {
      rc = table_grow(c, t, std::max<uint64_t>(t->n_keys + std::min<uint64_t>(cm, n - i), 2 * t->n_keys));
      continue;
    }
    hipError_t e = hipMemcpyAsync(dk, keys + i, m * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dc, counts + i, m * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_flags, 0, 16, c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "load upload"); break; }
    TableView v = table_view(t);
    {
      hipEvent_t a;
      timing_begin(c, &a);
      hipLaunchKernelGGL(k_load, dim3(std::min<uint64_t>(ceil_div(m, 256), 65536)), dim3(256), 0,
                         c->stream, dk, dc, m, v, c->d_flags);
      timing_end(c, "load", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_load"); break; }
    }
    e = hipMemcpyAsync(c->h_flags, c->d_flags, 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "load sync"); break; }
    if (c->h_flags[0] & 8u) { rc = fail(TK_ERR_ARG, "load: key wider than 2k bits"); break; }
    if (c->h_flags[0]) { rc = fail(TK_ERR_FULL, "load: table probe limit reached"); break; }
    uint64_t nnew = 0;
    memcpy(&nnew, &c->h_flags[2], 8);
    t->n_keys += nnew;
    i += m;
    if ((double)t->n_keys > t->max_load * (double)(t->nb * 8)) rc = table_grow(c, t, t->n_keys);
  }
  (void)hipStreamSynchronize(c->stream);
  dev_free(dk, cm * 8);
  dev_free(dc, cm * 4);
  return rc;
}

int tk_table_dump_size(tk_ctx* c, tk_table* t, uint32_t lo, uint32_t hi, uint64_t* n) {
  if (!c || !t || !n) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = table_set_final(c, t, true);
  if (rc) return rc;
  unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_flags + 8);
  TK_HIP(hipMemsetAsync(d, 0, 8, c->stream));
  TableView v = table_view(t);
  uint64_t ns = t->nb * 8;
  TK_LAUNCH(c, "dump_count", std::min<uint64_t>(ceil_div(ns, 256), 65536), 256, 0, k_dump_count,
            v, ns, (uint64_t)lo, (uint64_t)hi, d);
  uint64_t h = 0;
  TK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  *n = h;
  return TK_OK;
}

// Entries per sorted chunk of a dump: one sort call takes < 2^31 items, and a
// chunk holds 24 bytes per entry on the device (keys and counts, twice for
// the sort) within 40% of the free memory.  TK_DUMP_CHUNK sets a smaller
// limit (read per call), so that small tables take the multi-chunk path too.
static uint64_t dump_chunk_limit() {
  uint64_t lim = 0x7FFFFFFFull;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess)
    lim = std::min<uint64_t>(lim, std::max<uint64_t>(1u << 20, (uint64_t)(0.4 * (double)fr) / 24));
  else
    (void)hipGetLastError();
  if (const char* e = getenv("TK_DUMP_CHUNK")) {
    const uint64_t v = strtoull(e, nullptr, 10);
    if (v) lim = std::min(lim, v);
  }
  return lim;
}

// Collect, sort and download the entries with key in [klo, khi] (n of them)
// to keys/counts through the device buffers dk/dc.
static int dump_range(tk_ctx* c, tk_table* t, uint32_t lo, uint32_t hi, uint64_t klo, uint64_t khi, uint64_t n,
                      uint64_t* dk, uint32_t* dc, uint64_t* keys, uint32_t* counts) {
  unsigned long long* cur = reinterpret_cast<unsigned long long*>(c->d_flags + 8);
  TK_HIP(hipMemsetAsync(cur, 0, 8, c->stream));
  TableView v = table_view(t);
  const uint64_t ns = t->nb * 8;
  TK_LAUNCH(c, "dump", std::min<uint64_t>(ceil_div(ns, 256), 65536), 256, 0, k_dump, v, ns, (uint64_t)lo,
            (uint64_t)hi, klo, khi, dk, dc, n, cur);
  uint64_t got = 0;
  TK_HIP(hipMemcpyAsync(&got, cur, 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  if (got != n)
    return fail(TK_ERR_STATE, "dump: %llu entries in key range, %llu expected", (unsigned long long)got,
                (unsigned long long)n);
  int rc = sort_pairs_u64_u32(c, dk, dc, n, 2 * t->k);
  if (rc) return rc;
  TK_HIP(hipMemcpyAsync(keys, dk, n * 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipMemcpyAsync(counts, dc, n * 4, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

int tk_table_dump(tk_ctx* c, tk_table* t, uint32_t lo, uint32_t hi, uint64_t* keys,
                  uint32_t* counts, uint64_t cap, uint64_t* n_out) {
  if (!c || !t || !n_out || (cap && (!keys || !counts))) return fail(TK_ERR_ARG, "NULL argument");
  uint64_t n = 0;
  int rc = tk_table_dump_size(c, t, lo, hi, &n);
  if (rc) return rc;
  *n_out = n;
  if (n > cap) return fail(TK_ERR_CAP, "dump needs %llu entries, buffer holds %llu",
                           (unsigned long long)n, (unsigned long long)cap);
  if (!n) return TK_OK;
  const uint64_t limit = dump_chunk_limit();
  const uint64_t kmax = t->k >= 32 ? ~0ull : (1ull << (2 * t->k)) - 1;
  // Key ranges [klo, khi] of at most `limit` entries, ascending: a range over
  // the limit is binned by up to 12 more key bits and its bins grouped; a
  // bin still over the limit is binned again (distinct keys: a one-key bin
  // always fits).
  struct Range {
    uint64_t klo, khi, n;
  };
  std::vector<Range> ranges;
  std::vector<Range> todo{{0, kmax, n}};  // stack, top = next in key order
  void* pbins = nullptr;
  const uint64_t nbins_max = kHistoLds;
  if (n > limit && (rc = ensure_scratch(c, 1, nbins_max * 8, &pbins))) return rc;
  std::vector<uint64_t> bins(nbins_max);
  while (!todo.empty()) {
    const Range r = todo.back();
    todo.pop_back();
    if (r.n <= limit) {
      ranges.push_back(r);
      continue;
    }
    const uint64_t span = r.khi - r.klo;  // range holds span + 1 keys
    int bits = 64 - __builtin_clzll(span | 1);
    const uint32_t shift = (uint32_t)std::max(0, bits - 12);
    const uint64_t nb = (span >> shift) + 1;
    TK_HIP(hipMemsetAsync(pbins, 0, nb * 8, c->stream));
    TableView v = table_view(t);
    const uint64_t ns = t->nb * 8;
    TK_LAUNCH(c, "dump_bins", std::min<uint64_t>(ceil_div(ns, 256), 65536), 256, 0, k_dump_bins, v, ns,
              (uint64_t)lo, (uint64_t)hi, r.klo, r.khi, shift, (unsigned long long*)pbins);
    TK_HIP(hipMemcpyAsync(bins.data(), pbins, nb * 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    std::vector<Range> parts;  // this range's pieces in key order
    uint64_t first = 0, acc = 0, total = 0;
    auto bin_lo = [&](uint64_t b) { return r.klo + (b << shift); };
    auto bin_hi = [&](uint64_t b) { return b + 1 == nb ? r.khi : r.klo + ((b + 1) << shift) - 1; };
    for (uint64_t b = 0; b < nb; b++) {
      total += bins[b];
      if (bins[b] > limit) {
        if (acc) parts.push_back({bin_lo(first), bin_hi(b - 1), acc});
        parts.push_back({bin_lo(b), bin_hi(b), bins[b]});  // binned again
        first = b + 1;
        acc = 0;
        continue;
      }
      if (acc && acc + bins[b] > limit) {
        parts.push_back({bin_lo(first), bin_hi(b - 1), acc});
        first = b;
        acc = 0;
      }
      if (!acc) first = b;
      acc += bins[b];
    }
    if (acc) parts.push_back({bin_lo(first), bin_hi(nb - 1), acc});
    if (total != r.n)
      return fail(TK_ERR_STATE, "dump: binned %llu of %llu entries", (unsigned long long)total,
                  (unsigned long long)r.n);
    if (shift == 0)
      for (const Range& q : parts)
        if (q.n > limit) return fail(TK_ERR_STATE, "dump: %llu entries share one key", (unsigned long long)q.n);
    for (auto it = parts.rbegin(); it != parts.rend(); ++it) todo.push_back(*it);
  }
  uint64_t most = 0;
  for (auto& r : ranges) most = std::max(most, r.n);
  uint64_t* dk = nullptr;
  uint32_t* dc = nullptr;
  if ((rc = dev_alloc((void**)&dk, most * 8))) return rc;
  if ((rc = dev_alloc((void**)&dc, most * 4))) { dev_free(dk, most * 8); return rc; }
  uint64_t at = 0;
  for (size_t i = 0; i < ranges.size() && !rc; i++) {
    if (!ranges[i].n) continue;
    rc = dump_range(c, t, lo, hi, ranges[i].klo, ranges[i].khi, ranges[i].n, dk, dc, keys + at, counts + at);
    at += ranges[i].n;
  }
  if (!rc && at != n)
    rc = fail(TK_ERR_STATE, "dump: %llu of %llu entries in ranges", (unsigned long long)at, (unsigned long long)n);
  (void)hipStreamSynchronize(c->stream);
  dev_free(dk, most * 8);
  dev_free(dc, most * 4);
  return rc;
}

int tk_table_histo(tk_ctx* c, tk_table* t, uint64_t low, uint64_t high, uint64_t inc,
                   uint64_t* bins, uint64_t nbins) {
  if (!c || !t || !bins) return fail(TK_ERR_ARG, "NULL argument");
  if (inc == 0 || high < low) return fail(TK_ERR_ARG, "bad histo range");
  const uint64_t base = low > 1 ? (inc >= low ? 1 : low - inc) : 1;
  const uint64_t ceil_ = high + inc;
  const uint64_t nb = (ceil_ + inc - base) / inc;
  if (nb != nbins) return fail(TK_ERR_ARG, "nbins must be %llu", (unsigned long long)nb);
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = table_set_final(c, t, true);
  if (rc) return rc;
  void* p = nullptr;
  if ((rc = ensure_scratch(c, 1, nb * 8, &p))) return rc;
  TK_HIP(hipMemsetAsync(p, 0, nb * 8, c->stream));
  TableView v = table_view(t);
  uint64_t ns = t->nb * 8;
  TK_LAUNCH(c, "histo", std::min<uint64_t>(ceil_div(ns, 256), 4096), 256, 0, k_histo, v, ns, base,
            ceil_, inc, nb, (unsigned long long*)p);
  TK_HIP(hipMemcpyAsync(bins, p, nb * 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  return TK_OK;
}

int tk_table_prune(tk_ctx* c, tk_table* t, uint32_t min_count, const uint8_t* keep, uint64_t* n_kept) {
  if (!c || !t) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = table_set_final(c, t, true);
  if (rc) return rc;
  const uint64_t kt = (uint64_t)(t->k + 1) * (t->k + 1) * (t->k + 1);
  uint8_t* dkeep = nullptr;
  if (keep) {
    void* p = nullptr;
    if ((rc = ensure_scratch(c, 1, kt, &p))) return rc;
    dkeep = (uint8_t*)p;
    TK_HIP(hipMemcpyAsync(dkeep, keep, kt, hipMemcpyHostToDevice, c->stream));
  }
  unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_flags + 8);
  TK_HIP(hipMemsetAsync(d, 0, 8, c->stream));
  TableView v = table_view(t);
  uint64_t ns = t->nb * 8;
  TK_LAUNCH(c, "prune", std::min<uint64_t>(ceil_div(ns, 256), 65536), 256, 0, k_prune, v, ns, t->k,
            (uint64_t)min_count, dkeep, d);
  uint64_t h = 0;
  TK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  if (n_kept) *n_kept = h;
  return TK_OK;
}

int tk_table_lookup(tk_ctx* c, tk_table* t, const uint64_t* keys, uint64_t n, uint32_t* counts) {
  if (!c || !t || (n && (!keys || !counts))) return fail(TK_ERR_ARG, "NULL argument");
  if (!n) return TK_OK;
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = table_set_final(c, t, true);
  if (rc) return rc;
  uint64_t* dk = nullptr;
  uint32_t* dc = nullptr;
  if ((rc = dev_alloc((void**)&dk, n * 8))) return rc;
  if ((rc = dev_alloc((void**)&dc, n * 4))) { dev_free(dk, n * 8); return rc; }
  hipError_t e = hipMemcpyAsync(dk, keys, n * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipEvent_t a;
    timing_begin(c, &a);
    hipLaunchKernelGGL(k_lookup, dim3(std::min<uint64_t>(ceil_div(n, 256), 65536)), dim3(256), 0,
                       c->stream, table_view(t), dk, n, dc);
    timing_end(c, "lookup", a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(counts, dc, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_free(dk, n * 8);
  dev_free(dc, n * 4);
  if (e != hipSuccess) return hip_fail(e, "lookup");
  return TK_OK;
}

int tk_table_free(tk_table* t) {
  if (!t) return TK_OK;
  (void)hipSetDevice(t->device);
  if (t->d_slots) dev_free(t->d_slots, t->nb * 64);
  if (t->d_ovf) dev_free(t->d_ovf, t->ovf_cap * 16);
  delete t;
  return TK_OK;
}

// ---------------- stats ----------------
// LDS count-cache slots of the ordered scoring pass: TK_STATS_CACHE = 0
// (default), 2048 or 4096.  Off by default: on the 20M-pair workload the
// cached pass took 116 ms (2048) and 156 ms (4096) against 78 ms without,
// the LDS per block cutting the waves in flight more than the cache cuts
// table reads.
static int stats_cache_slots() {
  const char* e = getenv("TK_STATS_CACHE");
  if (!e) return 0;
  const int v = atoi(e);
  return v == 2048 ? 2048 : v == 0 ? 0 : 4096;
}

int tk_read_cov_stats_d(tk_ctx* c, tk_table* t, const tk_reads* r, const tk_stats_opts* o,
                        tk_read_stat* d_out) {
  if (!c || !t || !r || !o || (r->n_reads && !d_out)) return fail(TK_ERR_ARG, "NULL argument");
  if (o->nwin_policy != TK_NWIN_AS_MISSING && o->nwin_policy != TK_NWIN_SKIP)
    return fail(TK_ERR_ARG, "bad nwin_policy %d", o->nwin_policy);
  if (o->median_policy < 0 || o->median_policy > 2)
    return fail(TK_ERR_ARG, "bad median_policy %d", o->median_policy);
  if (set_device(c->device)) return TK_ERR_HIP;
  int rc = table_set_final(c, t, true);
  if (rc) return rc;
  if (!r->n_reads) return TK_OK;
  const uint32_t max_nwin = r->max_len >= (uint32_t)t->k ? r->max_len - t->k + 1 : 0;
  uint32_t slab = 0;
  int wpb = 0;
  if (slab_layout(max_nwin, &slab, &wpb))
    return fail(TK_ERR_ARG, "reads longer than %u bases are not supported", 16128 + t->k - 1);
  tk_table* use = t;
  StatsParams sp{o->min_count, o->missing_value, o->nwin_policy, o->median_policy};
  if ((rc = reads_order(c, r))) return rc;
  const int grid = (per_read_grid(c, r->n_reads, wpb) + 7) / 8 * 8;
  const int cache = stats_cache_slots();
  if (cache && r->d_order && max_nwin <= 128) {
    const uint64_t g = (ceil_div(r->n_reads, kCacheChunk) + 7) / 8 * 8;
    if (g > 0x7FFFFFF0ull) return fail(TK_ERR_ARG, "too many reads for one scoring pass");
    const size_t lds = (size_t)cache * 12 + 4 * 8 * kLookupScratch;
    if (cache == 2048)
      TK_LAUNCH(c, "stats", g, 256, lds, k_stats_cached<2048>, reads_view(r), t->k, t->canonical,
                table_view(use), sp, d_out);
    else
      TK_LAUNCH(c, "stats", g, 256, lds, k_stats_cached<4096>, reads_view(r), t->k, t->canonical,
                table_view(use), sp, d_out);
    return TK_OK;
  }
  const size_t lds = (size_t)slab * 4 * wpb + (size_t)wpb * 8 * kLookupScratch;
  const uint32_t grab = read_ticket_grab();
  uint32_t* tickets = grab ? c->d_tickets : nullptr;
  if (tickets) TK_LAUNCH(c, "tickets", 1, 256, 0, k_fill_u32, tickets, kTicketWords, 0u);
  if (max_nwin <= 128 && tickets && stats_pipe()) {
    const int pgrid = (per_read_grid(c, r->n_reads, 4) + 7) / 8 * 8;
    TK_LAUNCH(c, "stats", pgrid, 256, 4 * 8 * kLookupScratch, k_stats_pipe, reads_view(r), t->k, t->canonical,
              table_view(use), sp, d_out, tickets, grab);
  } else if (max_nwin <= 128 && tickets)
    TK_LAUNCH(c, "stats", grid, wpb * 64, lds, k_stats_short<true>, reads_view(r), t->k, t->canonical,
              table_view(use), sp, d_out, slab, tickets, grab);
  else if (max_nwin <= 128)
    TK_LAUNCH(c, "stats", grid, wpb * 64, lds, k_stats_short<false>, reads_view(r), t->k, t->canonical,
              table_view(use), sp, d_out, slab, tickets, grab);
  else if (tickets)
    TK_LAUNCH(c, "stats", grid, wpb * 64, lds, k_stats<true>, reads_view(r), t->k, t->canonical, table_view(use),
              sp, d_out, slab, tickets, grab);
  else
    TK_LAUNCH(c, "stats", grid, wpb * 64, lds, k_stats<false>, reads_view(r), t->k, t->canonical, table_view(use),
              sp, d_out, slab, tickets, grab);
  return TK_OK;
}

int tk_read_cov_stats(tk_ctx* c, tk_table* t, const tk_reads* r, const tk_stats_opts* o,
                      tk_read_stat* out) {
  if (!c || !r || (r->n_reads && !out)) return fail(TK_ERR_ARG, "NULL argument");
  if (!r->n_reads) return tk_read_cov_stats_d(c, t, r, o, nullptr);
  tk_read_stat* d = nullptr;
  size_t bytes = r->n_reads * sizeof(tk_read_stat);
  int rc = dev_alloc((void**)&d, bytes);
  if (rc) return rc;
  rc = tk_read_cov_stats_d(c, t, r, o, d);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "stats download");
  }
  (void)hipStreamSynchronize(c->stream);
  dev_free(d, bytes);
  return rc;
}

// ---------------- selection ----------------
}  // extern "C"
// The pair selection of nbkc_normalize.pl on device stats: flags of the pairs
// that draw a rand() value, their ranks (exclusive scan), then the draws.
// consumes/idx hold n words each; flags (16 words, cleared here by a kernel)
// get the error bits (word 0) and the selected count (words 8-9); scan_temp
// (scan_bytes) is the scan's temporary storage, or NULL for the context's.
namespace tk {
int select_run(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n, int max_cov,
               int min_cov, uint8_t* dsel, uint32_t* consumes, uint32_t* idx, void* scan_temp,
               size_t scan_bytes, uint32_t* flags, uint64_t* n_selected) {
  int rc = TK_OK;
  TK_LAUNCH(c, "select_flags", 1, 64, 0, k_fill_u32, flags, 16u, 0u);
  int grid = (int)std::min<uint64_t>(ceil_div(n, 256), 65536);
  TK_LAUNCH(c, "select_flags", grid, 256, 0, k_select_flags, dl, dr, n, min_cov, consumes, flags);
  if (scan_temp) rc = exclusive_scan_u32_in(c, consumes, idx, n, scan_temp, scan_bytes);
  else rc = exclusive_scan_u32(c, consumes, idx, n);
  if (rc) return rc;
  Lcg48Jump j;
  const uint64_t M = (1ull << 48) - 1;
  uint64_t A = 0x5DEECE66Dull, C = 0xB;
  for (int b = 0; b < 48; b++) {
    j.a[b] = A;
    j.c[b] = C;
    C = (A * C + C) & M;  // (A,C) o (A,C)
    A = (A * A) & M;
  }
  const uint64_t state0 = ((uint64_t)12345 << 16) | 0x330Eull;
  unsigned long long* d = reinterpret_cast<unsigned long long*>(flags + 8);
  const int sgrid = (int)std::min<uint64_t>(ceil_div(n, 256), (uint64_t)std::max(c->num_cu, 1) * 16);
  TK_LAUNCH(c, "select", sgrid, 256, 0, k_select, dl, dr, n, consumes, idx, j, state0, max_cov,
            dsel, d);
  TK_HIP(hipMemcpyAsync(c->h_flags, flags, 40, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  if (c->h_flags[0] & 1u)
    return fail(TK_ERR_ARG, "select: a read's mean coverage is below 1 (stats computed with missing_value 0?); "
                            "the merged-mean test is exact only for means >= 1");
  if (c->h_flags[0] & 2u)
    return fail(TK_ERR_ARG, "select: a pair with median 0 reaches the max_cov/median draw "
                            "(nbkc_normalize.pl divides by zero there)");
  uint64_t h = 0;
  memcpy(&h, &c->h_flags[8], 8);  // the count k_select adds at d
  if (n_selected) *n_selected = h;
  return TK_OK;
}

int select_check(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n, int max_cov,
                 int max_cv, uint32_t max_nwin, uint8_t* dsel) {
  if (!c || (n && (!dl || !dr || !dsel))) return fail(TK_ERR_ARG, "NULL argument");
  if (max_cov <= 0) return fail(TK_ERR_ARG, "max_cov must be positive");
  if ((double)max_cv < 1.2 * std::sqrt((double)max_nwin) + 1.0)
    return fail(TK_ERR_ARG, "max_cv=%d below the bound 1.2*sqrt(%u)+1 where the CV test is exact",
                max_cv, max_nwin);
  if (n > 0xFFFFFFFFull) return fail(TK_ERR_ARG, "too many pairs");
  return TK_OK;
}

// d_out[i] = d_stats[d_idx[i]] for i < n, on the context's stream (the
// resident normalization service's general jobs).
int gather_stats(tk_ctx* c, const tk_read_stat* d_stats, uint64_t n_reads, const uint64_t* d_idx, uint64_t n,
                 tk_read_stat* d_out) {
  if (!c || (n && (!d_stats || !d_idx || !d_out))) return fail(TK_ERR_ARG, "NULL argument");
  if (!n) return TK_OK;
  if (set_device(c->device)) return TK_ERR_HIP;
  const uint64_t grid = std::min<uint64_t>(ceil_div(n, 256), 65536);
  TK_LAUNCH(c, "gather_stats", grid, 256, 0, k_gather_stats, d_stats, n_reads, d_idx, n, d_out);
  return TK_OK;
}
}  // namespace tk

extern "C" {
int tk_select_pairs_d(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n,
                      int max_cov, int min_cov, int max_cv, uint32_t max_nwin, uint8_t* dsel,
                      uint64_t* n_selected) {
  int rc = select_check(c, dl, dr, n, max_cov, max_cv, max_nwin, dsel);
  if (rc) return rc;
  if (set_device(c->device)) return TK_ERR_HIP;
  if (n_selected) *n_selected = 0;
  if (!n) return TK_OK;
  void* p = nullptr;
  if ((rc = ensure_scratch(c, 1, n * 8 + 256, &p))) return rc;
  uint32_t* consumes = (uint32_t*)p;
  uint32_t* idx = consumes + ((n + 63) / 64) * 64;
  return select_run(c, dl, dr, n, max_cov, min_cov, dsel, consumes, idx, nullptr, 0, c->d_flags, n_selected);
}

int tk_select_pairs_d_noalloc(tk_ctx* c, const tk_read_stat* dl, const tk_read_stat* dr, uint64_t n,
                              int max_cov, int min_cov, int max_cv, uint32_t max_nwin, uint8_t* dsel,
                              uint64_t* n_selected) {
  int rc = select_check(c, dl, dr, n, max_cov, max_cv, max_nwin, dsel);
  if (rc) return rc;
  if (n > kNoAllocPairs)
    return fail(TK_ERR_ARG, "tk_select_pairs_d_noalloc: %llu pairs exceed the %llu its preallocated scratch holds",
                (unsigned long long)n, (unsigned long long)kNoAllocPairs);
  if (set_device(c->device)) return TK_ERR_HIP;
  if (n_selected) *n_selected = 0;
  if (!n) return TK_OK;
  uint8_t* base = (uint8_t*)scratch_ptr(c->device);
  if (!base) return fail(TK_ERR_STATE, "tk_select_pairs_d_noalloc: no scratch address for device %d", c->device);
  uint32_t* flags = (uint32_t*)base;
  uint32_t* consumes = (uint32_t*)(base + 256);
  uint32_t* idx = consumes + kNoAllocPairs;
  void* temp = (void*)(idx + kNoAllocPairs);
  return select_run(c, dl, dr, n, max_cov, min_cov, dsel, consumes, idx, temp, kNoAllocScanBytes, flags,
                    n_selected);
}

int tk_select_pairs(tk_ctx* c, const tk_read_stat* l, const tk_read_stat* r, uint64_t n, int max_cov,
                    int min_cov, int max_cv, uint8_t* sel, uint64_t* n_selected) {
  if (!c || (n && (!l || !r || !sel))) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  uint32_t max_nwin = 0;
  for (uint64_t i = 0; i < n; i++) max_nwin = std::max(max_nwin, std::max(l[i].nwin, r[i].nwin));
  size_t sb = n * sizeof(tk_read_stat);
  tk_read_stat *dl = nullptr, *dr = nullptr;
  uint8_t* ds = nullptr;
  int rc = dev_alloc((void**)&dl, sb);
  if (!rc) rc = dev_alloc((void**)&dr, sb);
  if (!rc) rc = dev_alloc((void**)&ds, n + 1);
  if (!rc && n) {
    hipError_t e = hipMemcpyAsync(dl, l, sb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, r, sb, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "select upload");
  }
  if (!rc) rc = tk_select_pairs_d(c, dl, dr, n, max_cov, min_cov, max_cv, max_nwin, ds, n_selected);
  if (!rc && n) {
    hipError_t e = hipMemcpyAsync(sel, ds, n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "select download");
  }
  (void)hipStreamSynchronize(c->stream);
  dev_free(dl, sb);
  dev_free(dr, sb);
  dev_free(ds, n + 1);
  return rc;
}

// ---------------- ReadsToTranscripts ----------------
int tk_ctable_build(tk_ctx* c, int k, int canonical, const tk_reads* b, const uint32_t* comp_ids,
                    tk_ctable** out) {
  if (!c || !b || !out || (b->n_reads && !comp_ids)) return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (k < 1 || k > 31) return fail(TK_ERR_ARG, "k=%d outside [1,31]", k);
  for (uint64_t i = 0; i < b->n_reads; i++)
    if (comp_ids[i] == 0xFFFFFFFFu) return fail(TK_ERR_ARG, "component id 0xFFFFFFFF is reserved");
  if (set_device(c->device)) return TK_ERR_HIP;
  uint64_t total = 0;
  int rc = reads_windows(c, b, k, &total);
  if (rc) return rc;
  tk_ctable* t = new tk_ctable();
  t->device = c->device;
  t->k = k;
  t->canonical = canonical ? 1 : 0;
  t->nbuckets = next_pow2(std::max<uint64_t>(ceil_div(2 * total, 4), 16));
  uint32_t* dcomp = nullptr;
  size_t cb = std::max<uint64_t>(b->n_reads, 1) * 4;
  do {
    if ((rc = dev_alloc((void**)&t->d_slots, t->nbuckets * 64))) break;
    if ((rc = dev_alloc((void**)&dcomp, cb))) break;
    hipError_t e = hipMemsetAsync(t->d_slots, 0, t->nbuckets * 64, c->stream);
    if (e == hipSuccess && b->n_reads)
      e = hipMemcpyAsync(dcomp, comp_ids, b->n_reads * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_flags, 0, 16, c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "ctable init"); break; }
    if (b->n_reads) {
      hipEvent_t a;
      timing_begin(c, &a);
      hipLaunchKernelGGL(k_ctable_build, dim3(ceil_div(b->n_reads, kReadsPerBlock)), dim3(256), 0,
                         c->stream, reads_view(b), dcomp, k, t->canonical, t->d_slots,
                         t->nbuckets - 1, c->d_flags);
      timing_end(c, "ctable_build", a);
      if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "k_ctable_build"); break; }
    }
    e = hipMemcpyAsync(c->h_flags, c->d_flags, 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "ctable sync"); break; }
    if (c->h_flags[0]) { rc = fail(TK_ERR_FULL, "ctable probe limit reached"); break; }
    memcpy(&t->n_keys, &c->h_flags[2], 8);
  } while (0);
  dev_free(dcomp, cb);
  t->bytes = t->nbuckets * 64;
  if (rc) { tk_ctable_free(t); return rc; }
  *out = t;
  return TK_OK;
}

int tk_ctable_info(const tk_ctable* t, uint64_t* n_keys, uint64_t* capacity) {
  if (!t) return fail(TK_ERR_ARG, "ctable is NULL");
  if (n_keys) *n_keys = t->n_keys;
  if (capacity) *capacity = t->nbuckets * 4;
  return TK_OK;
}

int tk_r2t_vote_reads_d(tk_ctx* c, const tk_ctable* t, const tk_reads* r, tk_r2t_vote* d_out) {
  if (!c || !t || !r || (r->n_reads && !d_out)) return fail(TK_ERR_ARG, "NULL argument");
  if (set_device(c->device)) return TK_ERR_HIP;
  if (!r->n_reads) return TK_OK;
  const uint32_t max_nwin = r->max_len >= (uint32_t)t->k ? r->max_len - t->k + 1 : 0;
  uint32_t slab = 0;
  int wpb = 0;
  if (slab_layout(max_nwin, &slab, &wpb))
    return fail(TK_ERR_ARG, "reads longer than %u bases are not supported", 16128 + t->k - 1);
  int rc = reads_order(c, r);
  if (rc) return rc;
  const int grid = (per_read_grid(c, r->n_reads, wpb) + 7) / 8 * 8;
  const uint32_t grab = read_ticket_grab();
  uint32_t* tickets = grab ? c->d_tickets : nullptr;
  if (tickets) TK_LAUNCH(c, "tickets", 1, 256, 0, k_fill_u32, tickets, kTicketWords, 0u);
  const size_t lds = (size_t)slab * 4 * wpb + (size_t)wpb * 256;
  if (tickets && max_nwin <= 128 && vote_pipe()) {
    const int pgrid = (per_read_grid(c, r->n_reads, 4) + 7) / 8 * 8;
    TK_LAUNCH(c, "vote", pgrid, 256, 0, k_vote_pipe, reads_view(r), t->k, t->canonical, t->d_slots,
              t->nbuckets - 1, d_out, tickets, grab);
  } else if (tickets)
    TK_LAUNCH(c, "vote", grid, wpb * 64, lds, k_vote<true>, reads_view(r), t->k, t->canonical, t->d_slots,
              t->nbuckets - 1, d_out, slab, tickets, grab);
  else
    TK_LAUNCH(c, "vote", grid, wpb * 64, lds, k_vote<false>, reads_view(r), t->k, t->canonical, t->d_slots,
              t->nbuckets - 1, d_out, slab, tickets, grab);
  return TK_OK;
}

int tk_r2t_vote_reads(tk_ctx* c, const tk_ctable* t, const tk_reads* r, tk_r2t_vote* out) {
  if (!c || !r || (r->n_reads && !out)) return fail(TK_ERR_ARG, "NULL argument");
  if (!r->n_reads) return TK_OK;
  tk_r2t_vote* d = nullptr;
  size_t bytes = r->n_reads * sizeof(tk_r2t_vote);
  int rc = dev_alloc((void**)&d, bytes);
  if (rc) return rc;
  rc = tk_r2t_vote_reads_d(c, t, r, d);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "vote download");
  }
  (void)hipStreamSynchronize(c->stream);
  dev_free(d, bytes);
  return rc;
}

int tk_ctable_free(tk_ctable* t) {
  if (!t) return TK_OK;
  (void)hipSetDevice(t->device);
  if (t->d_slots) dev_free(t->d_slots, t->nbuckets * 64);
  delete t;
  return TK_OK;
}

}  // extern "C"
