// tk_kset.hip -- occurrence counts of a fixed set of long or tagged k-mers.
//
// Two Chrysalis steps count, over the reads, k-mers that come from the
// Inchworm contigs rather than from the reads themselves:
//  * GraphFromFasta's weld support (Trinity:2180, -k 24 -kk 48): the 48-mers
//    that span a 24-mer two contigs share, counted in the reads, decide which
//    contigs are glued into one component;
//  * QuantifyGraph (Trinity:2352, -k 24): each component's de Bruijn nodes
//    (its 24-mers) and edges (25-mers) weighted by the windows of the reads
//    ReadsToTranscripts assigned to that component.
// Both are one table of 128-bit keys (hi, lo) built once and probed by every
// read window: a plain kk-mer (kk <= 63) is its 2kk-bit value split in two
// words; a tagged k-mer (k <= 32) is (tag of the read, k-mer), so one table
// holds the nodes of every component of a batch.
//
// Table: 64-B buckets of four 16-B slots {lo, hi | occupied}, load <= 0.5,
// power-of-two bucket count, linear bucket probing.  Keys must be distinct
// (checked after the build); each slot carries the key's input index, and
// counts are u64 per slot, gathered back to input order.
#include <string.h>

#include <algorithm>
#include <vector>

#include "tk_internal.h"

struct tk_kset {
  int device = 0;
  uint64_t n = 0;          // keys
  uint64_t nb = 0;         // buckets (4 slots each)
  uint64_t* d_slots = nullptr;   // 2 words per slot
  uint32_t* d_idx = nullptr;     // input index per slot
  unsigned long long* d_counts = nullptr;  // per slot
  uint64_t bytes = 0;
};

namespace tk {
namespace {

constexpr uint64_t kOcc = 1ull << 63;

#define KS_LAUNCH(ctx, name, grid, block, kernel, ...)                                         \
  do {                                                                                         \
    hipEvent_t _a;                                                                             \
    ::tk::timing_begin((ctx), &_a);                                                            \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (ctx)->stream, __VA_ARGS__);        \
    ::tk::timing_end((ctx), (name), _a);                                                       \
    hipError_t _le = hipGetLastError();                                                        \
    if (_le != hipSuccess) return ::tk::hip_fail(_le, "launch " name);                         \
  } while (0)

inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint64_t ks_hash(uint64_t hi, uint64_t lo) {
  return mix64(lo ^ mix64(hi + 0x9E3779B97F4A7C15ULL));
}

// Claim the first empty slot on the key's probe path (keys distinct: no
// compare needed).  flags[0] |= 1 when the table has no room.
__global__ void k_ks_build(const uint64_t* __restrict__ keys, uint64_t n, uint64_t* __restrict__ slots,
                           uint32_t* __restrict__ idx, uint64_t nb, uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = keys[2 * i], hi = keys[2 * i + 1];
    uint64_t b = ks_hash(hi, lo) & (nb - 1);
    bool done = false;
    for (uint64_t p = 0; p < nb && !done; p++) {
      for (int j = 0; j < 4 && !done; j++) {
        unsigned long long* w1 = (unsigned long long*)&slots[(b * 4 + j) * 2 + 1];
        if (*w1 != 0) continue;
        if (atomicCAS(w1, 0ull, (unsigned long long)(hi | kOcc)) == 0ull) {
          slots[(b * 4 + j) * 2] = lo;
          idx[b * 4 + j] = (uint32_t)i;
          done = true;
        }
      }
      b = (b + 1) & (nb - 1);
    }
    if (!done) atomicOr(&flags[0], 1u);
  }
}

// Slot holding (hi, lo), or ~0 when absent.
__device__ __forceinline__ uint64_t ks_find(const uint64_t* __restrict__ slots, uint64_t nb, uint64_t hi,
                                            uint64_t lo) {
  const uint64_t w1want = hi | kOcc;
  uint64_t b = ks_hash(hi, lo) & (nb - 1);
  for (uint64_t p = 0; p < nb; p++) {
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(slots + b * 8);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const ulonglong2 s = l2[j];
      if (s.y == 0) return ~0ull;
      if (s.y == w1want && s.x == lo) return b * 4 + j;
    }
    b = (b + 1) & (nb - 1);
  }
  return ~0ull;
}

// Every key found exactly once (a duplicate input key would sit in two slots).
__global__ void k_ks_verify(const uint64_t* __restrict__ keys, uint64_t n, const uint64_t* __restrict__ slots,
                            uint64_t nb, uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = keys[2 * i], hi = keys[2 * i + 1];
    const uint64_t w1want = hi | kOcc;
    uint64_t b = ks_hash(hi, lo) & (nb - 1);
    uint32_t found = 0;
    bool end = false;
    for (uint64_t p = 0; p < nb && !end; p++) {
      for (int j = 0; j < 4; j++) {
        const uint64_t w1 = slots[(b * 4 + j) * 2 + 1];
        if (w1 == 0) {
          end = true;
          break;
        }
        if (w1 == w1want && slots[(b * 4 + j) * 2] == lo) found++;
      }
      b = (b + 1) & (nb - 1);
    }
    if (found != 1) atomicOr(&flags[0], 2u);
  }
}

// One thread per read: roll the read's bases into the forward and
// reverse-complement 2kk-bit values (run of ACGT bases >= kk: valid window),
// and add one to the slot of each valid window's key.  With tags, the key is
// (tags[read], k-mer) and kk <= 32.
__global__ __launch_bounds__(256) void k_ks_count(ReadsView rv, int kk, int canonical,
                                                  const uint32_t* __restrict__ tags,
                                                  const uint64_t* __restrict__ slots, uint64_t nb,
                                                  unsigned long long* __restrict__ counts) {
  typedef unsigned __int128 u128;
  const u128 mask = ((u128)1 << (2 * kk)) - 1;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rv.n_reads;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t len = rv.len[r];
    const uint64_t wbase = rv.word_off[r];
    const uint64_t tag = tags ? (uint64_t)tags[r] : 0;
    u128 f = 0, rc = 0;
    uint32_t run = 0;
    uint64_t code_word = 0;
    uint32_t mask_word = 0;
    for (uint32_t i = 0; i < len; i++) {
      if ((i & 31) == 0) {
        code_word = rv.codes[wbase + (i >> 5)];
        mask_word = rv.nmask[wbase + (i >> 5)];
      }
      const uint32_t o = i & 31;
      const uint32_t c = (uint32_t)(code_word >> (62 - 2 * o)) & 3u;
      const bool bad = (mask_word >> (31 - o)) & 1u;
      run = bad ? 0 : run + 1;
      f = ((f << 2) | (u128)c) & mask;
      rc = (rc >> 2) | ((u128)(3u - c) << (2 * kk - 2));
      if (run < (uint32_t)kk) continue;
      const u128 key = (canonical && rc < f) ? rc : f;
      uint64_t hi, lo;
      if (tags) {
        hi = tag;
        lo = (uint64_t)key;
      } else {
        hi = (uint64_t)(key >> 64);
        lo = (uint64_t)key;
      }
      const uint64_t s = ks_find(slots, nb, hi, lo);
      if (s != ~0ull) atomicAdd(&counts[s], 1ull);
    }
  }
}

__global__ void k_ks_gather(const uint64_t* __restrict__ slots, const uint32_t* __restrict__ idx,
                            const unsigned long long* __restrict__ counts, uint64_t n_slots,
                            unsigned long long* __restrict__ out) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_slots;
       s += (uint64_t)gridDim.x * blockDim.x)
    if (slots[s * 2 + 1]) out[idx[s]] = counts[s];
}

__global__ void k_ks_zero(unsigned long long* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

}  // namespace
}  // namespace tk

extern "C" {

int tk_kset_create(tk_ctx* c, const uint64_t* keys, uint64_t n, tk_kset** out) {
  using namespace tk;
  if (!c || !out || (n && !keys)) return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (n >= 0xFFFFFFFFull) return fail(TK_ERR_ARG, "tk_kset_create: %llu keys (at most 2^32-2)", (unsigned long long)n);
  for (uint64_t i = 0; i < n; i++)
    if (keys[2 * i + 1] & (3ull << 62)) return fail(TK_ERR_ARG, "tk_kset_create: key %llu has its high word >= 2^62",
                                                     (unsigned long long)i);
  if (hipSetDevice(c->device) != hipSuccess) return fail(TK_ERR_HIP, "hipSetDevice");
  tk_kset* s = new tk_kset();
  s->device = c->device;
  s->n = n;
  uint64_t nb = 1;
  while (nb * 4 < 2 * n + 8) nb <<= 1;  // load <= 0.5
  s->nb = nb;
  const uint64_t ns = nb * 4;
  int rc = dev_alloc((void**)&s->d_slots, ns * 16);
  if (!rc) rc = dev_alloc((void**)&s->d_idx, ns * 4);
  if (!rc) rc = dev_alloc((void**)&s->d_counts, ns * 8);
  uint64_t* dk = nullptr;
  if (!rc && n) rc = dev_alloc((void**)&dk, n * 16);
  auto cleanup = [&]() {
    if (dk) dev_free(dk, n * 16);
  };
  s->bytes = ns * 28;
  if (rc) {
    cleanup();
    tk_kset_free(s);
    return rc;
  }
  hipError_t e = hipMemsetAsync(s->d_slots, 0, ns * 16, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->d_flags, 0, 4, c->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(dk, keys, n * 16, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) {
    cleanup();
    tk_kset_free(s);
    return hip_fail(e, "kset upload");
  }
  auto launch = [&]() -> int {
    if (!n) return TK_OK;
    const uint64_t grid = std::min<uint64_t>(cdiv(n, 256), 65536);
    KS_LAUNCH(c, "kset_build", grid, 256, k_ks_build, dk, n, s->d_slots, s->d_idx, nb, c->d_flags);
    KS_LAUNCH(c, "kset_verify", grid, 256, k_ks_verify, dk, n, s->d_slots, nb, c->d_flags);
    return TK_OK;
  };
  rc = launch();
  if (!rc) {
    e = hipMemcpyAsync(c->h_flags, c->d_flags, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "kset build");
    else if (c->h_flags[0] & 1u) rc = fail(TK_ERR_STATE, "tk_kset_create: table full");
    else if (c->h_flags[0] & 2u) rc = fail(TK_ERR_ARG, "tk_kset_create: keys are not distinct");
  }
  cleanup();
  if (rc) {
    tk_kset_free(s);
    return rc;
  }
  *out = s;
  return TK_OK;
}

int tk_kset_count(tk_ctx* c, tk_kset* s, const tk_reads* reads, int kk, int canonical, const uint32_t* tags,
                  uint64_t* counts) {
  using namespace tk;
  if (!c || !s || !reads || (s->n && !counts)) return fail(TK_ERR_ARG, "NULL argument");
  if (kk < 1 || kk > 63 || (tags && kk > 32))
    return fail(TK_ERR_ARG, "tk_kset_count: kk=%d (1..63, 1..32 with tags)", kk);
  if (hipSetDevice(c->device) != hipSuccess) return fail(TK_ERR_HIP, "hipSetDevice");
  const uint64_t ns = s->nb * 4;
  uint32_t* dtags = nullptr;
  unsigned long long* dout = nullptr;
  int rc = TK_OK;
  if (tags && reads->n_reads) rc = dev_alloc((void**)&dtags, reads->n_reads * 4);
  if (!rc) rc = dev_alloc((void**)&dout, std::max<uint64_t>(s->n, 1) * 8);
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    if (dtags) dev_free(dtags, reads->n_reads * 4);
    if (dout) dev_free(dout, std::max<uint64_t>(s->n, 1) * 8);
  };
  if (rc) {
    cleanup();
    return rc;
  }
  auto body = [&]() -> int {
    if (dtags) TK_HIP(hipMemcpyAsync(dtags, tags, reads->n_reads * 4, hipMemcpyHostToDevice, c->stream));
    KS_LAUNCH(c, "kset_zero", std::min<uint64_t>(cdiv(ns, 256), 65536), 256, k_ks_zero, s->d_counts, ns);
    ReadsView rv;
    rv.codes = reads->d_codes;
    rv.nmask = reads->d_nmask;
    rv.word_off = reads->d_word_off;
    rv.len = reads->d_len;
    rv.n_reads = reads->n_reads;
    if (reads->n_reads)
      KS_LAUNCH(c, "kset_count", std::min<uint64_t>(cdiv(reads->n_reads, 256), (uint64_t)std::max(c->num_cu, 1) * 32),
                256, k_ks_count, rv, kk, canonical, dtags, s->d_slots, s->nb, s->d_counts);
    if (s->n) {
      KS_LAUNCH(c, "kset_gather", std::min<uint64_t>(cdiv(ns, 256), 65536), 256, k_ks_gather, s->d_slots, s->d_idx,
                s->d_counts, ns, dout);
      TK_HIP(hipMemcpyAsync(counts, dout, s->n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    TK_HIP(hipStreamSynchronize(c->stream));
    return TK_OK;
  };
  rc = body();
  cleanup();
  return rc;
}

int tk_kset_free(tk_kset* s) {
  if (!s) return TK_OK;
  const uint64_t ns = s->nb * 4;
  (void)hipSetDevice(s->device);
  tk::dev_free(s->d_slots, ns * 16);
  tk::dev_free(s->d_idx, ns * 4);
  tk::dev_free(s->d_counts, ns * 8);
  delete s;
  return TK_OK;
}

}  // extern "C"
