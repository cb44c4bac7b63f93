// tk_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Read layout in HBM (tk_reads):
//   codes[w]  : 64-bit word, 32 bases, 2 bits each, first base in bits 63..62
//   nmask[w]  : 32-bit word, bit 31-i set when base i of codes[w] is not ACGT
//   word_off[r]: first word of read r (each read starts on a fresh word)
//   len[r]    : bases in read r
// One zero padding word follows the last read, so a window extractor may
// always read words w and w+1.
//
// Count table (tk_table): open addressing over 64-byte buckets of eight
// 64-bit slots.  A slot packs
//     [ count field : 63-2k bits | occupied : 1 bit | key : 2k bits ]
// so one 64-bit atomicAdd of (1 << (2k+1)) increments a k-mer and any carry
// leaves through bit 63 without touching the key.  A carry (field wrap) is
// recorded as 2^(63-2k) in a small side table keyed by k-mer ("ovf").
// Finalisation folds the side table back: the main field of every key with
// a side entry becomes all-ones and the side entry holds the full count.
// A field of 0 on an occupied, finalised slot is a pruned key (tombstone).
//
// Buckets are grouped into regions of 2^rbits consecutive buckets; the table
// holds nb = P * 2^rbits buckets.  A key's home bucket is
// mulhi(mix64(key), nb) and probing wraps inside the home region, so every
// key stays in one region.  Counting can then be done region by region with
// the active region resident in the Infinity Cache.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tk {

constexpr int kWave = 64;
constexpr int kBucketSlots = 8;
// u64 entries of LDS lookup scratch per wavefront (coop4_get with two keys per lane)
constexpr int kLookupScratch = 128;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// Reverse complement of a k-mer held in the low 2k bits.
__device__ __forceinline__ uint64_t revcomp(uint64_t v, int k) {
  uint64_t x = ~v;
  x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
  x = __builtin_bswap64(x);
  return x >> (64 - 2 * k);
}

struct ReadsView {
  const uint64_t* codes;
  const uint32_t* nmask;
  const uint64_t* word_off;
  const uint32_t* len;
  uint64_t n_reads;
  const uint32_t* order = nullptr;  // processing order of the table passes (nullptr = input order)
};

// Read processed at position i of the table passes.
__device__ __forceinline__ uint64_t read_at(const ReadsView& rv, uint64_t i) {
  return rv.order ? rv.order[i] : i;
}

// Workgroup b of a launch runs on XCD b % 8 (round-robin dispatch).  For a
// grid that is a multiple of 8 this maps b to a logical block such that each
// XCD takes one contiguous run of logical blocks, so work adjacent in the
// read order shares that XCD's L2.
__device__ __forceinline__ uint64_t xcd_block(uint64_t b, uint64_t grid) {
  return (b & 7) * (grid >> 3) + (b >> 3);
}

// Per-XCD position counters of the per-read table passes: eight u32, one per
// 256 B, cleared before each launch.
constexpr uint32_t kTicketStride = 64;
constexpr uint32_t kTicketWords = 8 * kTicketStride;

// Positions [0, n) of one XCD's share of a per-read pass, for one wave.
// kTicket: the XCD's waves take runs of `grab` consecutive positions from
// that XCD's counter *ctr (one returning atomic per run; the next run is
// taken while the current one is processed), so the positions in flight on
// the XCD stay one narrow front of the order however far single waves drift
// apart.  Otherwise the wave takes first, first + stride, ... (a fixed
// stride, along which waves drift apart by thousands of reads on large sets).
// Calls f(position) for each; every wave leaves once its next position is at
// or past n.
template <bool kTicket, class F>
__device__ __forceinline__ void xcd_positions(uint32_t* ctr, uint64_t n, uint32_t grab, uint64_t first,
                                              uint64_t stride, int lane, F&& f) {
  if (!kTicket) {
    for (uint64_t j = first; j < n; j += stride) f(j);
    return;
  }
  // positions fit 32 bits: a share is at most 2^31 / 8 reads
  const uint32_t n32 = (uint32_t)n;
  uint32_t t = 0;
  if (lane == 0) t = atomicAdd(ctr, grab);
  uint32_t b = __shfl(t, 0);
  while (b < n32) {
    t = 0;
    if (lane == 0) t = atomicAdd(ctr, grab);
    const uint32_t e = n32 - b > grab ? b + grab : n32;
    for (uint32_t j = b; j < e; j++) f(j);
    b = __shfl(t, 0);
  }
}

// The positions of xcd_positions<true>, one at a time: next() gives the
// wave's next position of [0, n), and n once the XCD's share is used up.
struct TicketStream {
  uint32_t* ctr;
  uint32_t n, grab;
  uint32_t cb = 0, ce = 0;  // current run [cb, ce)
  uint32_t tk = 0;          // lane 0: start of the run taken next
  bool done = false;
  int lane;
  __device__ __forceinline__ TicketStream(uint32_t* c, uint32_t n_, uint32_t g, int l)
      : ctr(c), n(n_), grab(g), lane(l) {
    if (lane == 0) tk = atomicAdd(ctr, grab);
  }
  __device__ __forceinline__ uint32_t next() {
    if (cb == ce) {
      if (done) return n;
      const uint32_t s = __shfl(tk, 0);
      if (s >= n) {
        done = true;
        return n;
      }
      cb = s;
      ce = n - s > grab ? s + grab : n;
      if (lane == 0) tk = atomicAdd(ctr, grab);
    }
    return cb++;
  }
};

// Forward k-mer starting at base `pos` of the read whose first word is wbase.
__device__ __forceinline__ void window_at(const ReadsView& rv, uint64_t wbase,
                                          uint32_t pos, int k, uint64_t& fwd,
                                          bool& valid) {
  const uint64_t w = wbase + (pos >> 5);
  const uint32_t o = pos & 31u;
  const uint64_t hi = rv.codes[w];
  const uint64_t lo = rv.codes[w + 1];
  const uint32_t mhi = rv.nmask[w];
  const uint32_t mlo = rv.nmask[w + 1];
  const uint64_t x = o ? ((hi << (2 * o)) | (lo >> (64 - 2 * o))) : hi;
  const uint32_t m = o ? ((mhi << o) | (mlo >> (32 - o))) : mhi;
  fwd = x >> (64 - 2 * k);
  valid = (m >> (32 - k)) == 0u;
}

// As window_at, plus the window's non-ACGT mask: bit k-1-i set when base
// pos+i is not ACGT (bit 0 = the window's last base).
__device__ __forceinline__ void window_at_mask(const ReadsView& rv, uint64_t wbase, uint32_t pos, int k,
                                               uint64_t& fwd, uint32_t& nmask) {
  const uint64_t w = wbase + (pos >> 5);
  const uint32_t o = pos & 31u;
  const uint64_t hi = rv.codes[w];
  const uint64_t lo = rv.codes[w + 1];
  const uint32_t mhi = rv.nmask[w];
  const uint32_t mlo = rv.nmask[w + 1];
  const uint64_t x = o ? ((hi << (2 * o)) | (lo >> (64 - 2 * o))) : hi;
  const uint32_t m = o ? ((mhi << o) | (mlo >> (32 - o))) : mhi;
  fwd = x >> (64 - 2 * k);
  nmask = m >> (32 - k);
}

// window_at over a read whose words are spread over the wave: lane j holds
// code word j of the read (cw) and its mask word (mw), for j up to the word
// after the read's last.  pos < 128, so words 0..4 are read.
__device__ __forceinline__ void window_lanes(uint64_t cw, uint32_t mw, uint32_t pos, int k, uint64_t& fwd,
                                             bool& valid) {
  const int w = (int)(pos >> 5);
  const uint32_t o = pos & 31u;
  const uint64_t hi = __shfl(cw, w);
  const uint64_t lo = __shfl(cw, w + 1);
  const uint32_t mhi = __shfl(mw, w);
  const uint32_t mlo = __shfl(mw, w + 1);
  const uint64_t x = o ? ((hi << (2 * o)) | (lo >> (64 - 2 * o))) : hi;
  const uint32_t m = o ? ((mhi << o) | (mlo >> (32 - o))) : mhi;
  fwd = x >> (64 - 2 * k);
  valid = (m >> (32 - k)) == 0u;
}

__device__ __forceinline__ uint64_t canon_key(uint64_t fwd, int k, int canonical) {
  if (!canonical) return fwd;
  const uint64_t rc = revcomp(fwd, k);
  return rc < fwd ? rc : fwd;
}

struct TableView {
  uint64_t* slots;    // base of the slot array (shared by the tables of a batch)
  uint64_t boff;      // first bucket of this table
  uint64_t nb;        // buckets
  uint64_t rmask;     // buckets per region - 1
  uint32_t rbits;     // log2(buckets per region)
  uint64_t occ;       // occupied bit (1 << 2k)
  uint64_t kmo;       // key + occupied bits
  uint64_t cmax;      // all-ones count field value
  uint32_t cshift;    // 2k + 1
  uint32_t cbits;     // 63 - 2k
  uint64_t* ovf;      // side table: pairs (tagged key, value)
  uint64_t omask;     // side capacity - 1
  uint32_t max_probe; // buckets probed before giving up
};

__device__ __forceinline__ uint64_t bucket0(const TableView& t, uint64_t key) {
  return t.boff + __umul64hi(mix64(key), t.nb);
}
__device__ __forceinline__ uint64_t bucket_next(const TableView& t, uint64_t b) {
  const uint64_t r = b - t.boff;
  return t.boff + ((r & ~t.rmask) | ((r + 1) & t.rmask));
}

// ---- side table (field carries) ----
__device__ __forceinline__ uint64_t ovf_slot0(const TableView& t, uint64_t tagged) {
  return mix64(tagged ^ 0x9E3779B97F4A7C15ULL) & t.omask;
}

__device__ inline void ovf_add(const TableView& t, uint64_t tagged, uint64_t units,
                               uint32_t* err) {
  uint64_t i = ovf_slot0(t, tagged);
  for (uint64_t p = 0; p <= t.omask; p++) {
    unsigned long long* kp = (unsigned long long*)&t.ovf[2 * i];
    uint64_t cur = *kp;
    if (cur == 0) {
      cur = atomicCAS(kp, 0ULL, (unsigned long long)tagged);
      if (cur == 0) cur = tagged;
    }
    if (cur == tagged) {
      atomicAdd((unsigned long long*)&t.ovf[2 * i + 1], (unsigned long long)units);
      return;
    }
    i = (i + 1) & t.omask;
  }
  atomicOr(err, 2u);
}

__device__ inline uint64_t* ovf_find(const TableView& t, uint64_t tagged) {
  uint64_t i = ovf_slot0(t, tagged);
  for (uint64_t p = 0; p <= t.omask; p++) {
    uint64_t cur = t.ovf[2 * i];
    if (cur == tagged) return &t.ovf[2 * i + 1];
    if (cur == 0) return nullptr;
    i = (i + 1) & t.omask;
  }
  return nullptr;
}

// Add r (< 2^cbits) to the count field of slot `s` and q (a multiple of
// 2^cbits) plus any carry to the side table.
__device__ __forceinline__ void field_add(const TableView& t, uint64_t* s, uint64_t tagged,
                                          uint64_t r, uint64_t q, uint32_t* err) {
  if (r) {
    const uint64_t old = atomicAdd((unsigned long long*)s, (unsigned long long)(r << t.cshift));
    if ((old >> t.cshift) + r > t.cmax) q += t.cmax + 1;
  }
  if (q) ovf_add(t, tagged, q, err);
}

// Insert-or-add c >= 1 occurrences of key.  Returns 1 when a new slot was
// claimed, 0 when the key was present, 2 when max_probe buckets were full
// (nothing added; the caller decides).
__device__ inline uint32_t table_add(const TableView& t, uint64_t key, uint64_t c,
                                     uint32_t* err) {
  const uint64_t tagged = key | t.occ;
  const uint64_t r = c & t.cmax;
  const uint64_t q = c - r;
  uint64_t b = bucket0(t, key);
  for (uint32_t p = 0; p < t.max_probe; p++) {
    uint64_t* line = t.slots + (b << 3);
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(line);
    ulonglong2 v0 = l2[0], v1 = l2[1], v2 = l2[2], v3 = l2[3];
    uint64_t sv[8] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      uint64_t s = sv[j];
      if (s == 0) {
        s = atomicCAS((unsigned long long*)&line[j], 0ULL,
                      (unsigned long long)(tagged | (r << t.cshift)));
        if (s == 0) {
          if (q) ovf_add(t, tagged, q, err);
          return 1u;
        }
      }
      if ((s & t.kmo) == tagged) {
        field_add(t, &line[j], tagged, r, q, err);
        return 0u;
      }
    }
    b = bucket_next(t, b);
  }
  return 2u;
}

// Count of key in a finalised table; 0 when absent or pruned.
__device__ __forceinline__ uint64_t table_get(const TableView& t, uint64_t key) {
  const uint64_t tagged = key | t.occ;
  uint64_t b = bucket0(t, key);
  for (uint32_t p = 0; p < t.max_probe; p++) {
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(t.slots + (b << 3));
    ulonglong2 v0 = l2[0], v1 = l2[1], v2 = l2[2], v3 = l2[3];
    uint64_t sv[8] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint64_t s = sv[j];
      if ((s & t.kmo) == tagged) {
        const uint64_t f = s >> t.cshift;
        if (f != t.cmax) return f;
        const uint64_t* o = ovf_find(t, tagged);
        return o ? *o : f;
      }
      if (s == 0) return 0;
    }
    b = bucket_next(t, b);
  }
  return 0;
}

// Count of the key held by an occupied slot value (finalised table).
__device__ __forceinline__ uint64_t slot_count(const TableView& t, uint64_t s) {
  const uint64_t f = s >> t.cshift;
  if (f != t.cmax) return f;
  const uint64_t* o = ovf_find(t, s & t.kmo);
  return o ? *o : f;
}

// ---- LDS tables of 64-B buckets (eight slots in the global slot format) ----
// Insert-or-add into an LDS table of nb buckets: `word` = tagged key | c <<
// cshift.  From bucket b (wrapping at nb) each step reads one bucket, picks without
// branching the slot holding the key or else the first empty one, and
// issues a single atomic there: add c, or claim with CAS (a lost claim
// rescans the same bucket).  Occupied slots stay a prefix of their bucket.
// Returns 2 when a slot was claimed, 1 when added (old = previous slot
// value), 0 when max_steps steps found no room.
__device__ __forceinline__ int lds_bucket_add(uint64_t* tab, uint32_t b, uint32_t nb, uint64_t word,
                                              uint64_t kmo, uint32_t cshift, uint32_t max_steps,
                                              uint64_t& old) {
  const uint64_t tagged = word & kmo;
  const uint64_t add = (word >> cshift) << cshift;
  for (uint32_t p = 0; p < max_steps; p++) {
    uint64_t* line = tab + ((uint64_t)b << 3);
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(line);
    const ulonglong2 v0 = l2[0], v1 = l2[1], v2 = l2[2], v3 = l2[3];
    const uint64_t sv[8] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
    int jm = 8, je = 8;
#pragma unroll
    for (int j = 7; j >= 0; j--) {
      jm = (sv[j] & kmo) == tagged ? j : jm;
      je = sv[j] == 0 ? j : je;
    }
    if (jm < 8) {
      old = atomicAdd((unsigned long long*)&line[jm], (unsigned long long)add);
      return 1;
    }
    if (je < 8) {
      const uint64_t v = atomicCAS((unsigned long long*)&line[je], 0ull, (unsigned long long)word);
      if (v == 0) return 2;
      if ((v & kmo) == tagged) {
        old = atomicAdd((unsigned long long*)&line[je], (unsigned long long)add);
        return 1;
      }
      continue;  // another key took the slot: rescan this bucket
    }
    b = b + 1 == nb ? 0 : b + 1;
  }
  return 0;
}

// lds_bucket_add_khi reading each bucket in two halves: the second four
// slots only when the first four hold neither the key nor an empty slot
// (occupied slots form a prefix, so a half that ends in an empty slot ends
// the bucket).  Same protocol and results.
__device__ __forceinline__ int lds_bucket_add_khi_half(uint64_t* tab, uint32_t b, uint32_t nb, uint64_t word,
                                                       uint32_t khi, uint32_t cshift, uint32_t max_steps,
                                                       uint64_t& old) {
  const uint32_t tlo = (uint32_t)word, thi = (uint32_t)(word >> 32) & khi;
  const uint64_t add = (word >> cshift) << cshift;
  for (uint32_t p = 0; p < max_steps; p++) {
    uint64_t* line = tab + ((uint64_t)b << 3);
    const uint4* l4 = reinterpret_cast<const uint4*>(line);
    int jm = 8, je = 8;
    {
      const uint4 q0 = l4[0], q1 = l4[1];
      const uint32_t lo[4] = {q0.x, q0.z, q1.x, q1.z};
      const uint32_t hi[4] = {q0.y, q0.w, q1.y, q1.w};
#pragma unroll
      for (int j = 3; j >= 0; j--) {
        jm = (lo[j] == tlo && (hi[j] & khi) == thi) ? j : jm;
        je = hi[j] == 0 ? j : je;
      }
    }
    if (jm == 8 && je == 8) {
      const uint4 q2 = l4[2], q3 = l4[3];
      const uint32_t lo[4] = {q2.x, q2.z, q3.x, q3.z};
      const uint32_t hi[4] = {q2.y, q2.w, q3.y, q3.w};
#pragma unroll
      for (int j = 3; j >= 0; j--) {
        jm = (lo[j] == tlo && (hi[j] & khi) == thi) ? j + 4 : jm;
        je = hi[j] == 0 ? j + 4 : je;
      }
    }
    if (jm < 8) {
      old = atomicAdd((unsigned long long*)&line[jm], (unsigned long long)add);
      return 1;
    }
    if (je < 8) {
      const uint64_t v = atomicCAS((unsigned long long*)&line[je], 0ull, (unsigned long long)word);
      if (v == 0) return 2;
      if ((uint32_t)v == tlo && ((uint32_t)(v >> 32) & khi) == thi) {
        old = atomicAdd((unsigned long long*)&line[je], (unsigned long long)add);
        return 1;
      }
      continue;  // another key took the slot: rescan this bucket
    }
    b = b + 1 == nb ? 0 : b + 1;
  }
  return 0;
}

// lds_bucket_add for 2k >= 32: the key's low 32 bits are all of the slot's
// low word and the occupied bit lies in the high word, so a slot matches
// when its low word equals the key's and its masked high word the tagged
// key's, and it is empty when its high word is zero.  khi = high word of
// the key + occupied mask.
__device__ __forceinline__ int lds_bucket_add_khi(uint64_t* tab, uint32_t b, uint32_t nb, uint64_t word,
                                                  uint32_t khi, uint32_t cshift, uint32_t max_steps,
                                                  uint64_t& old) {
  const uint32_t tlo = (uint32_t)word, thi = (uint32_t)(word >> 32) & khi;
  const uint64_t add = (word >> cshift) << cshift;
  for (uint32_t p = 0; p < max_steps; p++) {
    uint64_t* line = tab + ((uint64_t)b << 3);
    const uint4* l4 = reinterpret_cast<const uint4*>(line);
    const uint4 q0 = l4[0], q1 = l4[1], q2 = l4[2], q3 = l4[3];
    const uint32_t lo[8] = {q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, q3.x, q3.z};
    const uint32_t hi[8] = {q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, q3.y, q3.w};
    int jm = 8, je = 8;
#pragma unroll
    for (int j = 7; j >= 0; j--) {
      jm = (lo[j] == tlo && (hi[j] & khi) == thi) ? j : jm;
      je = hi[j] == 0 ? j : je;
    }
    if (jm < 8) {
      old = atomicAdd((unsigned long long*)&line[jm], (unsigned long long)add);
      return 1;
    }
    if (je < 8) {
      const uint64_t v = atomicCAS((unsigned long long*)&line[je], 0ull, (unsigned long long)word);
      if (v == 0) return 2;
      if ((uint32_t)v == tlo && ((uint32_t)(v >> 32) & khi) == thi) {
        old = atomicAdd((unsigned long long*)&line[je], (unsigned long long)add);
        return 1;
      }
      continue;  // another key took the slot: rescan this bucket
    }
    b = b + 1 == nb ? 0 : b + 1;
  }
  return 0;
}

// ---- LDS tables with an 8-bit tag per slot ----
// Tag of a key from its mix64 hash (low byte; the bucket comes from the high
// bits): 1..255, 0 marks a slot with no tag yet.
__device__ __forceinline__ uint32_t slot_tag(uint64_t h) {
  const uint32_t t8 = (uint32_t)h & 0xFFu;
  return t8 ? t8 : 1u;
}

// 0x80 in every byte of x that is zero, 0 elsewhere (exact: no carries
// between bytes).
__device__ __forceinline__ uint64_t zero_bytes(uint64_t x) {
  const uint64_t m = 0x7F7F7F7F7F7F7F7Full;
  return ~(((x & m) + m) | x | m);
}

// Insert-or-add into an LDS table of nb buckets of eight 8-B slots in the
// global slot format, with tags[8 * b + j] the tag of slot j of bucket b:
// `word` = tagged key | c << cshift.  A step reads the bucket's eight tags (8
// bytes), reads only the slots whose tag matches (to confirm the key), and
// otherwise claims the first slot with no tag by CAS, then writes its tag.
// A slot claimed but not yet tagged by another thread shows as untagged: the
// CAS then fails, and the slot is taken as the key's (add) or skipped.
// Occupied slots stay a prefix of their bucket.  Returns 2 when a slot was
// claimed, 1 when added (old = previous slot value), 0 when max_steps steps
// found no room.
__device__ __forceinline__ int lds_tag_add(uint64_t* tab, uint8_t* tags, uint32_t b, uint32_t nb, uint64_t word,
                                           uint32_t tag, uint64_t kmo, uint32_t cshift, uint32_t max_steps,
                                           uint64_t& old) {
  const uint64_t tagged = word & kmo;
  const uint64_t add = (word >> cshift) << cshift;
  const uint64_t rep = 0x0101010101010101ull * tag;
  uint64_t skip = 0;  // bytes of slots known to hold other keys
  for (uint32_t p = 0; p < max_steps; p++) {
    uint64_t* line = tab + ((uint64_t)b << 3);
    const uint64_t tw = *reinterpret_cast<const uint64_t*>(tags + ((uint64_t)b << 3));
    const uint64_t em = zero_bytes(tw) & ~skip;
    uint64_t mm = zero_bytes(tw ^ rep) & ~skip;
    const uint32_t je = em ? (uint32_t)__builtin_ctzll(em) >> 3 : 8u;
    while (mm) {
      const uint32_t j = (uint32_t)__builtin_ctzll(mm) >> 3;
      if (j >= je) break;
      const uint64_t v = line[j];
      if ((v & kmo) == tagged) {
        old = atomicAdd((unsigned long long*)&line[j], (unsigned long long)add);
        return 1;
      }
      skip |= 0xFFull << (8 * j);
      mm &= mm - 1;
    }
    if (je < 8) {
      const uint64_t v = atomicCAS((unsigned long long*)&line[je], 0ull, (unsigned long long)word);
      if (v == 0) {
        tags[((uint64_t)b << 3) + je] = (uint8_t)tag;
        return 2;
      }
      if ((v & kmo) == tagged) {
        old = atomicAdd((unsigned long long*)&line[je], (unsigned long long)add);
        return 1;
      }
      skip |= 0xFFull << (8 * je);
      continue;  // rescan this bucket past the slot another key took
    }
    b = b + 1 == nb ? 0 : b + 1;
    skip = 0;
  }
  return 0;
}

// ---- wave helpers ----
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bit g of the result is set when byte g of x is non-zero.
__device__ __forceinline__ uint32_t byte_any8(uint64_t x) {
  x |= x >> 4;
  x |= x >> 2;
  x |= x >> 1;
  x &= 0x0101010101010101ULL;
  return (uint32_t)((x * 0x0102040810204080ULL) >> 56);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
// Adds the block's sum of v to *out with one atomic (every thread of the
// block must call it; blockDim <= 1024).  Atomics on one address are
// serialized at the memory side, so one per wavefront of a large grid costs
// milliseconds.
__device__ __forceinline__ void block_sum_add(uint64_t v, unsigned long long* out) {
  __shared__ uint64_t s_part[16];
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t tot = 0;
    for (uint32_t w = 0; w < (blockDim.x + 63) / 64; w++) tot += s_part[w];
    if (tot) atomicAdd(out, (unsigned long long)tot);
  }
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

// Bit g of the result is set when nibble g (lanes 4g..4g+3) of x is non-zero.
__device__ __forceinline__ uint32_t nibble_bits(uint64_t x) {
  x |= x >> 1;
  x |= x >> 2;
  x &= 0x1111111111111111ULL;
  x = (x | (x >> 3)) & 0x0303030303030303ULL;
  x = (x | (x >> 6)) & 0x000F000F000F000FULL;
  x = (x | (x >> 12)) & 0x000000FF000000FFULL;
  return (uint32_t)((x | (x >> 24)) & 0xFFFFu);
}

// Cooperative lookup of 64*NK keys, NK per lane (key i of a lane is key
// number 64*i + lane), four lanes per key.  In round q the lanes 4g..4g+3
// read the 64-byte home bucket of key 16q+g, sixteen bytes (two slots) each:
// one coalesced request per key, as with eight lanes, but half the load,
// shuffle and compare instructions.  All 4*NK loads of a lane are issued
// before any is inspected.  Whether the bucket matched or held an empty slot
// reaches the key's own lane through the round's ballots; the matching lane
// passes the raw count field through `val` (64*NK x u64, LDS).  A key whose
// home bucket is full without a match continues with a full probe.  out[i]
// is the count of key i (0 = absent or pruned); lanes with want[i] = false
// take part in the rounds (their bucket is not read) and get 0.
template <int NK>
__device__ __forceinline__ void coop4_get(const TableView& t, const uint64_t (&key)[NK],
                                          const bool (&want)[NK], uint64_t* val, int lane,
                                          uint64_t (&out)[NK]) {
  constexpr int R = 4 * NK;
  const int g = lane >> 2, sl = lane & 3;
  uint64_t b[NK], tg[NK];
#pragma unroll
  for (int i = 0; i < NK; i++) {
    b[i] = bucket0(t, key[i]);
    tg[i] = key[i] | t.occ;
  }
  ulonglong2 sv[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    const uint64_t bq = __shfl(b[q >> 2], (q & 3) * 16 + g);
    const int wq = __shfl((int)want[q >> 2], (q & 3) * 16 + g);
    sv[q] = wq ? reinterpret_cast<const ulonglong2*>(t.slots + (bq << 3))[sl] : make_ulonglong2(0, 0);
  }
  uint64_t mbs[R], ebs[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    const uint64_t tq = __shfl(tg[q >> 2], (q & 3) * 16 + g);
    const bool m0 = (sv[q].x & t.kmo) == tq;
    const bool m1 = (sv[q].y & t.kmo) == tq;
    if (m0 | m1) val[16 * q + g] = (m0 ? sv[q].x : sv[q].y) >> t.cshift;
    mbs[q] = __ballot(m0 | m1);
    ebs[q] = __ballot((sv[q].x == 0) | (sv[q].y == 0));
  }
  wave_sync();
  const int qq = lane >> 4, sh = (lane & 15) * 4;
#pragma unroll
  for (int i = 0; i < NK; i++) {
    const uint64_t mq = qq == 0 ? mbs[4 * i] : qq == 1 ? mbs[4 * i + 1] : qq == 2 ? mbs[4 * i + 2] : mbs[4 * i + 3];
    const uint64_t eq = qq == 0 ? ebs[4 * i] : qq == 1 ? ebs[4 * i + 1] : qq == 2 ? ebs[4 * i + 2] : ebs[4 * i + 3];
    const bool hit = (mq >> sh) & 0xFull;
    const bool empty = (eq >> sh) & 0xFull;
    uint64_t f = 0;
    if (want[i]) {
      if (hit) {
        f = val[64 * i + lane];
        if (f == t.cmax) {
          const uint64_t* o = ovf_find(t, tg[i]);
          f = o ? *o : f;
        }
      } else if (!empty) {
        f = table_get(t, key[i]);
      }
    }
    out[i] = f;
  }
  wave_sync();
}

// Cooperative +1 of the 64 keys held one per lane (want = lane has a key).
// Eight lanes per key: in round q the lanes 8g..8g+7 read the eight slots of
// the home bucket of lane 8q+g's key (one coalesced 64-byte request per key,
// eight loads in flight per lane).  In each group the lane holding
// the key adds 1; otherwise the lowest empty slot of the snapshot is claimed
// with CAS.  The snapshot may be stale, but a slot's key never changes once
// set and every "empty" is confirmed by the CAS, so no key is inserted twice.
// All eight rounds' atomics are issued before any result is inspected, so a
// wavefront waits for one atomic round trip per 64 keys, not eight.
// Keys left unresolved (lost CAS race, full home bucket) are finished by
// their own lane with table_add; `spilled` reports a key that found no room
// within max_probe buckets.  Returns the number of new slots this lane
// claimed.
__device__ __forceinline__ uint32_t coop_add64(const TableView& t, uint64_t key, bool want,
                                               int lane, uint32_t* err, bool& spilled) {
  const uint64_t tagged = key | t.occ;
  const uint64_t b = bucket0(t, key);
  const int g = lane >> 3, sl = lane & 7;
  const uint64_t one = 1ull << t.cshift;
  uint64_t todo = __ballot(want);
  uint64_t sv[8], bq[8], tq[8], old[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    bq[q] = __shfl(b, q * 8 + g);
    sv[q] = t.slots[(bq[q] << 3) + sl];
  }
  uint32_t kind = 0;  // per round, 2 bits: 1 = add issued, 2 = CAS issued
#pragma unroll
  for (int q = 0; q < 8; q++) {
    tq[q] = __shfl(tagged, q * 8 + g);
    old[q] = 0;
    const bool act = (todo >> (q * 8 + g)) & 1ull;
    const bool m = act && (sv[q] & t.kmo) == tq[q];
    const bool e = act && sv[q] == 0;
    const uint32_t mb = (uint32_t)((__ballot(m) >> (8 * g)) & 0xFFu);
    const uint32_t eb = (uint32_t)((__ballot(e) >> (8 * g)) & 0xFFu);
    unsigned long long* slot = (unsigned long long*)(t.slots + (bq[q] << 3) + sl);
    if (m) {
      old[q] = atomicAdd(slot, (unsigned long long)one);
      kind |= 1u << (2 * q);
    } else if (act && !mb && eb && sl == __builtin_ctz(eb)) {
      old[q] = atomicCAS(slot, 0ULL, (unsigned long long)(tq[q] | one));
      kind |= 2u << (2 * q);
    }
  }
  uint32_t claimed = 0;
  uint64_t carry = 0;  // bit w: the +1 of lane w's key wrapped its count field
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const uint32_t kq = (kind >> (2 * q)) & 3u;
    bool done = false, wrap = false;
    if (kq == 1u) {
      wrap = (old[q] >> t.cshift) == t.cmax;
      done = true;
    } else if (kq == 2u) {
      if (old[q] == 0) {
        claimed++;
        done = true;
      } else if ((old[q] & t.kmo) == tq[q]) {
        unsigned long long* slot = (unsigned long long*)(t.slots + (bq[q] << 3) + sl);
        const uint64_t o2 = atomicAdd(slot, (unsigned long long)one);
        wrap = (o2 >> t.cshift) == t.cmax;
        done = true;
      }
    }
    const uint32_t gd = byte_any8(__ballot(done));
    todo &= ~((uint64_t)gd << (8 * q));
    carry |= (uint64_t)byte_any8(__ballot(wrap)) << (8 * q);
  }
  // the key's own lane records a carry in its own table's side table
  if ((carry >> lane) & 1ull) ovf_add(t, tagged, t.cmax + 1, err);
  spilled = false;
  if ((todo >> lane) & 1ull) {
    const uint32_t r = table_add(t, key, 1, err);
    if (r == 2u) spilled = true;
    else claimed += r;
  }
  return claimed;
}

// Cooperative +1 of the 64 keys held one per lane, four lanes per key (the
// round layout of coop4_get: in round q the lanes 4g..4g+3 hold the sixteen
// bytes 16*sl.. of the home bucket of key 16q+g, lane 16q+g's key).  In each
// group the lane holding the key adds 1 to it; otherwise the lowest empty
// slot of the snapshot is claimed with CAS.  Same protocol and guarantees as
// coop_add64 (no key inserted twice; all rounds' atomics in flight before any
// result is inspected; unresolved keys finished by their own lane).
// `c` (1 <= c <= cmax) is the number of occurrences the lane adds.
__device__ __forceinline__ uint32_t coop4_add(const TableView& t, uint64_t key, bool want, int lane,
                                              uint32_t* err, bool& spilled, uint64_t c = 1) {
  const uint64_t tagged = key | t.occ;
  const uint64_t b = bucket0(t, key);
  const int g = lane >> 2, sl = lane & 3;
  const uint64_t one = c << t.cshift;
  uint64_t todo = __ballot(want);
  ulonglong2 sv[4];
  uint64_t bq[4], tq[4], old[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    bq[q] = __shfl(b, q * 16 + g);
    sv[q] = reinterpret_cast<const ulonglong2*>(t.slots + (bq[q] << 3))[sl];
  }
  uint32_t kind = 0;  // per round, 3 bits: 1 = add issued, 2 = CAS issued; bit 2 = second slot
#pragma unroll
  for (int q = 0; q < 4; q++) {
    tq[q] = __shfl(tagged, q * 16 + g);
    old[q] = 0;
    const bool act = (todo >> (q * 16 + g)) & 1ull;
    const bool m0 = act && (sv[q].x & t.kmo) == tq[q];
    const bool m1 = act && (sv[q].y & t.kmo) == tq[q];
    const bool e0 = act && sv[q].x == 0;
    const bool e1 = act && sv[q].y == 0;
    const uint32_t mb = (uint32_t)((__ballot(m0 | m1) >> (4 * g)) & 0xFu);
    const uint32_t eb = (uint32_t)((__ballot(e0 | e1) >> (4 * g)) & 0xFu);
    const uint64_t oneq = __shfl(one, q * 16 + g);
    if (m0 | m1) {
      const int j = m0 ? 0 : 1;
      old[q] = atomicAdd((unsigned long long*)(t.slots + (bq[q] << 3) + 2 * sl + j), (unsigned long long)oneq);
      kind |= (1u | ((uint32_t)j << 2)) << (3 * q);
    } else if (act && !mb && eb && sl == __builtin_ctz(eb)) {
      const int j = e0 ? 0 : 1;
      old[q] = atomicCAS((unsigned long long*)(t.slots + (bq[q] << 3) + 2 * sl + j), 0ULL,
                         (unsigned long long)(tq[q] | oneq));
      kind |= (2u | ((uint32_t)j << 2)) << (3 * q);
    }
  }
  uint32_t claimed = 0;
  uint64_t carry = 0;  // bit w: the +1 of lane w's key wrapped its count field
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint32_t kq = (kind >> (3 * q)) & 3u;
    const int j = (kind >> (3 * q + 2)) & 1u;
    bool done = false, wrap = false;
    const uint64_t cq = __shfl(c, q * 16 + g);
    if (kq == 1u) {
      wrap = (old[q] >> t.cshift) + cq > t.cmax;
      done = true;
    } else if (kq == 2u) {
      if (old[q] == 0) {
        claimed++;
        done = true;
      } else if ((old[q] & t.kmo) == tq[q]) {
        const uint64_t o2 = atomicAdd((unsigned long long*)(t.slots + (bq[q] << 3) + 2 * sl + j),
                                      (unsigned long long)(cq << t.cshift));
        wrap = (o2 >> t.cshift) + cq > t.cmax;
        done = true;
      }
    }
    todo &= ~((uint64_t)nibble_bits(__ballot(done)) << (16 * q));
    carry |= (uint64_t)nibble_bits(__ballot(wrap)) << (16 * q);
  }
  if ((carry >> lane) & 1ull) ovf_add(t, tagged, t.cmax + 1, err);
  spilled = false;
  if ((todo >> lane) & 1ull) {
    const uint32_t r = table_add(t, key, c, err);
    if (r == 2u) spilled = true;
    else claimed += r;
  }
  return claimed;
}

// Locality sort key of a read, fed its bases in order: bits 39..8 hold the
// smallest mixed hash over its canonical 16-mers (a minimizer; an N base
// reads as A, the key only orders reads), bits 7..0 the read's start relative
// to that 16-mer on the 16-mer's canonical strand.  Reads that share the
// minimizer overlap one locus of the transcriptome and sort by where they
// start along it, so neighbours in the order share most k-mers: their table
// lines are then reused from L2, and their windows merge before counting.
// Reads shorter than 16 bases sort last.
struct MinKey {
  uint32_t f = 0, r = 0, best = 0xFFFFFFFFu, bpos = 0;
  bool bfwd = true;
  __device__ __forceinline__ void push(uint32_t b, uint32_t pos) {
    f = (f << 2) | b;
    r = (r >> 2) | ((3u - b) << 30);
    if (pos >= 15) {
      const bool fw = f <= r;
      uint32_t x = fw ? f : r;
      x *= 0x9E3779B1u;
      x ^= x >> 16;
      x *= 0x85EBCA6Bu;
      x ^= x >> 13;
      if (x < best) {
        best = x;
        bpos = pos - 15;
        bfwd = fw;
      }
    }
  }
  __device__ __forceinline__ uint64_t key(uint32_t L) const {
    uint32_t rel = bfwd ? bpos : L - 16 - bpos;
    rel = rel > 255u ? 255u : rel;
    return ((uint64_t)best << 8) | (255u - rel);
  }
};
constexpr int kOrderKeyBits = 40;

}  // namespace tk
