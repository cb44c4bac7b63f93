// tk_perread.h -- per-read device routines shared by the single-sample
// kernels (tk_lib.hip) and the batched-component kernels (tk_batch.hip).
#pragma once
#include "tk_device.h"
#include "../../include/trinity_gpu.h"

namespace tk {

struct StatsParams {
  uint32_t min_count;
  uint32_t missing;
  int32_t nwin_policy;
  int32_t median_policy;
};

// kth smallest (0-based) of vals[0..n) by radix select over bits hb..0.
__device__ inline uint32_t wave_select(const uint32_t* vals, uint32_t n, uint32_t kth, int hb,
                                int lane) {
  uint32_t prefix = 0, pmask = 0;
  for (int bit = hb; bit >= 0; --bit) {
    uint32_t c = 0;
    for (uint32_t i = lane; i < n; i += 64) {
      const uint32_t v = vals[i];
      c += (((v & pmask) == prefix) && !((v >> bit) & 1u)) ? 1u : 0u;
    }
    c = wave_sum_u32(c);
    if (kth >= c) {
      kth -= c;
      prefix |= 1u << bit;
    }
    pmask |= 1u << bit;
  }
  return prefix;
}

// Value of one window for the coverage statistics: `has` = the window takes
// part (fastaToKmerCoverageStats.cpp skips or counts invalid windows by
// policy), v = its count, or `missing` below min_count.
__device__ __forceinline__ void window_value(const StatsParams& sp, bool in_read, bool valid, uint64_t c,
                                             bool& has, uint32_t& v) {
  has = false;
  v = 0;
  if (!in_read) return;
  if (valid) {
    v = (c != 0 && c >= sp.min_count) ? (c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c) : sp.missing;
    has = true;
  } else if (sp.nwin_policy == TK_NWIN_AS_MISSING) {
    v = sp.missing;
    has = true;
  }
}

// kth smallest (0-based) of the values held two per lane (ha/va, hb_/vb; a
// value takes part when its flag is set), radix select over bits hb..0.
__device__ __forceinline__ uint32_t wave_select2(bool ha, uint32_t va, bool hb_, uint32_t vb, uint32_t kth,
                                                 int hb) {
  uint32_t prefix = 0, pmask = 0;
  for (int bit = hb; bit >= 0; --bit) {
    const bool za = ha && (va & pmask) == prefix && !((va >> bit) & 1u);
    const bool zb = hb_ && (vb & pmask) == prefix && !((vb >> bit) & 1u);
    const uint32_t c = (uint32_t)__popcll(__ballot(za)) + (uint32_t)__popcll(__ballot(zb));
    if (kth >= c) {
      kth -= c;
      prefix |= 1u << bit;
    }
    pmask |= 1u << bit;
  }
  return prefix;
}

__device__ __forceinline__ void sq_add(uint64_t v, uint64_t& lo, uint64_t& hi) {
  const uint64_t sq = v * v;
  lo += sq;
  hi += (lo < sq);
}

__device__ __forceinline__ void wave_sum_u128(uint64_t& lo, uint64_t& hi) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t ol = __shfl_xor(lo, o);
    const uint64_t oh = __shfl_xor(hi, o);
    const uint64_t nl = lo + ol;
    hi = hi + oh + (nl < lo);
    lo = nl;
  }
}

// The statistics of a read of at most 128 windows, two per lane with their
// values in registers: win(pos, fwd, valid) gives the forward k-mer at
// window `pos` (< 128) and whether it is all ACGT; only windows below nwin
// count.  Every lane calls win for both of its positions.
template <class Win>
__device__ __forceinline__ void stats_short_core(uint32_t nwin, int k, int canonical, const TableView& t,
                                                 const StatsParams& sp, uint64_t* tmp, int lane, Win&& win,
                                                 uint32_t& n, uint32_t& med, uint64_t& sum, uint64_t& sqlo,
                                                 uint64_t& sqhi) {
  const uint32_t pa = lane, pb = 64 + lane;
  uint64_t fa = 0, fb = 0;
  bool va = false, vb = false;
  win(pa, fa, va);
  win(pb, fb, vb);
  va = va && pa < nwin;
  vb = vb && pb < nwin;
  const uint64_t keys[2] = {va ? canon_key(fa, k, canonical) : 0, vb ? canon_key(fb, k, canonical) : 0};
  const bool want[2] = {va, vb};
  uint64_t c[2];
  coop4_get<2>(t, keys, want, tmp, lane, c);
  bool ha, hb_;
  uint32_t xa, xb;
  window_value(sp, pa < nwin, va, c[0], ha, xa);
  window_value(sp, pb < nwin, vb, c[1], hb_, xb);
  n = (uint32_t)__popcll(__ballot(ha)) + (uint32_t)__popcll(__ballot(hb_));
  sum = (uint64_t)xa + xb;
  sqlo = sqhi = 0;
  sq_add(xa, sqlo, sqhi);
  sq_add(xb, sqlo, sqhi);
  sum = wave_sum_u64(sum);
  wave_sum_u128(sqlo, sqhi);
  med = 0;
  if (n) {
    const uint32_t vmax = wave_max_u32(xa > xb ? xa : xb);
    const int hb = 31 - __clz(vmax | 1u);
    if (sp.median_policy == TK_MEDIAN_LOWER) {
      med = wave_select2(ha, xa, hb_, xb, (n - 1) / 2, hb);
    } else if (sp.median_policy == TK_MEDIAN_MEAN_FLOOR && !(n & 1)) {
      const uint64_t a = wave_select2(ha, xa, hb_, xb, n / 2 - 1, hb);
      const uint64_t b = wave_select2(ha, xa, hb_, xb, n / 2, hb);
      med = (uint32_t)((a + b) / 2);
    } else {
      med = wave_select2(ha, xa, hb_, xb, n / 2, hb);
    }
  }
}

__device__ __forceinline__ void write_stat(int lane, uint32_t med, uint32_t n, uint64_t sum, uint64_t sqlo,
                                           uint64_t sqhi, tk_read_stat* out) {
  if (lane == 0) {
    tk_read_stat st;
    st.median = med;
    st.nwin = n;
    st.sum = sum;
    st.sumsq_lo = sqlo;
    st.sumsq_hi = sqhi;
    *out = st;
  }
}

// fastaToKmerCoverageStats for read r, one wavefront: exact integer sums and
// the median by radix select.  Reads of up to 128 windows keep their values
// in registers (two per lane); longer reads stage them in `vals` (LDS, >=
// nwin entries) 128 windows at a time.  tmp: 128 x u64 LDS scratch of the
// cooperative lookup.  Lane 0 writes *out.
// kShortOnly: the caller guarantees nwin <= 128 (the long-read path is not
// compiled, which frees registers).
template <bool kShortOnly = false>
__device__ inline void stats_read(const ReadsView& rv, uint64_t r, int k, int canonical,
                                  const TableView& t, const StatsParams& sp, uint32_t* vals,
                                  uint64_t* tmp, int lane, tk_read_stat* out) {
  const uint32_t L = rv.len[r];
  const uint32_t nwin = L >= (uint32_t)k ? L - k + 1 : 0;
  const uint64_t wb = rv.word_off[r];
  uint32_t n = 0, vmax = 0, med = 0;
  uint64_t sum = 0, sqlo = 0, sqhi = 0;
  if (kShortOnly || nwin <= 128) {
    stats_short_core(nwin, k, canonical, t, sp, tmp, lane,
                     [&](uint32_t pos, uint64_t& f, bool& v) {
                       f = 0;
                       v = false;
                       if (pos < nwin) window_at(rv, wb, pos, k, f, v);
                     },
                     n, med, sum, sqlo, sqhi);
  } else {
    // windows are taken 128 at a time (two per lane), values appended to
    // vals in window order
    for (uint32_t base = 0; base < nwin; base += 128) {
      const uint32_t pa = base + lane, pb = base + 64 + lane;
      uint64_t fa = 0, fb = 0;
      bool va = false, vb = false;
      if (pa < nwin) window_at(rv, wb, pa, k, fa, va);
      if (pb < nwin) window_at(rv, wb, pb, k, fb, vb);
      const uint64_t keys[2] = {va ? canon_key(fa, k, canonical) : 0, vb ? canon_key(fb, k, canonical) : 0};
      const bool want[2] = {va, vb};
      uint64_t c[2];
      coop4_get<2>(t, keys, want, tmp, lane, c);
#pragma unroll
      for (int half = 0; half < 2; half++) {
        bool has;
        uint32_t v;
        window_value(sp, (half ? pb : pa) < nwin, half ? vb : va, c[half], has, v);
        const uint64_t m = __ballot(has);
        if (has) {
          vals[n + __popcll(m & ((1ull << lane) - 1))] = v;
          sum += v;
          sq_add(v, sqlo, sqhi);
          vmax = v > vmax ? v : vmax;
        }
        n += (uint32_t)__popcll(m);
      }
    }
    wave_sync();
    sum = wave_sum_u64(sum);
    wave_sum_u128(sqlo, sqhi);
    if (n) {
      vmax = wave_max_u32(vmax);
      const int hb = 31 - __clz(vmax | 1u);
      if (sp.median_policy == TK_MEDIAN_LOWER) {
        med = wave_select(vals, n, (n - 1) / 2, hb, lane);
      } else if (sp.median_policy == TK_MEDIAN_MEAN_FLOOR && !(n & 1)) {
        const uint64_t a = wave_select(vals, n, n / 2 - 1, hb, lane);
        const uint64_t b = wave_select(vals, n, n / 2, hb, lane);
        med = (uint32_t)((a + b) / 2);
      } else {
        med = wave_select(vals, n, n / 2, hb, lane);
      }
    }
  }
  write_stat(lane, med, n, sum, sqlo, sqhi, out);
  wave_sync();
}

// Block-shared LDS cache of table counts (tagged key -> count), for reads
// taken in the locality order: a read's windows are mostly its
// predecessors', so the block looks a key up in HBM once and its later reads
// find it here.  A slot's key is claimed by CAS and its value written after;
// a reader that finds the key before the value (kNotReady) treats it as a
// miss, so a lookup never returns a wrong count.  Slots are never reused.
constexpr uint32_t kNotReady = 0xFFFFFFFFu;
constexpr int kCacheProbe = 16;

struct CountCache {
  uint64_t* keys;   // 0 = empty
  uint32_t* vals;   // kNotReady until written
  uint32_t mask;    // slots - 1
};

__device__ __forceinline__ uint32_t cache_home(uint64_t tagged, uint32_t mask) {
  return (uint32_t)((tagged * 0x9E3779B97F4A7C15ULL) >> 40) & mask;
}

__device__ __forceinline__ bool cache_get(const CountCache& c, uint64_t tagged, uint32_t& v) {
  uint32_t h = cache_home(tagged, c.mask);
  for (int p = 0; p < kCacheProbe; p++) {
    const uint64_t k = c.keys[h];
    if (k == tagged) {
      v = c.vals[h];
      return v != kNotReady;
    }
    if (k == 0) return false;
    h = (h + 1) & c.mask;
  }
  return false;
}

__device__ __forceinline__ void cache_put(const CountCache& c, uint64_t tagged, uint32_t v) {
  uint32_t h = cache_home(tagged, c.mask);
  for (int p = 0; p < kCacheProbe; p++) {
    uint64_t k = c.keys[h];
    if (k == 0) {
      k = atomicCAS((unsigned long long*)&c.keys[h], 0ULL, (unsigned long long)tagged);
      if (k == 0) {
        c.vals[h] = v;
        return;
      }
    }
    if (k == tagged) return;
    h = (h + 1) & c.mask;
  }
}

// stats_read for a read of at most 128 windows, with the block's count
// cache in front of the table: only keys the cache does not hold are looked
// up (coop4_get reads no bucket for the others), and those are added to it.
__device__ inline void stats_read_cached(const ReadsView& rv, uint64_t r, int k, int canonical,
                                         const TableView& t, const StatsParams& sp, const CountCache& cc,
                                         uint64_t* tmp, int lane, tk_read_stat* out) {
  const uint32_t L = rv.len[r];
  const uint32_t nwin = L >= (uint32_t)k ? L - k + 1 : 0;
  const uint64_t wb = rv.word_off[r];
  const uint32_t pa = lane, pb = 64 + lane;
  uint64_t fa = 0, fb = 0;
  bool va = false, vb = false;
  if (pa < nwin) window_at(rv, wb, pa, k, fa, va);
  if (pb < nwin) window_at(rv, wb, pb, k, fb, vb);
  const uint64_t keys[2] = {va ? canon_key(fa, k, canonical) : 0, vb ? canon_key(fb, k, canonical) : 0};
  uint32_t cva = 0, cvb = 0;
  const bool hita = va && cache_get(cc, keys[0] | t.occ, cva);
  const bool hitb = vb && cache_get(cc, keys[1] | t.occ, cvb);
  const bool want[2] = {va && !hita, vb && !hitb};
  uint64_t c[2];
  coop4_get<2>(t, keys, want, tmp, lane, c);
  if (want[0] && c[0] < kNotReady) cache_put(cc, keys[0] | t.occ, (uint32_t)c[0]);
  if (want[1] && c[1] < kNotReady) cache_put(cc, keys[1] | t.occ, (uint32_t)c[1]);
  if (hita) c[0] = cva;
  if (hitb) c[1] = cvb;
  bool ha, hb_;
  uint32_t xa, xb;
  window_value(sp, pa < nwin, va, c[0], ha, xa);
  window_value(sp, pb < nwin, vb, c[1], hb_, xb);
  const uint32_t n = (uint32_t)__popcll(__ballot(ha)) + (uint32_t)__popcll(__ballot(hb_));
  uint64_t sum = (uint64_t)xa + xb, sqlo = 0, sqhi = 0;
  sq_add(xa, sqlo, sqhi);
  sq_add(xb, sqlo, sqhi);
  sum = wave_sum_u64(sum);
  wave_sum_u128(sqlo, sqhi);
  uint32_t med = 0;
  if (n) {
    const uint32_t vmax = wave_max_u32(xa > xb ? xa : xb);
    const int hb = 31 - __clz(vmax | 1u);
    if (sp.median_policy == TK_MEDIAN_LOWER) {
      med = wave_select2(ha, xa, hb_, xb, (n - 1) / 2, hb);
    } else if (sp.median_policy == TK_MEDIAN_MEAN_FLOOR && !(n & 1)) {
      const uint64_t a = wave_select2(ha, xa, hb_, xb, n / 2 - 1, hb);
      const uint64_t b = wave_select2(ha, xa, hb_, xb, n / 2, hb);
      med = (uint32_t)((a + b) / 2);
    } else {
      med = wave_select2(ha, xa, hb_, xb, n / 2, hb);
    }
  }
  if (lane == 0) {
    tk_read_stat st;
    st.median = med;
    st.nwin = n;
    st.sum = sum;
    st.sumsq_lo = sqlo;
    st.sumsq_hi = sqhi;
    *out = st;
  }
  wave_sync();
}

// Component table: 64-byte buckets of four 16-byte slots (tagged key, ~comp);
// buckets [boff, boff + bmask] belong to one table.
__device__ inline void ctable_add(uint64_t* slots, uint64_t boff, uint64_t bmask, uint64_t occ,
                                  uint64_t key, uint32_t comp, uint32_t& mynew, uint32_t* err) {
  const uint64_t tagged = key | occ;
  uint64_t b = mix64(key) & bmask;
  for (uint64_t p = 0; p <= bmask && p < 4096; p++) {
    uint64_t* line = slots + ((boff + b) << 3);
    for (int j = 0; j < 4; j++) {
      uint64_t s = line[2 * j];
      if (s == 0) {
        s = atomicCAS((unsigned long long*)&line[2 * j], 0ULL, (unsigned long long)tagged);
        if (s == 0) { mynew++; s = tagged; }
      }
      if (s == tagged) {
        atomicMax((unsigned int*)&line[2 * j + 1], ~comp);
        return;
      }
    }
    b = (b + 1) & bmask;
  }
  atomicOr(err, 1u);
}

__device__ __forceinline__ uint32_t ctable_get(const uint64_t* slots, uint64_t boff, uint64_t bmask,
                                               uint64_t occ, uint64_t key) {
  const uint64_t tagged = key | occ;
  uint64_t b = mix64(key) & bmask;
  for (uint64_t p = 0; p <= bmask && p < 4096; p++) {
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(slots + ((boff + b) << 3));
    // the first slot holding the key or empty decides; the slots are taken
    // one by one into registers (an array indexed after early returns went
    // to scratch memory)
    uint32_t res = 0xFFFFFFFFu;
    bool done = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const ulonglong2 v = l2[j];
      const bool hit = !done && v.x == tagged;
      const bool empty = !done && v.x == 0;
      res = hit ? ~(uint32_t)v.y : res;
      done = done || hit || empty;
    }
    if (done) return res;
    b = (b + 1) & bmask;
  }
  return 0xFFFFFFFFu;
}

// Packed component table (k <= 25): 64-byte buckets of eight 8-byte slots
// [vmax - comp : 63-2k | occupied : 1 | key : 2k], vmax = 2^(63-2k) - 1, so
// a new key costs one CAS and the largest word of one key holds its
// smallest comp; comp <= vmax.
__device__ inline void pctable_add(uint64_t* slots, uint64_t boff, uint64_t bmask, int k, uint64_t key,
                                   uint32_t comp, uint32_t* err) {
  const uint64_t occ = 1ull << (2 * k), km = (occ << 1) - 1, tagged = key | occ;
  const uint64_t vmax = (1ull << (63 - 2 * k)) - 1;
  const uint64_t word = ((vmax - comp) << (2 * k + 1)) | tagged;
  uint64_t b = mix64(key) & bmask;
  for (uint64_t p = 0; p <= bmask && p < 4096; p++) {
    uint64_t* line = slots + ((boff + b) << 3);
    for (int j = 0; j < 8; j++) {
      uint64_t s = line[j];
      if (s == 0) {
        s = atomicCAS((unsigned long long*)&line[j], 0ULL, (unsigned long long)word);
        if (s == 0) return;
      }
      if ((s & km) == tagged) {
        if (s < word) atomicMax((unsigned long long*)&line[j], (unsigned long long)word);
        return;
      }
    }
    b = (b + 1) & bmask;
  }
  atomicOr(err, 1u);
}

__device__ __forceinline__ uint32_t pctable_comp(uint64_t s, int k) {
  return (uint32_t)(((1ull << (63 - 2 * k)) - 1) - (s >> (2 * k + 1)));
}

__device__ inline uint32_t pctable_get(const uint64_t* slots, uint64_t boff, uint64_t bmask, int k,
                                       uint64_t key) {
  const uint64_t occ = 1ull << (2 * k), km = (occ << 1) - 1, tagged = key | occ;
  uint64_t b = mix64(key) & bmask;
  for (uint64_t p = 0; p <= bmask && p < 4096; p++) {
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(slots + ((boff + b) << 3));
    // first slot holding the key or empty decides (slots in registers, as
    // in ctable_get)
    uint32_t res = 0xFFFFFFFFu;
    bool done = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const ulonglong2 v = l2[j];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint64_t w = h ? v.y : v.x;
        const bool hit = !done && (w & km) == tagged;
        const bool empty = !done && w == 0;
        res = hit ? pctable_comp(w, k) : res;
        done = done || hit || empty;
      }
    }
    if (done) return res;
    b = (b + 1) & bmask;
  }
  return 0xFFFFFFFFu;
}

// coop_cget64 for the packed table: the four lanes of a group read two
// 8-byte slots each of the key's home bucket.
__device__ __forceinline__ uint32_t coop_pget64(const uint64_t* slots, uint64_t boff, uint64_t bmask, int k,
                                                uint64_t key, bool want, uint32_t* tmp, int lane) {
  const uint64_t occ = 1ull << (2 * k), km = (occ << 1) - 1, tagged = key | occ;
  const uint64_t b = boff + (mix64(key) & bmask);
  const int g = lane >> 2, sl = lane & 3;
  tmp[lane] = 0xFFFFFFFFu;
  ulonglong2 sv[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint64_t bq = __shfl(b, q * 16 + g);
    sv[q] = reinterpret_cast<const ulonglong2*>(slots + (bq << 3))[sl];
  }
  wave_sync();
  uint64_t pending = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint64_t tq = __shfl(tagged, q * 16 + g);
    const bool mx = (sv[q].x & km) == tq, my = (sv[q].y & km) == tq;
    if (mx) tmp[q * 16 + g] = pctable_comp(sv[q].x, k);
    else if (my) tmp[q * 16 + g] = pctable_comp(sv[q].y, k);
    const uint32_t gm = nibble_bits(__ballot(mx || my));
    const uint32_t ge = nibble_bits(__ballot(sv[q].x == 0 || sv[q].y == 0));
    pending |= (uint64_t)(~gm & ~ge & 0xFFFFu) << (16 * q);
  }
  wave_sync();
  uint32_t c = tmp[lane];
  if (want && ((pending >> lane) & 1ull)) c = pctable_get(slots, boff, bmask, k, key);
  wave_sync();
  return want ? c : 0xFFFFFFFFu;
}

// Cooperative component-table lookup of the 64 keys held one per lane: in
// round q the four lanes 4g..4g+3 read the four 16-byte slots of the bucket
// of lane 16q+g's key (one 64-byte request per key, four loads in flight per
// lane).  Returns the component of the calling lane's key (0xFFFFFFFF when
// absent).  tmp: 64 x u32 LDS scratch.
__device__ __forceinline__ uint32_t coop_cget64(const uint64_t* slots, uint64_t boff, uint64_t bmask,
                                                uint64_t occ, uint64_t key, bool want, uint32_t* tmp,
                                                int lane) {
  const uint64_t tagged = key | occ;
  const uint64_t b = boff + (mix64(key) & bmask);
  const int g = lane >> 2, sl = lane & 3;
  tmp[lane] = 0xFFFFFFFFu;
  ulonglong2 sv[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint64_t bq = __shfl(b, q * 16 + g);
    sv[q] = reinterpret_cast<const ulonglong2*>(slots + (bq << 3))[sl];
  }
  wave_sync();
  uint64_t pending = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint64_t tq = __shfl(tagged, q * 16 + g);
    const bool m = sv[q].x == tq;
    if (m) tmp[q * 16 + g] = ~(uint32_t)sv[q].y;
    const uint32_t gm = nibble_bits(__ballot(m));
    const uint32_t ge = nibble_bits(__ballot(sv[q].x == 0));
    pending |= (uint64_t)(~gm & ~ge & 0xFFFFu) << (16 * q);
  }
  wave_sync();
  uint32_t c = tmp[lane];
  if (want && ((pending >> lane) & 1ull)) c = ctable_get(slots, boff, bmask, occ, key);
  wave_sync();
  return want ? c : 0xFFFFFFFFu;
}

// ReadsToTranscripts vote of read r, one wavefront: component ids of the
// hitting windows staged in `hits` (LDS), distinct ids walked in ascending
// order so the first strictly largest count wins (ties -> smallest id).
// kPacked: the component table is the packed layout (pctable_add).
template <bool kPacked = false>
__device__ inline void vote_read(const ReadsView& rv, uint64_t r, int k, int canonical,
                                 const uint64_t* slots, uint64_t boff, uint64_t bmask,
                                 uint32_t* hits, uint32_t* tmp, int lane, tk_r2t_vote* out) {
  const uint64_t occ = 1ull << (2 * k);
    const uint32_t L = rv.len[r];
    const uint32_t nwin = L >= (uint32_t)k ? L - k + 1 : 0;
    const uint64_t wb = rv.word_off[r];
    uint32_t n = 0;
    for (uint32_t base = 0; base < nwin; base += 64) {
      const uint32_t p = base + lane;
      uint64_t fwd = 0;
      bool valid = false;
      if (p < nwin) window_at(rv, wb, p, k, fwd, valid);
      const uint64_t key = valid ? canon_key(fwd, k, canonical) : 0;
      const uint32_t c = kPacked ? coop_pget64(slots, boff, bmask, k, key, valid, tmp, lane)
                                 : coop_cget64(slots, boff, bmask, occ, key, valid, tmp, lane);
      const bool has = c != 0xFFFFFFFFu;
      const uint64_t m = __ballot(has);
      if (has) hits[n + __popcll(m & ((1ull << lane) - 1))] = c;
      n += (uint32_t)__popcll(m);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Walk the distinct components in ascending order; keep the first with
    // the strictly largest number of hits (ties -> smallest id).
    uint32_t best = 0xFFFFFFFFu, bestn = 0;
    uint64_t prev = 0;  // candidates are > prev-1
    bool first = true;
    while (n) {
      uint32_t cand = 0xFFFFFFFFu;
      for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t v = hits[i];
        if ((first || v > (uint32_t)prev) && v < cand) cand = v;
      }
      cand = wave_min_u32(cand);
      if (cand == 0xFFFFFFFFu) break;
      uint32_t cnt = 0;
      for (uint32_t i = lane; i < n; i += 64) cnt += hits[i] == cand;
      cnt = wave_sum_u32(cnt);
      if (cnt > bestn) { bestn = cnt; best = cand; }
      prev = cand;
      first = false;
    }
    if (lane == 0) {
      tk_r2t_vote v;
      v.comp = best;
      v.hits = bestn;
      v.nwin = nwin;
      *out = v;
    }
  wave_sync();
}

}  // namespace tk
