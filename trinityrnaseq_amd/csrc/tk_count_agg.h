// tk_count_agg.h -- the block-merging count pass (k_count_agg), shared by
// the single-table count (tk_lib.hip) and the batched components
// (tk_batch.hip), whose blocks each merge one component's reads.
#pragma once
#include <algorithm>

#include "tk_device.h"

namespace tk {

constexpr int kReadsPerBlock = 64;

// Block-aggregated count for reads taken in the locality order, where the
// 64 reads of a block largely overlap one locus.  Every device-scope atomic
// on the table is a round trip to the memory side (the eight XCDs' L2s are
// not coherent with each other), so the count pass is bound by the number
// of atomics, not by bytes: the block first merges its windows, 256 per
// round, in an LDS hash table of words `tagged key | occurrences << (2k+1)`
// (the global slot format), and only when the table could pass 3/4 load
// (or a count kMaxRun) does each distinct key reach the global table, once,
// with its number of occurrences, through the cooperative four-lane add.
// Each wavefront compacts and flushes its own quarter of the LDS table.
// Needs cmax >= kMaxRun (k <= 25), so neither the LDS word nor one global
// add carries more than once.
//
// EMIT: instead of adding them to the table, a flush appends the block's
// merged words to `out` for the region-resident passes below.  The grid is
// then persistent (each block takes groups of 64 reads grid-stride) and a
// block reserves the stream `chunk` words at a time (chunk >= SLOTS, the
// most one flush emits), so the shared cursor sees one atomic per chunk,
// not one per flush.  A flush that does not fit below `cap` is added to the
// table as above; every reserved word below `cap` that holds no merged word
// is zero, so out[0, min(*cursor, cap)) is always defined (0 = no word).

// WPR windows per thread between two block barriers (1 or 2): up to 256 * WPR
// new keys per round, so the table holds at most kMaxDistinct + 256 * WPR <
// SLOTS keys and a flush comes before any count passes kMaxRun.
//
// Groups says which reads a block merges and into which table: count() groups,
// get(g, r0, nr, t) gives group g's first position r0 in the read order, its
// nr <= kReadsPerBlock reads and the table its flushes add to (every table of
// one pass has the same key and count layout).
template <int SLOTS, bool EMIT, bool KHI, int WPR, class Groups>
__device__ __forceinline__ void count_agg(ReadsView rv, const Groups& groups, int k, int canonical, TableView t,
                                          uint32_t* __restrict__ flags, uint64_t* __restrict__ out, uint64_t cap,
                                          uint64_t chunk, unsigned long long* __restrict__ cursor,
                                          int table_ok = 1) {
  extern __shared__ uint64_t s_tab[];  // SLOTS words, 0 = empty
  __shared__ uint32_t s_off[kReadsPerBlock + 1];
  __shared__ uint64_t s_wb[kReadsPerBlock];
  __shared__ uint32_t s_new, s_distinct;
  __shared__ uint32_t s_nd[4];
  __shared__ unsigned long long s_cur, s_end, s_z0, s_z1;  // EMIT: reserved chunk, range to zero
  constexpr uint32_t kMaxDistinct = SLOTS / 4 * 3;  // LDS load <= 3/4
  constexpr uint32_t kMaxRun = 8064;                // occurrences per merge <= cmax (k <= 25)
  constexpr uint32_t kQuarter = SLOTS / 4;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint64_t ngroups = groups.count();
  const uint64_t g0 = EMIT ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
  const uint64_t gstride = EMIT ? gridDim.x : ngroups;
  for (int i = tid; i < SLOTS; i += blockDim.x) s_tab[i] = 0;
  if (tid == 0) {
    s_new = 0;
    s_distinct = 0;
    s_cur = s_end = 0;
  }
  const uint64_t one = 1ull << t.cshift;
  uint64_t* quarter = s_tab + (uint32_t)wv * kQuarter;
  uint32_t mynew = 0;
  for (uint64_t grp = g0; grp < ngroups; grp += gstride) {
    uint64_t r0 = 0;
    int nr = 0;
    groups.get(grp, r0, nr, t);
    __syncthreads();  // the previous group's reads of s_off / s_wb are done
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
      if (tid == 0) s_off[0] = 0;
    }
    __syncthreads();
    // Four threads per read, each rolling over one quarter of its windows:
    // per round a thread shifts one base into its forward and reverse
    // k-mers (a code word is loaded once per 32 bases) and tracks the run of
    // ACGT bases ending there (a window is valid when the run is >= k).
    // The two code / mask words holding the thread's first window are loaded
    // here; later words when the rolling position reaches them.  The current
    // words are kept shifted so that the next base sits in their top bits.
    const int rj = tid >> 2, rp = tid & 3;
    uint32_t rs = 0, rcnt = 0;
    uint64_t rwb = 0, c0 = 0, c1 = 0;
    uint32_t m0 = 0, m1 = 0;
    if (rj < nr) {
      const uint32_t nw = s_off[rj + 1] - s_off[rj];
      const uint32_t q = (nw + 3) >> 2;
      rs = min(nw, (uint32_t)rp * q);
      rcnt = min(nw, rs + q) - rs;
      rwb = s_wb[rj];
      if (rcnt) {
        // words rs/32 and rs/32 + 1: the second is at most the word after
        // the read's last one (the next read's first word or the padding)
        c0 = rv.codes[rwb + (rs >> 5)];
        m0 = rv.nmask[rwb + (rs >> 5)];
        c1 = rv.codes[rwb + (rs >> 5) + 1];
        m1 = rv.nmask[rwb + (rs >> 5) + 1];
      }
    }
    const uint32_t wmax = wave_max_u32(rcnt);
    __syncthreads();  // every thread has read s_off
    if (lane == 0) s_nd[wv] = wmax;
    __syncthreads();
    const uint32_t rounds = max(max(s_nd[0], s_nd[1]), max(s_nd[2], s_nd[3]));
    const uint64_t kmask = (1ull << (2 * k)) - 1;
    const uint32_t khi = (uint32_t)(t.kmo >> 32);
    // the thread's first window, and its rolling state after it
    uint64_t fwd = 0, rcm = 0, cw = 0;
    uint32_t cm = 0, run = 0, left = 0, wnext = 0;
    if (rcnt) {
      const uint32_t o = rs & 31u;
      const uint64_t x = o ? ((c0 << (2 * o)) | (c1 >> (64 - 2 * o))) : c0;
      const uint32_t mw = o ? ((m0 << o) | (m1 >> (32 - o))) : m0;
      fwd = x >> (64 - 2 * k);
      const uint32_t nm = mw >> (32 - k);
      rcm = revcomp(fwd, k);
      run = nm ? (uint32_t)__builtin_ctz(nm) : (uint32_t)k;
      // the next base, rs + k, is in word rs/32 or rs/32 + 1 (k <= 32)
      const uint32_t pos = rs + k, po = pos & 31u;
      const bool second = (pos >> 5) != (rs >> 5);
      cw = (second ? c1 : c0) << (2 * po);
      cm = (second ? m1 : m0) << po;
      left = 32 - po;
      wnext = (pos >> 5) + 1;
    }
    uint32_t since = 0;  // windows merged since the last flush (block-uniform)
    static_assert(SLOTS / 4 * 3 + 256 * WPR < SLOTS, "the LDS table could fill");
    for (uint32_t r = 0; r < rounds; r += WPR) {
      // 1. WPR windows per thread into the LDS table
      uint32_t nnew = 0;
#pragma unroll
      for (int u = 0; u < WPR; u++) {
      const uint32_t ru = r + u;
      bool isnew = false;
      if (ru < rcnt) {
        if (ru > 0) {
          if (left == 0) {
            cw = rv.codes[rwb + wnext];
            cm = rv.nmask[rwb + wnext];
            left = 32;
            wnext++;
          }
          const uint32_t b = (uint32_t)(cw >> 62);
          const uint32_t isn = cm >> 31;
          cw <<= 2;
          cm <<= 1;
          left--;
          fwd = ((fwd << 2) | b) & kmask;
          rcm = (rcm >> 2) | ((uint64_t)(b ^ 3u) << (2 * k - 2));
          run = isn ? 0u : run + 1;
        }
        if (run >= (uint32_t)k) {
          const uint64_t tagged = ((canonical && rcm < fwd) ? rcm : fwd) | t.occ;
          const uint32_t lo = (uint32_t)tagged, hi = (uint32_t)(tagged >> 32);
          const uint32_t hb = __umulhi((lo ^ (lo >> 15) ^ (hi << 7)) * 0x9E3779B1u, SLOTS / 8);
          uint64_t old;
          // the table never holds more than kMaxDistinct + 256 * WPR < SLOTS keys
          isnew = (KHI ? (EMIT
                              ? lds_bucket_add_khi_half(s_tab, hb, SLOTS / 8, tagged | one, khi, t.cshift, 4 * SLOTS, old)
                              : lds_bucket_add_khi(s_tab, hb, SLOTS / 8, tagged | one, khi, t.cshift, 4 * SLOTS, old))
                       : lds_bucket_add(s_tab, hb, SLOTS / 8, tagged | one, t.kmo, t.cshift, 4 * SLOTS, old)) == 2;
        }
      }
      nnew += isnew ? 1u : 0u;
      }
      const uint32_t wnew = WPR == 1 ? (uint32_t)__popcll(__ballot(nnew != 0)) : wave_sum_u32(nnew);
      if (lane == 0 && wnew) atomicAdd(&s_distinct, wnew);
      since += 256 * WPR;
      __syncthreads();
      // Flush when another round could take the table past 3/4 load or a
      // count past kMaxRun occurrences, and after the last round.
      const uint32_t distinct = s_distinct;
      if (r + WPR >= rounds || distinct + 256 * WPR > kMaxDistinct || since + 256 * WPR > kMaxRun) {
        __syncthreads();  // every thread has read s_distinct
        if (tid == 0) s_distinct = 0;
        // 2. each wavefront compacts its quarter in place (a wavefront's LDS
        //    reads of a batch precede its writes, and writes land at or below
        //    the batch being read)...
        uint32_t nd = 0;
        for (uint32_t i0 = 0; i0 < kQuarter; i0 += 64) {
          const uint64_t v = quarter[i0 + lane];
          const uint64_t m = __ballot(v != 0);
          if (v) quarter[nd + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = v;
          nd += (uint32_t)__popcll(m);
        }
        wave_sync();
        if (EMIT) {
          // 3a. ...and appends them to the block's chunk of the stream (a
          //     new chunk when they do not fit; the old one's rest zeroed)
          if (lane == 0) s_nd[wv] = nd;
          __syncthreads();
          const uint32_t sum = s_nd[0] + s_nd[1] + s_nd[2] + s_nd[3];
          if (tid == 0) {
            s_z0 = s_z1 = 0;
            if (s_cur + sum > s_end) {
              s_z0 = s_cur;
              s_z1 = s_end < cap ? s_end : cap;
              s_cur = atomicAdd(cursor, (unsigned long long)chunk);
              s_end = s_cur + chunk;
            }
          }
          __syncthreads();
          const uint64_t cur = s_cur, lim = s_end < cap ? s_end : cap;
          for (uint64_t i = s_z0 + tid; i < s_z1; i += blockDim.x) out[i] = 0;
          if (cur + sum <= lim) {
            uint64_t at = cur;
            for (int i = 0; i < wv; i++) at += s_nd[i];
            for (uint32_t e0 = lane; e0 < nd; e0 += 64) out[at + e0] = quarter[e0];
            nd = 0;
          }
          __syncthreads();  // every thread has read s_cur
          if (tid == 0 && cur + sum <= lim) s_cur = cur + sum;
        }
        // 3. ...and adds its distinct keys to the global table, 64 at a time
        //    (EMIT: only a flush past cap, one lane per key; while the table
        //    is not yet cleared, table_ok = 0, such a flush is dropped and
        //    flags[6] set: the caller clears the table and merges again)
        if (EMIT && !table_ok && nd) {
          if (lane == 0) atomicOr(&flags[6], 1u);
          nd = 0;
        }
        for (uint32_t e0 = 0; e0 < nd; e0 += 64) {
          const bool want = e0 + lane < nd;
          const uint64_t v = want ? quarter[e0 + lane] : 0;
          if (EMIT) {
            if (want) {
              const uint32_t r = table_add(t, v & (t.occ - 1), v >> t.cshift, &flags[0]);
              if (r == 2u) atomicOr(&flags[0], 1u);
              else mynew += r;
            }
          } else {
            bool lost = false;
            mynew += coop4_add(t, v & (t.occ - 1), want, lane, &flags[0], lost, want ? v >> t.cshift : 1);
            if (lost) atomicOr(&flags[0], 1u);
          }
        }
        wave_sync();
        for (uint32_t i = lane; i < kQuarter; i += 64) quarter[i] = 0;
        __syncthreads();
        since = 0;
      }
    }
  }
  if (EMIT) {
    __syncthreads();
    const uint64_t lim = s_end < cap ? s_end : cap;
    for (uint64_t i = s_cur + tid; i < lim; i += blockDim.x) out[i] = 0;
  }
  mynew = wave_sum_u32(mynew);
  if (lane == 0 && mynew) atomicAdd(&s_new, mynew);
  __syncthreads();
  if (tid == 0 && s_new)
    atomicAdd(reinterpret_cast<unsigned long long*>(&flags[2]), (unsigned long long)s_new);
}

// Groups of kReadsPerBlock consecutive positions of [r_begin, r_end), one table.
struct RangeGroups {
  uint64_t r_begin, r_end;
  __device__ __forceinline__ uint64_t count() const { return (r_end - r_begin + kReadsPerBlock - 1) / kReadsPerBlock; }
  __device__ __forceinline__ void get(uint64_t g, uint64_t& r0, int& nr, TableView&) const {
    r0 = r_begin + g * kReadsPerBlock;
    nr = (int)std::min<uint64_t>(kReadsPerBlock, r_end - r0);
  }
};

template <int SLOTS, bool EMIT, bool KHI = true, int WPR = 1>
__global__ __launch_bounds__(256, EMIT ? (SLOTS <= 3072 ? 6 : 5) : 1) void k_count_agg(
    ReadsView rv, uint64_t r_begin, uint64_t r_end, int k, int canonical, TableView t, uint32_t* __restrict__ flags,
    uint64_t* __restrict__ out, uint64_t cap, uint64_t chunk, unsigned long long* __restrict__ cursor,
    int table_ok) {
  count_agg<SLOTS, EMIT, KHI, WPR>(rv, RangeGroups{r_begin, r_end}, k, canonical, t, flags, out, cap, chunk, cursor,
                                   table_ok);
}

}  // namespace tk
