// tk_batch.hip -- many small independent components in one pass.
//
// Trinity's phase 2 runs one `Trinity --trinity_complete` per Chrysalis
// component through ParaFly (Trinity:2510-2529,3652;
// write_partitioned_trinity_cmds.pl:79-93); each counts its reads
// (inchworm --reads, Trinity:2669-2672) and assigns them to its contigs with
// ReadsToTranscripts -p 50 (Trinity:1044-1046,2241).  Here every component of
// a batch gets its own count table, side table and component table carved
// out of three shared allocations, and one launch per stage covers all
// components:
//   clear -> count -> fold carries -> coverage stats -> contig table -> vote
// The sequence is captured once into a hipGraph (no memset nodes: clearing is
// a kernel) and replayed.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tk_count_agg.h"
#include "tk_internal.h"
#include "tk_perread.h"

namespace tk {

// Per-component layout, read by every lane from the component of its read.
struct CompLayout {
  uint64_t boff;     // first count-table bucket
  uint64_t nb;       // count-table buckets (power of two, one region)
  uint64_t ovf_off;  // first side-table pair
  uint64_t omask;    // side-table pairs - 1
  uint64_t ct_boff;  // first component-table bucket
  uint64_t ct_mask;  // component-table buckets - 1
  uint32_t rbits;
  uint32_t flags;    // kCompCtInLds: the component table is built in LDS (k_batch_ctable_lds)
};

constexpr uint32_t kCompCtInLds = 1;
// Component tables of at most kCtLdsBuckets 64-B buckets (64 KB) are built
// by one workgroup in LDS and written out whole (32 KB when the 64 KB of
// dynamic LDS cannot be granted).
constexpr uint64_t kCtLdsBuckets = 1024;

__device__ __forceinline__ TableView comp_view(const TableView& base, const CompLayout& L) {
  TableView v = base;
  v.boff = L.boff;
  v.nb = L.nb;
  v.rbits = L.rbits;
  v.rmask = L.nb - 1;
  v.ovf = base.ovf + 2 * L.ovf_off;
  v.omask = L.omask;
  v.max_probe = (uint32_t)(L.nb < 4096 ? L.nb : 4096);
  return v;
}

__global__ void k_fill_zero(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_fill_zero32(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

constexpr int kBatchReads = 64;

// Count: one thread per window of 64 consecutive reads (which may belong to
// different components; each lane probes its own component's table).
__global__ __launch_bounds__(256) void k_batch_count(ReadsView rv, const uint32_t* __restrict__ read_comp,
                                                     const CompLayout* __restrict__ lay, int k,
                                                     int canonical, TableView base,
                                                     uint32_t* __restrict__ flags) {
  __shared__ uint32_t s_off[kBatchReads + 1];
  __shared__ uint64_t s_wb[kBatchReads];
  __shared__ uint32_t s_comp[kBatchReads];
  const uint64_t r0 = (uint64_t)blockIdx.x * kBatchReads;
  if (r0 >= rv.n_reads) return;
  const int nr = (int)(rv.n_reads - r0 < kBatchReads ? rv.n_reads - r0 : kBatchReads);
  const int tid = threadIdx.x;
  if (tid < kBatchReads) {
    uint32_t nw = 0;
    if (tid < nr) {
      const uint32_t L = rv.len[r0 + tid];
      nw = L >= (uint32_t)k ? L - k + 1 : 0;
      s_wb[tid] = rv.word_off[r0 + tid];
      s_comp[tid] = read_comp[r0 + tid];
    }
    uint32_t x = nw;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (tid >= o) x += y;
    }
    s_off[tid + 1] = x;
    if (tid == 0) s_off[0] = 0;
  }
  __syncthreads();
  const uint32_t total = s_off[nr];
  for (uint32_t w0 = 0; w0 < total; w0 += blockDim.x) {
    const uint32_t w = w0 + tid;
    uint64_t fwd = 0;
    bool valid = false;
    int j = 0;
    if (w < total) {
      int lo = 0, hi = nr;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= w) lo = mid; else hi = mid;
      }
      j = lo;
      window_at(rv, s_wb[j], w - s_off[j], k, fwd, valid);
    }
    const TableView t = comp_view(base, lay[s_comp[j]]);
    const uint64_t key = valid ? canon_key(fwd, k, canonical) : 0;
    bool lost = false;
    coop_add64(t, key, valid, tid & 63, &flags[0], lost);
    if (lost) atomicOr(&flags[0], 1u);
  }
}

// The count with the block merge of k_count_agg (k <= 25): each block takes
// a group of at most 64 consecutive reads of ONE component, merges their
// windows in its LDS table and adds each distinct key once, with its
// occurrences, to that component's table.  A component of many reads at high
// coverage repeats most k-mers within a group, so the table sees a fraction
// of the per-window atomics of k_batch_count.
struct BatchGroup {
  uint64_t r0;    // first read
  uint32_t nr;    // reads (<= kReadsPerBlock)
  uint32_t comp;  // component
};

struct BatchGroups {
  const BatchGroup* g;
  uint64_t n;
  const CompLayout* lay;
  TableView base;
  __device__ __forceinline__ uint64_t count() const { return n; }
  __device__ __forceinline__ void get(uint64_t i, uint64_t& r0, int& nr, TableView& t) const {
    const BatchGroup x = g[i];
    r0 = x.r0;
    nr = (int)x.nr;
    t = comp_view(base, lay[x.comp]);
  }
};

constexpr int kBatchAggSlots = 4096;

template <bool KHI, int WPR>
__global__ __launch_bounds__(256, 1) void k_batch_count_agg(ReadsView rv, const BatchGroup* __restrict__ groups,
                                                            uint64_t n_groups, const CompLayout* __restrict__ lay,
                                                            int k, int canonical, TableView base,
                                                            uint32_t* __restrict__ flags) {
  count_agg<kBatchAggSlots, false, KHI, WPR>(rv, BatchGroups{groups, n_groups, lay, base}, k, canonical, base, flags,
                                             nullptr, 0, 0, nullptr);
}

// Locality key of each read within its component: the component in the top
// bits, then the read's minimizer key (MinKey), so that sorting by it keeps
// every component's reads together and puts reads of one locus side by side
// (the merge then sees their shared windows within one group).
__global__ __launch_bounds__(256) void k_batch_order_keys(ReadsView rv, const uint32_t* __restrict__ read_comp,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
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
    keys[r] = ((uint64_t)read_comp[r] << kOrderKeyBits) | mk.key(L);
    idx[r] = (uint32_t)r;
  }
}

// Whether the batched count merges per block (TK_BATCH_AGG=0: one atomic per
// window, k_batch_count), read when the batch is created.  Needs k <= 25 (a
// merged count fits the slot field); without the read order (more than 2^24
// components or 2^31 reads) the batch counts per window.
static bool batch_agg(int k) {
  const char* e = getenv("TK_BATCH_AGG");
  return !(e && e[0] == '0') && 63 - 2 * k >= 13;
}

// Finalise every component's side table (see k_ovf_fold in tk_lib.hip).
__global__ void k_batch_fold(TableView base, const CompLayout* __restrict__ lay, uint64_t n_comp,
                             uint64_t n_pairs, uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t tagged = base.ovf[2 * i];
    if (!tagged) continue;
    uint64_t lo = 0, hi = n_comp;  // component owning pair i
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (lay[mid].ovf_off <= i) lo = mid; else hi = mid;
    }
    const TableView t = comp_view(base, lay[lo]);
    uint64_t b = bucket0(t, tagged & (t.occ - 1));
    uint64_t* s = nullptr;
    for (uint32_t p = 0; p < t.max_probe && !s; p++) {
      uint64_t* line = t.slots + (b << 3);
      for (int j = 0; j < 8; j++) {
        if ((line[j] & t.kmo) == tagged) { s = &line[j]; break; }
        if (line[j] == 0) { p = t.max_probe; break; }
      }
      b = bucket_next(t, b);
    }
    if (!s) { atomicOr(&flags[0], 4u); continue; }
    const uint64_t total = base.ovf[2 * i + 1] + (*s >> t.cshift);
    base.ovf[2 * i + 1] = total;
    *s = total ? ((*s & t.kmo) | (t.cmax << t.cshift)) : (*s & t.kmo);
  }
}

// kShort: every read has at most 128 windows (the long-read path is not compiled).
template <bool kShort>
__global__ __launch_bounds__(256) void k_batch_stats(ReadsView rv, const uint32_t* __restrict__ read_comp,
                                                     const CompLayout* __restrict__ lay, int k,
                                                     int canonical, TableView base, StatsParams sp,
                                                     tk_read_stat* __restrict__ out, uint32_t slab) {
  extern __shared__ uint32_t s_dyn[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t* vals = s_dyn + (uint64_t)wv * slab;
  uint64_t* tmp = reinterpret_cast<uint64_t*>(s_dyn + wpb * slab) + wv * kLookupScratch;
  for (uint64_t r = (uint64_t)blockIdx.x * wpb + wv; r < rv.n_reads; r += (uint64_t)gridDim.x * wpb) {
    const TableView t = comp_view(base, lay[read_comp[r]]);
    stats_read<kShort>(rv, r, k, canonical, t, sp, vals, tmp, lane, out + r);
  }
}

// Contig k-mers of each component -> index of the contig within its component
// (kPacked: pctable_add's one-word slots, else ctable_add's two-word slots).
template <bool kPacked>
__global__ void k_batch_ctable(ReadsView cv, const uint32_t* __restrict__ contig_comp,
                               const uint32_t* __restrict__ contig_local,
                               const CompLayout* __restrict__ lay, int k, int canonical,
                               uint64_t* __restrict__ ct, uint32_t* __restrict__ flags) {
  const uint64_t occ = 1ull << (2 * k);
  const int lane = threadIdx.x & 63;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t dummy = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); r < cv.n_reads;
       r += (uint64_t)gridDim.x * wpb) {
    const CompLayout L = lay[contig_comp[r]];
    if (L.flags & kCompCtInLds) continue;  // built by k_batch_ctable_lds
    const uint32_t local = contig_local[r];
    const uint32_t len = cv.len[r];
    const uint32_t nwin = len >= (uint32_t)k ? len - k + 1 : 0;
    const uint64_t wb = cv.word_off[r];
    for (uint32_t p = lane; p < nwin; p += 64) {
      uint64_t fwd;
      bool valid;
      window_at(cv, wb, p, k, fwd, valid);
      if (!valid) continue;
      if (kPacked)
        pctable_add(ct, L.ct_boff, L.ct_mask, k, canon_key(fwd, k, canonical), local, &flags[0]);
      else
        ctable_add(ct, L.ct_boff, L.ct_mask, occ, canon_key(fwd, k, canonical), local, dummy, &flags[0]);
    }
  }
}

// The component tables of small components (kCompCtInLds), one workgroup
// each: the table is built in LDS with the same insert protocol and probe
// sequence as in global memory (every contig k-mer's word claimed with a
// CAS, the smallest contig index kept), then written out whole, so each
// contig window costs an LDS atomic instead of a memory-side one.  Waves
// take the component's contigs in turn, lanes their windows.
template <bool kPacked>
__global__ __launch_bounds__(256) void k_batch_ctable_lds(ReadsView cv, const uint32_t* __restrict__ comps,
                                                          const uint64_t* __restrict__ contig_off,
                                                          const CompLayout* __restrict__ lay, int k, int canonical,
                                                          uint64_t* __restrict__ ct, uint32_t* __restrict__ flags) {
  extern __shared__ uint64_t s_ct[];
  const uint32_t c = comps[blockIdx.x];
  const CompLayout L = lay[c];
  const uint64_t words = (L.ct_mask + 1) << 3;
  for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) s_ct[i] = 0;
  __syncthreads();
  const uint64_t occ = 1ull << (2 * k);
  const int lane = threadIdx.x & 63;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t dummy = 0;
  const uint64_t c0 = contig_off[c], c1 = contig_off[c + 1];
  for (uint64_t r = c0 + (threadIdx.x >> 6); r < c1; r += wpb) {
    const uint32_t local = (uint32_t)(r - c0);
    const uint32_t len = cv.len[r];
    const uint32_t nwin = len >= (uint32_t)k ? len - k + 1 : 0;
    const uint64_t wb = cv.word_off[r];
    for (uint32_t p = lane; p < nwin; p += 64) {
      uint64_t fwd;
      bool valid;
      window_at(cv, wb, p, k, fwd, valid);
      if (!valid) continue;
      if (kPacked)
        pctable_add(s_ct, 0, L.ct_mask, k, canon_key(fwd, k, canonical), local, &flags[0]);
      else
        ctable_add(s_ct, 0, L.ct_mask, occ, canon_key(fwd, k, canonical), local, dummy, &flags[0]);
    }
  }
  __syncthreads();
  uint64_t* g = ct + (L.ct_boff << 3);
  for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) g[i] = s_ct[i];
}

template <bool kPacked>
__global__ __launch_bounds__(256) void k_batch_vote(ReadsView rv, const uint32_t* __restrict__ read_comp,
                                                    const CompLayout* __restrict__ lay, int k,
                                                    int canonical, const uint64_t* __restrict__ ct,
                                                    tk_r2t_vote* __restrict__ out, uint32_t slab) {
  extern __shared__ uint32_t s_dyn[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t wpb = blockDim.x >> 6;
  uint32_t* hits = s_dyn + (uint64_t)wv * slab;
  uint32_t* tmp = s_dyn + wpb * slab + wv * 64;
  for (uint64_t r = (uint64_t)blockIdx.x * wpb + wv; r < rv.n_reads; r += (uint64_t)gridDim.x * wpb) {
    const CompLayout L = lay[read_comp[r]];
    vote_read<kPacked>(rv, r, k, canonical, ct, L.ct_boff, L.ct_mask, hits, tmp, lane, out + r);
  }
}

static inline uint64_t p2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace tk

using namespace tk;

struct tk_batch {
  int device = 0;
  int k = 25;
  int canonical = 1;
  uint64_t n_comp = 0;
  const tk_reads* reads = nullptr;
  const tk_reads* contigs = nullptr;
  double window_load = 0.7;
  int stats_waves = 4;
  bool use_graph = true;
  bool ct_packed = false;  // one-word component-table slots (k <= 25, < 2^(63-2k) contigs per component)
  uint32_t max_nwin = 0;
  // device
  CompLayout* d_lay = nullptr;
  uint32_t* d_read_comp = nullptr;
  uint32_t* d_contig_comp = nullptr;
  uint32_t* d_contig_local = nullptr;
  BatchGroup* d_groups = nullptr;  // block-merge groups of the count
  uint64_t n_groups = 0;
  uint32_t* d_order = nullptr;     // reads of each component in locality order (merged count)
  uint64_t* d_contig_off = nullptr;  // contigs of each component (n_comp + 1)
  uint32_t* d_ct_lds = nullptr;      // components whose table is built in LDS
  uint64_t n_ct_lds = 0, ct_lds_bytes = 0;
  uint64_t* d_slots = nullptr;
  uint64_t* d_ovf = nullptr;
  uint64_t* d_ct = nullptr;
  uint64_t n_slots_words = 0, n_ovf_pairs = 0, n_ct_words = 0;
  uint64_t bytes = 0;
  // captured launch sequence
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  tk_read_stat* g_stats = nullptr;
  tk_r2t_vote* g_votes = nullptr;
  tk_stats_opts g_opts{};
  // host copies for re-layout
  std::vector<uint64_t> read_off, contig_off;
  std::vector<uint64_t> comp_windows, comp_cwindows;
};

static void batch_release_device(tk_batch* b) {
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  dev_free(b->d_slots, b->n_slots_words * 8);
  dev_free(b->d_ovf, b->n_ovf_pairs * 16);
  dev_free(b->d_ct, b->n_ct_words * 8);
  dev_free(b->d_lay, b->n_comp * sizeof(CompLayout));
  dev_free(b->d_ct_lds, std::max<uint64_t>(b->n_ct_lds, 1) * 4);
  b->d_slots = b->d_ovf = b->d_ct = nullptr;
  b->d_lay = nullptr;
  b->d_ct_lds = nullptr;
  b->n_ct_lds = b->ct_lds_bytes = 0;
}

// Whether k_batch_ctable_lds may take 64 KB of dynamic LDS (the attribute is
// set once per process; a function-local static is initialised once even
// when several threads get here together).
static bool ct_lds_64k() {
  static const bool ok = [] {
    const bool r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_ctable_lds<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 65536) == hipSuccess &&
                   hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_ctable_lds<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 65536) == hipSuccess;
    (void)hipGetLastError();
    return r;
  }();
  return ok;
}

// Lay the per-component tables out for the current window_load.
static int batch_layout(tk_ctx* c, tk_batch* b) {
  batch_release_device(b);
  std::vector<CompLayout> lay(b->n_comp);
  uint64_t boff = 0, ooff = 0, coff = 0;
  const uint32_t cbits = 63 - 2 * b->k;
  for (uint64_t i = 0; i < b->n_comp; i++) {
    CompLayout& L = lay[i];
    const uint64_t w = b->comp_windows[i];
    const uint64_t slots = (uint64_t)ceil((double)w / b->window_load) + 8;
    L.nb = p2((slots + 7) / 8);
    L.rbits = 0;
    while ((1ull << L.rbits) < L.nb) L.rbits++;
    L.boff = boff;
    boff += L.nb;
    const uint64_t oc = p2(2 * (w >> cbits) + 16);
    L.ovf_off = ooff;
    L.omask = oc - 1;
    ooff += oc;
    // buckets of eight packed or four two-word slots, at most half full
    const uint64_t cb = p2(std::max<uint64_t>(b->ct_packed ? (2 * b->comp_cwindows[i] + 7) / 8
                                                            : (2 * b->comp_cwindows[i] + 3) / 4, 4));
    L.ct_boff = coff;
    L.ct_mask = cb - 1;
    coff += cb;
    L.flags = cb <= kCtLdsBuckets ? kCompCtInLds : 0u;
  }
  b->n_slots_words = boff * 8;
  b->n_ovf_pairs = ooff;
  b->n_ct_words = coff * 8;
  // components whose table is built in LDS (TK_BATCH_CT_LDS=0: none); one
  // launch takes them all, with the LDS of the largest
  std::vector<uint32_t> small;
  {
    const char* ev = getenv("TK_BATCH_CT_LDS");
    const bool on = !(ev && ev[0] == '0');
    const uint64_t max_buckets = ct_lds_64k() ? kCtLdsBuckets : kCtLdsBuckets / 2;
    for (uint64_t i = 0; i < b->n_comp; i++) {
      if (!on || lay[i].ct_mask + 1 > max_buckets) lay[i].flags &= ~kCompCtInLds;
      if ((lay[i].flags & kCompCtInLds) && b->contig_off[i + 1] > b->contig_off[i]) {
        small.push_back((uint32_t)i);
        b->ct_lds_bytes = std::max<uint64_t>(b->ct_lds_bytes, (lay[i].ct_mask + 1) * 64);
      }
    }
  }
  b->n_ct_lds = small.size();
  int rc = dev_alloc((void**)&b->d_slots, b->n_slots_words * 8);
  if (!rc) rc = dev_alloc((void**)&b->d_ovf, b->n_ovf_pairs * 16);
  if (!rc) rc = dev_alloc((void**)&b->d_ct, b->n_ct_words * 8);
  if (!rc) rc = dev_alloc((void**)&b->d_lay, b->n_comp * sizeof(CompLayout));
  if (!rc) rc = dev_alloc((void**)&b->d_ct_lds, std::max<uint64_t>(b->n_ct_lds, 1) * 4);
  if (rc) return rc;
  TK_HIP(hipMemcpyAsync(b->d_lay, lay.data(), b->n_comp * sizeof(CompLayout), hipMemcpyHostToDevice,
                        c->stream));
  if (b->n_ct_lds)
    TK_HIP(hipMemcpyAsync(b->d_ct_lds, small.data(), b->n_ct_lds * 4, hipMemcpyHostToDevice, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  b->bytes = b->n_slots_words * 8 + b->n_ovf_pairs * 16 + b->n_ct_words * 8;
  return TK_OK;
}

static TableView batch_base_view(const tk_batch* b) {
  TableView v;
  memset(&v, 0, sizeof v);
  v.slots = b->d_slots;
  v.occ = 1ull << (2 * b->k);
  v.kmo = (v.occ << 1) - 1;
  v.cshift = 2 * b->k + 1;
  v.cbits = 63 - 2 * b->k;
  v.cmax = (1ull << v.cbits) - 1;
  v.ovf = b->d_ovf;
  return v;
}

static ReadsView rview(const tk_reads* r) {
  ReadsView v;
  v.codes = r->d_codes;
  v.nmask = r->d_nmask;
  v.word_off = r->d_word_off;
  v.len = r->d_len;
  v.n_reads = r->n_reads;
  v.order = nullptr;
  return v;
}

// Enqueue the whole batch on c->stream (capturable: kernels only).
static int batch_enqueue(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* d_stats,
                         tk_r2t_vote* d_votes) {
  const int cap = std::max(c->num_cu, 1) * 8;
  auto grid = [&](uint64_t n, uint64_t per) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + per - 1) / per, cap)); };
  hipLaunchKernelGGL(k_fill_zero32, dim3(1), dim3(64), 0, c->stream, c->d_flags, 64u);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid(b->n_slots_words, 256)), dim3(256), 0, c->stream, b->d_slots, b->n_slots_words);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid(2 * b->n_ovf_pairs, 256)), dim3(256), 0, c->stream, b->d_ovf, 2 * b->n_ovf_pairs);
  const TableView base = batch_base_view(b);
  const ReadsView rv = rview(b->reads);
  if (rv.n_reads) {
    const unsigned ag = (unsigned)((b->n_groups + 7) / 8 * 8);  // xcd_block maps grids of 8k blocks
    ReadsView ro = rv;
    ro.order = b->d_order;
    if (b->d_order && b->k >= 16)
      hipLaunchKernelGGL((k_batch_count_agg<true, 2>), dim3(ag), dim3(256), kBatchAggSlots * 8, c->stream, ro,
                         b->d_groups, b->n_groups, b->d_lay, b->k, b->canonical, base, c->d_flags);
    else if (b->d_order)
      hipLaunchKernelGGL((k_batch_count_agg<false, 1>), dim3(ag), dim3(256), kBatchAggSlots * 8, c->stream, ro,
                         b->d_groups, b->n_groups, b->d_lay, b->k, b->canonical, base, c->d_flags);
    else
      hipLaunchKernelGGL(k_batch_count, dim3((rv.n_reads + kBatchReads - 1) / kBatchReads), dim3(256), 0,
                         c->stream, rv, b->d_read_comp, b->d_lay, b->k, b->canonical, base, c->d_flags);
    hipLaunchKernelGGL(k_batch_fold, dim3(grid(b->n_ovf_pairs, 256)), dim3(256), 0, c->stream, base, b->d_lay,
                       b->n_comp, b->n_ovf_pairs, c->d_flags);
  }
  uint32_t slab = std::max<uint32_t>(64, (b->max_nwin + 63) / 64 * 64);
  int wpb = b->stats_waves;
  while (wpb > 1 && ((uint64_t)slab * 4 + 8 * kLookupScratch) * wpb > 65536) wpb >>= 1;
  if (d_stats && rv.n_reads) {
    StatsParams sp{o->min_count, o->missing_value, o->nwin_policy, o->median_policy};
    if (b->max_nwin <= 128)
      hipLaunchKernelGGL(k_batch_stats<true>, dim3(grid(rv.n_reads, wpb)), dim3(64 * wpb),
                         (size_t)slab * 4 * wpb + (size_t)wpb * 8 * kLookupScratch, c->stream, rv, b->d_read_comp,
                         b->d_lay, b->k, b->canonical, base, sp, d_stats, slab);
    else
      hipLaunchKernelGGL(k_batch_stats<false>, dim3(grid(rv.n_reads, wpb)), dim3(64 * wpb),
                         (size_t)slab * 4 * wpb + (size_t)wpb * 8 * kLookupScratch, c->stream, rv, b->d_read_comp,
                         b->d_lay, b->k, b->canonical, base, sp, d_stats, slab);
  }
  if (d_votes && rv.n_reads) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid(b->n_ct_words, 256)), dim3(256), 0, c->stream, b->d_ct, b->n_ct_words);
    const ReadsView cv = rview(b->contigs);
    const size_t vlds = (size_t)slab * 4 * wpb + (size_t)wpb * 256;
    if (b->n_ct_lds && b->ct_packed)
      hipLaunchKernelGGL(k_batch_ctable_lds<true>, dim3((unsigned)b->n_ct_lds), dim3(256), b->ct_lds_bytes, c->stream,
                         cv, b->d_ct_lds, b->d_contig_off, b->d_lay, b->k, b->canonical, b->d_ct, c->d_flags);
    else if (b->n_ct_lds)
      hipLaunchKernelGGL(k_batch_ctable_lds<false>, dim3((unsigned)b->n_ct_lds), dim3(256), b->ct_lds_bytes,
                         c->stream, cv, b->d_ct_lds, b->d_contig_off, b->d_lay, b->k, b->canonical, b->d_ct,
                         c->d_flags);
    if (b->ct_packed) {
      if (cv.n_reads)
        hipLaunchKernelGGL(k_batch_ctable<true>, dim3(grid(cv.n_reads, 4)), dim3(256), 0, c->stream, cv,
                           b->d_contig_comp, b->d_contig_local, b->d_lay, b->k, b->canonical, b->d_ct, c->d_flags);
      hipLaunchKernelGGL(k_batch_vote<true>, dim3(grid(rv.n_reads, wpb)), dim3(64 * wpb), vlds, c->stream, rv,
                         b->d_read_comp, b->d_lay, b->k, b->canonical, b->d_ct, d_votes, slab);
    } else {
      if (cv.n_reads)
        hipLaunchKernelGGL(k_batch_ctable<false>, dim3(grid(cv.n_reads, 4)), dim3(256), 0, c->stream, cv,
                           b->d_contig_comp, b->d_contig_local, b->d_lay, b->k, b->canonical, b->d_ct, c->d_flags);
      hipLaunchKernelGGL(k_batch_vote<false>, dim3(grid(rv.n_reads, wpb)), dim3(64 * wpb), vlds, c->stream, rv,
                         b->d_read_comp, b->d_lay, b->k, b->canonical, b->d_ct, d_votes, slab);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "batch launch");
  return TK_OK;
}

static int batch_check(tk_ctx* c) {
  TK_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 16, hipMemcpyDeviceToHost, c->stream));
  TK_HIP(hipStreamSynchronize(c->stream));
  if (c->h_flags[0]) return fail(TK_ERR_FULL, "batch: table error flags=%u", c->h_flags[0]);
  return TK_OK;
}

static int batch_run_eager(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* d_stats,
                           tk_r2t_vote* d_votes) {
  hipEvent_t a;
  timing_begin(c, &a);
  int rc = batch_enqueue(c, b, o, d_stats, d_votes);
  timing_end(c, "batch", a);
  if (rc) return rc;
  return batch_check(c);
}

static int batch_run_graph(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* d_stats,
                           tk_r2t_vote* d_votes) {
  const bool same = b->exec && b->g_stats == d_stats && b->g_votes == d_votes &&
                    !memcmp(&b->g_opts, o, sizeof *o);
  if (!same) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    TK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(c, b, o, d_stats, d_votes);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    TK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_stats = d_stats;
    b->g_votes = d_votes;
    b->g_opts = *o;
  }
  hipEvent_t a;
  timing_begin(c, &a);
  TK_HIP(hipGraphLaunch(b->exec, c->stream));
  timing_end(c, "batch", a);
  return batch_check(c);
}

extern "C" {

int tk_batch_create(tk_ctx* c, int k, int canonical, const tk_reads* reads, const uint64_t* read_comp_off,
                    const tk_reads* contigs, const uint64_t* contig_comp_off, uint64_t n_comp,
                    const tk_batch_opts* opts, tk_batch** out) {
  if (!c || !reads || !contigs || !read_comp_off || !contig_comp_off || !out)
    return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (k < 1 || k > 31) return fail(TK_ERR_ARG, "k=%d outside [1,31]", k);
  if (!n_comp) return fail(TK_ERR_ARG, "empty batch");
  if (read_comp_off[0] != 0 || read_comp_off[n_comp] != reads->n_reads)
    return fail(TK_ERR_ARG, "read_comp_off must span reads [0, %llu)", (unsigned long long)reads->n_reads);
  if (contig_comp_off[0] != 0 || contig_comp_off[n_comp] != contigs->n_reads)
    return fail(TK_ERR_ARG, "contig_comp_off must span contigs [0, %llu)", (unsigned long long)contigs->n_reads);
  for (uint64_t i = 0; i < n_comp; i++)
    if (read_comp_off[i + 1] < read_comp_off[i] || contig_comp_off[i + 1] < contig_comp_off[i])
      return fail(TK_ERR_ARG, "component offsets not monotone at %llu", (unsigned long long)i);
  TK_HIP(hipSetDevice(c->device));
  tk_batch* b = new tk_batch();
  b->device = c->device;
  b->k = k;
  b->canonical = canonical ? 1 : 0;
  b->n_comp = n_comp;
  b->reads = reads;
  b->contigs = contigs;
  if (opts) {
    if (opts->window_load > 0) b->window_load = std::min(0.9, opts->window_load);
    if (opts->stats_waves > 0) b->stats_waves = std::min(4, opts->stats_waves);
    b->use_graph = opts->use_graph != 0;
  }
  b->read_off.assign(read_comp_off, read_comp_off + n_comp + 1);
  b->contig_off.assign(contig_comp_off, contig_comp_off + n_comp + 1);
  // per-component windows from the read lengths
  std::vector<uint32_t> rl(reads->n_reads), cl(contigs->n_reads);
  int rc = TK_OK;
  hipError_t e = hipSuccess;
  // copies on the context's stream (non-blocking: a null-stream copy would not
  // wait for work queued there, and a reused block may still be in use by it)
  if (reads->n_reads) e = hipMemcpyAsync(rl.data(), reads->d_len, reads->n_reads * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && contigs->n_reads)
    e = hipMemcpyAsync(cl.data(), contigs->d_len, contigs->n_reads * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { delete b; return hip_fail(e, "batch lengths"); }
  b->comp_windows.assign(n_comp, 0);
  b->comp_cwindows.assign(n_comp, 0);
  std::vector<uint32_t> rcomp(reads->n_reads), ccomp(contigs->n_reads), clocal(contigs->n_reads);
  for (uint64_t i = 0; i < n_comp; i++) {
    for (uint64_t r = read_comp_off[i]; r < read_comp_off[i + 1]; r++) {
      rcomp[r] = (uint32_t)i;
      b->comp_windows[i] += rl[r] >= (uint32_t)k ? rl[r] - k + 1 : 0;
      b->max_nwin = std::max<uint32_t>(b->max_nwin, rl[r] >= (uint32_t)k ? rl[r] - k + 1 : 0);
    }
    for (uint64_t r = contig_comp_off[i]; r < contig_comp_off[i + 1]; r++) {
      ccomp[r] = (uint32_t)i;
      clocal[r] = (uint32_t)(r - contig_comp_off[i]);
      b->comp_cwindows[i] += cl[r] >= (uint32_t)k ? cl[r] - k + 1 : 0;
    }
  }
  if ((uint64_t)(std::max<uint32_t>(64, (b->max_nwin + 63) / 64 * 64)) * 4 + 8 * kLookupScratch > 65536) {
    delete b;
    return fail(TK_ERR_ARG, "reads longer than %u bases are not supported", 16128 + k - 1);
  }
  // block-merge groups: each component's reads in runs of kReadsPerBlock
  std::vector<BatchGroup> groups;
  for (uint64_t i = 0; i < n_comp; i++)
    for (uint64_t r = read_comp_off[i]; r < read_comp_off[i + 1]; r += kReadsPerBlock)
      groups.push_back({r, (uint32_t)std::min<uint64_t>(kReadsPerBlock, read_comp_off[i + 1] - r), (uint32_t)i});
  b->n_groups = groups.size();
  rc = dev_alloc((void**)&b->d_groups, std::max<uint64_t>(b->n_groups, 1) * sizeof(BatchGroup));
  if (!rc && b->n_groups &&
      (e = hipMemcpyAsync(b->d_groups, groups.data(), b->n_groups * sizeof(BatchGroup), hipMemcpyHostToDevice,
                          c->stream)) != hipSuccess)
    rc = hip_fail(e, "upload count groups");
  if (!rc) rc = dev_alloc((void**)&b->d_contig_off, (n_comp + 1) * 8);
  if (!rc && (e = hipMemcpyAsync(b->d_contig_off, contig_comp_off, (n_comp + 1) * 8, hipMemcpyHostToDevice,
                                 c->stream)) != hipSuccess)
    rc = hip_fail(e, "upload contig offsets");
  if (!rc) rc = dev_alloc((void**)&b->d_read_comp, std::max<uint64_t>(reads->n_reads, 1) * 4);
  if (!rc) rc = dev_alloc((void**)&b->d_contig_comp, std::max<uint64_t>(contigs->n_reads, 1) * 4);
  if (!rc) rc = dev_alloc((void**)&b->d_contig_local, std::max<uint64_t>(contigs->n_reads, 1) * 4);
  if (!rc && reads->n_reads)
    if ((e = hipMemcpyAsync(b->d_read_comp, rcomp.data(), reads->n_reads * 4, hipMemcpyHostToDevice, c->stream)) !=
        hipSuccess)
      rc = hip_fail(e, "upload read components");
  if (!rc && contigs->n_reads) {
    if ((e = hipMemcpyAsync(b->d_contig_comp, ccomp.data(), contigs->n_reads * 4, hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b->d_contig_local, clocal.data(), contigs->n_reads * 4, hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
      rc = hip_fail(e, "upload contig components");
  }
  // the host vectors are freed on return
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess && !rc) rc = hip_fail(e, "batch uploads");
  // one-word component-table slots when every contig index fits beside the
  // key (TK_BATCH_CT_PACKED=0: always the two-word slots)
  {
    const char* ev = getenv("TK_BATCH_CT_PACKED");
    uint64_t max_contigs = 0;
    for (uint64_t i = 0; i < n_comp; i++)
      max_contigs = std::max<uint64_t>(max_contigs, contig_comp_off[i + 1] - contig_comp_off[i]);
    b->ct_packed = !(ev && ev[0] == '0') && k <= 25 && max_contigs <= (1ull << (63 - 2 * k));
  }
  // the merged count's read order: by component, then by locality key
  // (component ids fit above the 40 key bits; one sort call takes < 2^31 items)
  if (!rc && batch_agg(k) && b->n_groups && reads->n_reads <= 0x7FFFFFFFull && n_comp <= (1ull << 24)) {
    uint64_t* keys = nullptr;
    rc = dev_alloc((void**)&keys, reads->n_reads * 8);
    if (!rc) rc = dev_alloc((void**)&b->d_order, reads->n_reads * 4);
    if (!rc) {
      const uint64_t grid = std::min<uint64_t>((reads->n_reads + 255) / 256, (uint64_t)std::max(c->num_cu, 1) * 16);
      hipLaunchKernelGGL(k_batch_order_keys, dim3(grid), dim3(256), 0, c->stream, rview(reads), b->d_read_comp,
                         keys, b->d_order);
      if ((e = hipGetLastError()) != hipSuccess) rc = hip_fail(e, "k_batch_order_keys");
    }
    int cbits = 0;
    while ((1ull << cbits) < n_comp) cbits++;
    if (!rc) rc = sort_pairs_u64_u32(c, keys, b->d_order, reads->n_reads, kOrderKeyBits + cbits);
    if (keys) {
      (void)hipStreamSynchronize(c->stream);
      dev_free(keys, reads->n_reads * 8);
    }
  }
  if (!rc) rc = batch_layout(c, b);
  if (rc) { tk_batch_free(b); return rc; }
  *out = b;
  return TK_OK;
}

int tk_batch_run_d(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* d_stats, tk_r2t_vote* d_votes) {
  if (!c || !b || !o) return fail(TK_ERR_ARG, "NULL argument");
  if (o->nwin_policy != TK_NWIN_AS_MISSING && o->nwin_policy != TK_NWIN_SKIP)
    return fail(TK_ERR_ARG, "bad nwin_policy %d", o->nwin_policy);
  if (o->median_policy < 0 || o->median_policy > 2) return fail(TK_ERR_ARG, "bad median_policy %d", o->median_policy);
  TK_HIP(hipSetDevice(c->device));
  if (b->use_graph) return batch_run_graph(c, b, o, d_stats, d_votes);
  return batch_run_eager(c, b, o, d_stats, d_votes);
}

int tk_batch_run(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* stats, tk_r2t_vote* votes) {
  if (!c || !b || !o) return fail(TK_ERR_ARG, "NULL argument");
  const uint64_t n = b->reads->n_reads;
  tk_read_stat* ds = nullptr;
  tk_r2t_vote* dv = nullptr;
  int rc = TK_OK;
  if (stats) rc = dev_alloc((void**)&ds, std::max<uint64_t>(n, 1) * sizeof(tk_read_stat));
  if (!rc && votes) rc = dev_alloc((void**)&dv, std::max<uint64_t>(n, 1) * sizeof(tk_r2t_vote));
  const bool g = b->use_graph;
  b->use_graph = false;  // buffers are per call here: run eagerly
  if (!rc) rc = tk_batch_run_d(c, b, o, ds, dv);
  b->use_graph = g;
  if (!rc && stats && n) {
    hipError_t e = hipMemcpyAsync(stats, ds, n * sizeof(tk_read_stat), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "batch stats download");
  }
  if (!rc && votes && n) {
    hipError_t e = hipMemcpyAsync(votes, dv, n * sizeof(tk_r2t_vote), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "batch votes download");
  }
  dev_free(ds, std::max<uint64_t>(n, 1) * sizeof(tk_read_stat));
  dev_free(dv, std::max<uint64_t>(n, 1) * sizeof(tk_r2t_vote));
  return rc;
}

int tk_batch_autotune(tk_ctx* c, tk_batch* b, const tk_stats_opts* o, tk_read_stat* d_stats,
                      tk_r2t_vote* d_votes, double* best_ms) {
  if (!c || !b || !o) return fail(TK_ERR_ARG, "NULL argument");
  TK_HIP(hipSetDevice(c->device));
  static const double kLoads[] = {0.35, 0.7};
  static const int kWaves[] = {2, 4};
  double best = 1e300, best_load = b->window_load;
  int best_waves = b->stats_waves;
  hipEvent_t e0, e1;
  TK_HIP(hipEventCreate(&e0));
  TK_HIP(hipEventCreate(&e1));
  int rc = TK_OK;
  for (double load : kLoads) {
    b->window_load = load;
    if ((rc = batch_layout(c, b))) break;
    for (int w : kWaves) {
      b->stats_waves = w;
      if ((rc = batch_run_eager(c, b, o, d_stats, d_votes))) break;  // warm
      (void)hipEventRecord(e0, c->stream);
      if ((rc = batch_enqueue(c, b, o, d_stats, d_votes))) break;
      (void)hipEventRecord(e1, c->stream);
      if ((rc = batch_check(c))) break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; best_load = load; best_waves = w; }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->window_load = best_load;
  b->stats_waves = best_waves;
  if ((rc = batch_layout(c, b))) return rc;
  if (best_ms) *best_ms = best;
  return TK_OK;
}

int tk_batch_info(const tk_batch* b, uint64_t* device_bytes, double* window_load, int* stats_waves) {
  if (!b) return fail(TK_ERR_ARG, "batch is NULL");
  if (device_bytes) *device_bytes = b->bytes;
  if (window_load) *window_load = b->window_load;
  if (stats_waves) *stats_waves = b->stats_waves;
  return TK_OK;
}

int tk_batch_free(tk_batch* b) {
  if (!b) return TK_OK;
  (void)hipSetDevice(b->device);
  batch_release_device(b);
  dev_free(b->d_read_comp, std::max<uint64_t>(b->reads->n_reads, 1) * 4);
  dev_free(b->d_contig_comp, std::max<uint64_t>(b->contigs->n_reads, 1) * 4);
  dev_free(b->d_contig_local, std::max<uint64_t>(b->contigs->n_reads, 1) * 4);
  dev_free(b->d_groups, std::max<uint64_t>(b->n_groups, 1) * sizeof(BatchGroup));
  dev_free(b->d_order, b->reads->n_reads * 4);
  dev_free(b->d_contig_off, (b->n_comp + 1) * 8);
  delete b;
  return TK_OK;
}

}  // extern "C"
