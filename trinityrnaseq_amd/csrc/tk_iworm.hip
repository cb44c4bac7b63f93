// tk_iworm.hip -- Inchworm greedy contig construction over a count table.
//
// Replaces the assembly step of `inchworm --kmers <dump> --run_inchworm -K 25
// -L 25 --DS` (Trinity:2654-2711): seeds in order of count (descending, then
// key), each extended to the right and then to the left by the most abundant
// unused k-mer that overlaps the contig end by k-1 bases (ties: A<C<G<T),
// every k-mer used at most once.  The result is the serial greedy's, bit for
// bit, computed by many walks at once:
//
//  * seeds: the table's keys with count >= min_seed_cov (and a seed-entropy
//    composition table) sorted by key, then stably by inverted count;
//  * rounds: the next B seeds still unused are walked at once, one four-lane
//    group per walk (lane b looks up the candidate ending, or starting, with
//    base b).  Walk w of round base S has rank S+1+w; a walk claims each
//    k-mer it takes with atomicMin(claim[slot], rank), so a k-mer goes to the
//    earliest walk that wants it and a later walk that held it is marked lost.
//    A walk treats k-mers claimed by earlier ranks (committed rounds, or
//    earlier walks of this round) as used;
//  * commit: in rank order, walks that finished without losing a k-mer, and
//    seeds an earlier walk had already taken (skipped), form the committed
//    prefix; the first lost or overflowing walk ends it.  By induction every
//    committed walk equals the serial greedy walk: the earlier walks it saw
//    as claimers are committed with exactly those k-mers, and any k-mer an
//    earlier serial walk takes was taken from it in this round (it would be
//    lost).  Walks past the prefix release their claims and are walked again
//    in a later round; the first walk of a round always commits, so every
//    round progresses.
//
// Claims are one u32 per table slot (tables of < 2^32 slots).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "tk_internal.h"

struct tk_contigs {
  int k = 25;
  std::string seq;
  std::vector<uint64_t> off{0};
  std::vector<uint64_t> sums;
  std::vector<uint64_t> seeds;
  uint64_t rounds = 0;
  uint64_t walks = 0;   // walks launched (committed or not)
  uint64_t steps = 0;   // k-mers appended by committed walks
};

namespace tk {

namespace {

constexpr uint32_t kFree = 0xFFFFFFFFu;
enum : uint32_t { kRunning = 0, kDone = 1, kSkip = 2, kLost = 3, kOverflow = 4 };

#define IW_LAUNCH(ctx, name, grid, block, kernel, ...)                                         \
  do {                                                                                         \
    hipEvent_t _a;                                                                             \
    ::tk::timing_begin((ctx), &_a);                                                            \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (ctx)->stream, __VA_ARGS__);        \
    ::tk::timing_end((ctx), (name), _a);                                                       \
    hipError_t _le = hipGetLastError();                                                        \
    if (_le != hipSuccess) return ::tk::hip_fail(_le, "launch " name);                         \
  } while (0)

inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Count of key (0 when absent or pruned) and its global slot index (~0 when absent).
__device__ __forceinline__ uint64_t table_find(const TableView& t, uint64_t key, uint32_t* slot) {
  const uint64_t tagged = key | t.occ;
  uint64_t b = bucket0(t, key);
  for (uint32_t p = 0; p < t.max_probe; p++) {
    const ulonglong2* l2 = reinterpret_cast<const ulonglong2*>(t.slots + (b << 3));
    const ulonglong2 v0 = l2[0], v1 = l2[1], v2 = l2[2], v3 = l2[3];
    const uint64_t sv[8] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint64_t s = sv[j];
      if ((s & t.kmo) == tagged) {
        *slot = (uint32_t)((b << 3) + j);
        return slot_count(t, s);
      }
      if (s == 0) {
        *slot = 0xFFFFFFFFu;
        return 0;
      }
    }
    b = bucket_next(t, b);
  }
  *slot = 0xFFFFFFFFu;
  return 0;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
  const int lo = __shfl_xor((int)(uint32_t)v, mask, 4);
  const int hi = __shfl_xor((int)(uint32_t)(v >> 32), mask, 4);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Seed candidates: (key, slot) of every slot with count >= min_cov whose
// composition passes keep (NULL: all), appended at *cursor.
__global__ void k_iw_seeds(TableView t, uint64_t n_slots, int k, uint64_t min_cov,
                           const uint8_t* __restrict__ keep, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ slots, uint64_t cap, unsigned long long* __restrict__ cursor) {
  const int lane = threadIdx.x & 63;
  const int K1 = k + 1;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n_slots;
       base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool take = false;
    uint64_t key = 0;
    if (i < n_slots) {
      const uint64_t s = t.slots[i];
      if (s && (s >> t.cshift) != 0) {
        const uint64_t v = slot_count(t, s);
        key = s & (t.occ - 1);
        take = v >= min_cov && v > 0;
        if (take && keep) {
          int n[4] = {0, 0, 0, 0};
          uint64_t x = key;
          for (int j = 0; j < k; j++) {
            n[x & 3]++;
            x >>= 2;
          }
          take = keep[(n[0] * K1 + n[1]) * K1 + n[2]] != 0;
        }
      }
    }
    const uint64_t m = __ballot(take);
    if (!m) continue;
    unsigned long long start = 0;
    if (lane == 0) start = atomicAdd(cursor, (unsigned long long)__popcll(m));
    start = __shfl(start, 0);
    if (take) {
      const uint64_t pos = start + __popcll(m & ((1ull << lane) - 1));
      if (pos < cap) {  // the host refuses a count past cap
        keys[pos] = key;
        slots[pos] = (uint32_t)i;
      }
    }
  }
}

// keys[i] = 2^32-1 - min(count of slot slots[i], 2^32-1): sorting by it
// (stably) orders seeds by count descending.
__global__ void k_iw_inv_count(TableView t, const uint32_t* __restrict__ slots, uint64_t n,
                               uint64_t* __restrict__ keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t v = slot_count(t, t.slots[slots[i]]);
    v = v > 0xFFFFFFFFull ? 0xFFFFFFFFull : v;
    keys[i] = 0xFFFFFFFFull - v;
  }
}

__global__ void k_iw_fill(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Committed claims keep their order only relative to new ranks: all become 1.
__global__ void k_iw_renumber(uint32_t* __restrict__ claim, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (claim[i] != kFree) claim[i] = 1u;
}

// One workgroup: the first `want` seeds from position `cursor` on whose slot
// is unclaimed.  out[0] = number picked; out64[1] = the position after the
// last one examined.  Each thread examines four consecutive positions per
// pass (their seed and claim loads in flight together), ranked by a block
// scan of the threads' counts, so a pass covers 4,096 positions.
constexpr uint32_t kPickPer = 4;

__global__ __launch_bounds__(1024) void k_iw_pick(const uint32_t* __restrict__ seeds, uint64_t n_seeds,
                                                  uint64_t cursor, const uint32_t* __restrict__ claim,
                                                  uint32_t want, uint32_t* __restrict__ picked_slot,
                                                  uint64_t* __restrict__ picked_pos,
                                                  unsigned long long* __restrict__ out) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_total;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t span = (uint64_t)kPickPer * blockDim.x;
  uint32_t got = 0;
  uint64_t base = cursor;
  uint64_t end = cursor;
  while (got < want && base < n_seeds) {
    const uint64_t p0 = base + (uint64_t)kPickPer * threadIdx.x;
    uint32_t sl[kPickPer];
    bool fr[kPickPer];
#pragma unroll
    for (uint32_t j = 0; j < kPickPer; j++) sl[j] = p0 + j < n_seeds ? seeds[p0 + j] : 0u;
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPickPer; j++) {
      fr[j] = p0 + j < n_seeds && claim[sl[j]] == kFree;
      cnt += fr[j] ? 1u : 0u;
    }
    // inclusive scan of the counts within the wave, then across the waves
    uint32_t x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t acc = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        const uint32_t c = s_wave[w];
        s_wave[w] = acc;
        acc += c;
      }
      s_total = acc;
    }
    __syncthreads();
    uint32_t rank = got + s_wave[wave] + x - cnt;
#pragma unroll
    for (uint32_t j = 0; j < kPickPer; j++) {
      if (!fr[j]) continue;
      if (rank < want) {
        picked_slot[rank] = sl[j];
        picked_pos[rank] = p0 + j;
        if (rank == want - 1) out[1] = p0 + j + 1;  // the last picked seed ends the scan
      }
      rank++;
    }
    const uint32_t total = s_total;
    if (got + total >= want) {
      got = want;
    } else {
      got += total;
      end = std::min<uint64_t>(base + span, n_seeds);
    }
    base += span;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = got;
    if (got < want) out[1] = end;
  }
}

struct WalkParams {
  TableView t;
  uint64_t kmask;
  uint32_t S;       // ranks of this round: S+1 ... S+n
  uint32_t n;       // walks
  uint32_t lmax;    // path entries per walk and side
  int k;
  int canonical;
};

constexpr uint32_t kDeps = 8;  // recorded dependencies per walk (more: the walk is redone)

// Record that walk w treated a k-mer claimed by earlier walk `dep` of this
// round as used (the decision holds only if that walk commits a contig).
__device__ __forceinline__ void iw_dep(uint32_t* __restrict__ deps, uint32_t* __restrict__ dep_n, uint32_t w,
                                       uint32_t dep, uint32_t& last) {
  if (dep == last) return;
  last = dep;
  const uint32_t j = atomicAdd(&dep_n[w], 1u);
  if (j < kDeps) deps[w * kDeps + j] = dep;
}

// One four-lane group per walk: claim the seed, extend right, then left.
__global__ __launch_bounds__(256) void k_iw_walk(WalkParams p, const uint32_t* __restrict__ picked_slot,
                                                 uint32_t* __restrict__ claim, uint32_t* __restrict__ lost,
                                                 uint32_t* __restrict__ status, uint32_t* __restrict__ lens,
                                                 uint64_t* __restrict__ sums, uint32_t* __restrict__ path_slot,
                                                 uint8_t* __restrict__ path_base, uint32_t* __restrict__ deps,
                                                 uint32_t* __restrict__ dep_n) {
  const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
  if (w >= p.n) return;  // whole groups leave together
  const int lane = threadIdx.x & 63;
  const int b = lane & 3;
  const int g0 = lane & ~3;
  const TableView& t = p.t;
  const uint32_t r = p.S + 1 + w;
  const uint32_t seed_slot = picked_slot[w];
  const uint64_t seed_word = t.slots[seed_slot];
  const uint64_t seed_key = seed_word & (t.occ - 1);
  uint64_t sum = slot_count(t, seed_word);
  uint32_t st = kRunning;
  uint32_t last_dep = 0xFFFFFFFFu;
  if (b == 0) {
    const uint32_t old = atomicMin(&claim[seed_slot], r);
    if (old < r) {
      st = kSkip;
      if (old > p.S) iw_dep(deps, dep_n, w, old - p.S - 1, last_dep);
    } else if (old != kFree && old != r) {
      atomicOr(&lost[old - p.S - 1], 1u);
    }
  }
  st = __shfl(st, g0);
  uint32_t lens_side[2] = {0, 0};
  for (int side = 0; side < 2 && st == kRunning; side++) {
    uint64_t x = seed_key;
    uint32_t len = 0;
    uint32_t* pslot = path_slot + ((uint64_t)w * 2 + side) * p.lmax;
    uint8_t* pbase = path_base + ((uint64_t)w * 2 + side) * p.lmax;
    for (;;) {
      // this walk's lost flag, loaded first so that its latency hides behind the lookup
      uint32_t l = 0;
      if (b == 0) l = __hip_atomic_load(&lost[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t y = side == 0 ? (((x << 2) | (uint64_t)b) & p.kmask)
                                   : (((uint64_t)b << (2 * (p.k - 1))) | (x >> 2));
      uint32_t slot;
      const uint64_t c = table_find(t, canon_key(y, p.k, p.canonical), &slot);
      // Candidates are tried by count (then base) with the claim itself: a
      // claim that returns an earlier rank (committed, this walk's own, or an
      // earlier walk of the round) drops the candidate, so the k-mer taken is
      // the best one no earlier walk holds -- one round trip per step instead
      // of reading the claim word first.
      bool avail = c > 0;
      bool took = false;
      uint32_t best_b = 4;
      uint64_t best_c = 0;
      for (int attempt = 0; attempt < 4; attempt++) {
        // best remaining candidate of the group: count, then smallest base
        uint64_t key = avail ? (((c < (1ull << 61) ? c : (1ull << 61)) << 2) | (uint64_t)(3 - b)) : 0;
        key = std::max(key, shfl_xor_u64(key, 1));
        key = std::max(key, shfl_xor_u64(key, 2));
        if (key == 0) break;
        best_b = 3 - (uint32_t)(key & 3);
        best_c = key >> 2;
        if (len == p.lmax) {  // no room to record the claim: stop unclaimed
          st = kOverflow;
          break;
        }
        uint32_t old = 0;
        if ((uint32_t)b == best_b) old = atomicMin(&claim[slot], r);
        old = __shfl(old, g0 | (int)best_b);
        if (old > r) {
          if ((uint32_t)b == best_b && old != kFree) atomicOr(&lost[old - p.S - 1], 1u);
          took = true;
          break;
        }
        if ((uint32_t)b == best_b) {  // held by an earlier rank (or this walk)
          avail = false;
          if (old > p.S && old < r) iw_dep(deps, dep_n, w, old - p.S - 1, last_dep);
        }
      }
      if (!took) break;
      const uint32_t bslot = (uint32_t)__shfl((int)slot, g0 | (int)best_b);
      if (b == 0) {
        pslot[len] = bslot;
        pbase[len] = (uint8_t)best_b;
      }
      len++;
      // the count field of a taken slot is exact up to 2^61 (clamped above)
      sum += best_c;
      x = side == 0 ? (((x << 2) | (uint64_t)best_b) & p.kmask)
                    : (((uint64_t)best_b << (2 * (p.k - 1))) | (x >> 2));
      l = __shfl(l, g0);
      if (l) {
        st = kLost;
        break;
      }
    }
    lens_side[side] = len;
  }
  if (b == 0) {
    status[w] = st == kRunning ? kDone : st;
    lens[2 * w] = lens_side[0];
    lens[2 * w + 1] = lens_side[1];
    sums[w] = sum;
  }
}

// The seed slot's final claim of every walk (its own rank unless an earlier
// walk took the seed).
__global__ void k_iw_owner(const uint32_t* __restrict__ picked_slot, uint32_t n, const uint32_t* __restrict__ claim,
                           uint32_t* __restrict__ owner) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x)
    owner[w] = claim[picked_slot[w]];
}

// Flagged walks give their claims back (seed and path).
__global__ void k_iw_release(uint32_t n, uint32_t S, uint32_t lmax, const uint32_t* __restrict__ picked_slot,
                             const uint32_t* __restrict__ release, const uint32_t* __restrict__ lens,
                             const uint32_t* __restrict__ path_slot, uint32_t* __restrict__ claim) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * 2;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w = (uint32_t)(i >> 1);
    const int side = (int)(i & 1);
    const uint32_t r = S + 1 + w;
    if (!release[w]) continue;
    if (side == 0) atomicCAS(&claim[picked_slot[w]], r, kFree);
    const uint32_t* ps = path_slot + ((uint64_t)w * 2 + side) * lmax;
    for (uint32_t j = 0; j < lens[2 * w + side]; j++) atomicCAS(&claim[ps[j]], r, kFree);
  }
}

// One block per committed contig: left bases reversed, the seed, right bases.
__global__ __launch_bounds__(256) void k_iw_emit(const uint32_t* __restrict__ walks, const uint64_t* __restrict__ out_off,
                                                 const uint32_t* __restrict__ picked_slot, TableView t, int k,
                                                 uint32_t lmax, const uint32_t* __restrict__ lens,
                                                 const uint8_t* __restrict__ path_base, char* __restrict__ out) {
  const uint32_t w = walks[blockIdx.x];
  const uint32_t nr = lens[2 * w], nl = lens[2 * w + 1];
  const uint64_t key = t.slots[picked_slot[w]] & (t.occ - 1);
  const uint8_t* pr = path_base + ((uint64_t)w * 2) * lmax;
  const uint8_t* pl = path_base + ((uint64_t)w * 2 + 1) * lmax;
  char* o = out + out_off[blockIdx.x];
  const uint64_t L = (uint64_t)nl + k + nr;
  for (uint64_t i = threadIdx.x; i < L; i += blockDim.x) {
    uint32_t code;
    if (i < nl) code = pl[nl - 1 - i];
    else if (i < (uint64_t)nl + k) code = (uint32_t)((key >> (2 * (k - 1 - (i - nl)))) & 3);
    else code = pr[i - nl - k];
    o[i] = "ACGT"[code];
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) dev_free(p, bytes);
  }
  int alloc(size_t b) {
    if (p) dev_free(p, bytes);
    p = nullptr;
    bytes = b;
    return dev_alloc(&p, b ? b : 8);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

int inchworm_run(tk_ctx* c, tk_table* t, uint32_t min_seed_cov, const uint8_t* seed_keep, tk_contigs* res) {
  const uint64_t n_slots = t->nb * 8;
  if (n_slots >= 0xFFFFFFFFull)
    return fail(TK_ERR_ARG, "inchworm: tables of 2^32 or more slots are not supported (%llu slots)",
                (unsigned long long)n_slots);
  const int k = t->k;
  TableView v = table_view(t);
  int rc;
  // ---- seeds in order: key ascending, then (stably) count descending
  DevBuf dkeep, dseedk, dseeds, dclaim;
  const int K1 = k + 1;
  if (seed_keep) {
    if ((rc = dkeep.alloc((size_t)K1 * K1 * K1))) return rc;
    TK_HIP(hipMemcpyAsync(dkeep.p, seed_keep, (size_t)K1 * K1 * K1, hipMemcpyHostToDevice, c->stream));
  }
  uint64_t n_cand = 0;
  {
    unsigned long long* cur = reinterpret_cast<unsigned long long*>(c->d_flags + 8);
    TK_HIP(hipMemsetAsync(cur, 0, 8, c->stream));
    // upper bound on seeds: the table's keys
    const uint64_t cap = std::max<uint64_t>(t->n_keys, 1);
    if ((rc = dseedk.alloc(cap * 8)) || (rc = dseeds.alloc(cap * 4))) return rc;
    IW_LAUNCH(c, "iworm_seeds", std::min<uint64_t>(cdiv(n_slots, 256), 65536), 256, k_iw_seeds, v, n_slots, k,
              (uint64_t)min_seed_cov, seed_keep ? dkeep.as<uint8_t>() : nullptr, dseedk.as<uint64_t>(),
              dseeds.as<uint32_t>(), cap, cur);
    TK_HIP(hipMemcpyAsync(&n_cand, cur, 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    if (n_cand > cap) return fail(TK_ERR_STATE, "inchworm: %llu seeds in a table of %llu keys",
                                  (unsigned long long)n_cand, (unsigned long long)cap);
    if ((rc = sort_pairs_u64_u32(c, dseedk.as<uint64_t>(), dseeds.as<uint32_t>(), n_cand, 2 * k))) return rc;
    if (n_cand)
      IW_LAUNCH(c, "iworm_order", std::min<uint64_t>(cdiv(n_cand, 256), 65536), 256, k_iw_inv_count, v,
                dseeds.as<uint32_t>(), n_cand, dseedk.as<uint64_t>());
    if ((rc = sort_pairs_u64_u32(c, dseedk.as<uint64_t>(), dseeds.as<uint32_t>(), n_cand, 32))) return rc;
  }
  if ((rc = dclaim.alloc(n_slots * 4))) return rc;
  IW_LAUNCH(c, "iworm_fill", std::min<uint64_t>(cdiv(n_slots, 256), 65536), 256, k_iw_fill,
            dclaim.as<uint32_t>(), n_slots, kFree);
  // ---- rounds
  // TK_IWORM_WALKS caps the walks per round (default 16,384, starting at
  // 1,024) and TK_IWORM_LMAX sets the first path capacity per walk side
  // (default 4,096 k-mers): small values force the conflict and regrowth
  // paths on small inputs.
  uint32_t Bmax = 16384, lmax = 4096;
  if (const char* e = getenv("TK_IWORM_WALKS")) Bmax = (uint32_t)std::min<unsigned long>(std::max(1ul, strtoul(e, nullptr, 10)), 16384ul);
  if (const char* e = getenv("TK_IWORM_LMAX")) lmax = (uint32_t)std::min<unsigned long>(std::max(4ul, strtoul(e, nullptr, 10)), 1ul << 20);
  const uint32_t Bmin = std::min<uint32_t>(64, Bmax);
  // path buffers hold Bcap walks: at most 2^28 path entries in all
  auto cap_for = [&](uint32_t lm) {
    return std::max<uint32_t>(Bmin, (uint32_t)std::min<uint64_t>(Bmax, (1ull << 28) / lm));
  };
  uint32_t Bcap = cap_for(lmax);
  uint32_t B = std::min<uint32_t>(1024, Bcap);
  DevBuf dpick, dpos, dout, dlost, dstat, dlens, dsums, dpslot, dpbase, dwalks, dwoff, dchars, ddeps, ddepn,
      downer, drel;
  if ((rc = dpick.alloc(Bmax * 4)) || (rc = dpos.alloc(Bmax * 8)) || (rc = dout.alloc(16)) ||
      (rc = dlost.alloc(Bmax * 4)) || (rc = dstat.alloc(Bmax * 4)) || (rc = dlens.alloc(Bmax * 8)) ||
      (rc = dsums.alloc(Bmax * 8)) || (rc = dwalks.alloc(Bmax * 4)) || (rc = dwoff.alloc(Bmax * 8 + 8)) ||
      (rc = ddeps.alloc((size_t)Bmax * kDeps * 4)) || (rc = ddepn.alloc(Bmax * 4)) || (rc = downer.alloc(Bmax * 4)) ||
      (rc = drel.alloc(Bmax * 4)))
    return rc;
  auto alloc_paths = [&]() -> int {
    int r2 = dpslot.alloc((size_t)Bcap * 2 * lmax * 4);
    if (!r2) r2 = dpbase.alloc((size_t)Bcap * 2 * lmax);
    return r2;
  };
  if ((rc = alloc_paths())) return rc;
  // the per-round host copies live in page-locked memory, so each is one DMA
  // (copies to pageable memory are staged and wait on the host)
  struct Pinned {
    void* p = nullptr;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
  } pin;
  // u32: stat, lost, lens (2), depn, owner, rel, walks, deps (kDeps) per walk;
  // u64: sums, pos, woff per walk, woff's end, out (2)
  const size_t u32s = (size_t)Bmax * (8 + kDeps), u64s = (size_t)Bmax * 3 + 1 + 2;
  TK_HIP(hipHostMalloc(&pin.p, u64s * 8 + u32s * 4, hipHostMallocDefault));
  uint64_t* h_sums = static_cast<uint64_t*>(pin.p);
  uint64_t* h_pos = h_sums + Bmax;
  uint64_t* h_woff = h_pos + Bmax;  // Bmax + 1
  unsigned long long* h_out = reinterpret_cast<unsigned long long*>(h_woff + Bmax + 1);  // 2
  uint32_t* h_stat = reinterpret_cast<uint32_t*>(h_out + 2);
  uint32_t* h_lost = h_stat + Bmax;
  uint32_t* h_lens = h_lost + Bmax;  // 2 * Bmax
  uint32_t* h_depn = h_lens + 2 * (size_t)Bmax;
  uint32_t* h_owner = h_depn + Bmax;
  uint32_t* h_rel = h_owner + Bmax;
  uint32_t* h_walks = h_rel + Bmax;
  uint32_t* h_deps = h_walks + Bmax;  // Bmax * kDeps
  std::vector<uint8_t> late(Bmax);
  uint64_t cursor = 0;
  uint32_t S = 0;
  std::string chunk;
  for (;;) {
    if ((uint64_t)S + 2ull * Bmax + 2 >= kFree) {
      IW_LAUNCH(c, "iworm_renumber", std::min<uint64_t>(cdiv(n_slots, 256), 65536), 256, k_iw_renumber,
                dclaim.as<uint32_t>(), n_slots);
      S = 1;
    }
    h_out[0] = h_out[1] = 0;
    IW_LAUNCH(c, "iworm_pick", 1, 1024, k_iw_pick, dseeds.as<uint32_t>(), n_cand, cursor, dclaim.as<uint32_t>(), B,
              dpick.as<uint32_t>(), dpos.as<uint64_t>(), dout.as<unsigned long long>());
    TK_HIP(hipMemcpyAsync(h_out, dout.p, 16, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    const uint32_t n = (uint32_t)h_out[0];
    if (n == 0) break;
    res->rounds++;
    res->walks += n;
    IW_LAUNCH(c, "iworm_fill", cdiv(n, 256), 256, k_iw_fill, dlost.as<uint32_t>(), (uint64_t)n, 0u);
    IW_LAUNCH(c, "iworm_fill", cdiv(n, 256), 256, k_iw_fill, ddepn.as<uint32_t>(), (uint64_t)n, 0u);
    WalkParams wp;
    wp.t = v;
    wp.kmask = k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    wp.S = S;
    wp.n = n;
    wp.lmax = lmax;
    wp.k = k;
    wp.canonical = t->canonical;
    IW_LAUNCH(c, "iworm_walk", cdiv((uint64_t)n * 4, 256), 256, k_iw_walk, wp, dpick.as<uint32_t>(),
              dclaim.as<uint32_t>(), dlost.as<uint32_t>(), dstat.as<uint32_t>(), dlens.as<uint32_t>(),
              dsums.as<uint64_t>(), dpslot.as<uint32_t>(), dpbase.as<uint8_t>(), ddeps.as<uint32_t>(),
              ddepn.as<uint32_t>());
    IW_LAUNCH(c, "iworm_owner", cdiv(n, 256), 256, k_iw_owner, dpick.as<uint32_t>(), n, dclaim.as<uint32_t>(),
              downer.as<uint32_t>());
    TK_HIP(hipMemcpyAsync(h_depn, ddepn.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_deps, ddeps.p, (size_t)n * kDeps * 4, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_owner, downer.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_stat, dstat.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_lost, dlost.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_lens, dlens.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_sums, dsums.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipMemcpyAsync(h_pos, dpos.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    // Committed prefix, in rank order.  A walk whose seed an earlier walk of
    // the round took ("late skip") is the serial skip of that seed when the
    // taker commits a contig; its own claims are void.  Any decision that
    // treated a k-mer as used because an earlier walk of the round held it
    // (the dependencies) holds only if that walk commits a contig, i.e. is
    // not a late skip.  Seeds claimed at the start by an earlier walk are
    // skips under the same condition.  The first walk that lost a k-mer
    // without losing its seed, outgrew its buffers or recorded too many
    // dependencies ends the prefix.
    uint32_t P = n;
    for (uint32_t i = 0; i < n; i++) {
      late[i] = 0;
      bool deps_ok = h_depn[i] <= kDeps;
      for (uint32_t j = 0; deps_ok && j < h_depn[i]; j++) {
        const uint32_t d = h_deps[(size_t)i * kDeps + j];
        deps_ok = d < i && !late[d];
      }
      const uint32_t r = S + 1 + i;
      if (h_stat[i] == kSkip) {
        if (!deps_ok) { P = i; break; }
        continue;
      }
      if (h_owner[i] != r) {  // seed taken during the round by walk h_owner[i] - S - 1 < i
        const uint32_t o = h_owner[i] - S - 1;
        if (h_owner[i] <= S || o >= i || late[o]) { P = i; break; }
        late[i] = 1;
        continue;
      }
      if (h_stat[i] == kDone && !h_lost[i] && deps_ok) continue;
      P = i;
      break;
    }
    // release: late skips of the prefix and every walk past it
    bool any_rel = false;
    for (uint32_t i = 0; i < n; i++) {
      h_rel[i] = (i >= P && h_stat[i] != kSkip) || (i < P && late[i]) ? 1u : 0u;
      any_rel = any_rel || h_rel[i];
    }
    if (any_rel) {
      TK_HIP(hipMemcpyAsync(drel.p, h_rel, n * 4, hipMemcpyHostToDevice, c->stream));
      IW_LAUNCH(c, "iworm_release", std::min<uint64_t>(cdiv((uint64_t)n * 2, 256), 4096), 256, k_iw_release, n, S,
                lmax, dpick.as<uint32_t>(), drel.as<uint32_t>(), dlens.as<uint32_t>(), dpslot.as<uint32_t>(),
                dclaim.as<uint32_t>());
    }
    if (P == 0) {
      // the first walk outgrew its path buffers: larger buffers, fewer walks
      if (h_stat[0] != kOverflow) return fail(TK_ERR_STATE, "inchworm: first walk of a round did not commit");
      if (lmax >= (1u << 28)) return fail(TK_ERR_STATE, "inchworm: walk longer than 2^28 k-mers");
      TK_HIP(hipStreamSynchronize(c->stream));
      lmax *= 2;
      Bcap = cap_for(lmax);
      B = std::min(B, Bcap);
      if ((rc = alloc_paths())) return rc;
      S += n;
      continue;
    }
    // contigs of the committed walks, in rank order
    uint32_t nw = 0;
    uint64_t total = 0;
    for (uint32_t i = 0; i < P; i++) {
      if (h_stat[i] != kDone || late[i]) continue;
      h_walks[nw] = i;
      h_woff[nw] = total;
      const uint64_t L = (uint64_t)h_lens[2 * i] + h_lens[2 * i + 1] + k;
      total += L;
      res->sums.push_back(h_sums[i]);
      res->seeds.push_back(0);  // filled from the sequence below
      res->steps += L - k;
      nw++;
    }
    if (nw) {
      if (dchars.bytes < total && (rc = dchars.alloc(std::max<uint64_t>(total, 1 << 20)))) return rc;
      TK_HIP(hipMemcpyAsync(dwalks.p, h_walks, (size_t)nw * 4, hipMemcpyHostToDevice, c->stream));
      TK_HIP(hipMemcpyAsync(dwoff.p, h_woff, (size_t)nw * 8, hipMemcpyHostToDevice, c->stream));
      IW_LAUNCH(c, "iworm_emit", nw, 256, k_iw_emit, dwalks.as<uint32_t>(), dwoff.as<uint64_t>(),
                dpick.as<uint32_t>(), v, k, lmax, dlens.as<uint32_t>(), dpbase.as<uint8_t>(), dchars.as<char>());
      chunk.resize(total);
      TK_HIP(hipMemcpyAsync(&chunk[0], dchars.p, total, hipMemcpyDeviceToHost, c->stream));
      TK_HIP(hipStreamSynchronize(c->stream));
      const uint64_t base_n = res->off.size() - 1;
      for (uint32_t j = 0; j < nw; j++) {
        const uint32_t i = h_walks[j];
        const uint64_t L = (uint64_t)h_lens[2 * i] + h_lens[2 * i + 1] + k;
        // seed key: the k bases after the left extension
        uint64_t key = 0;
        const char* s = chunk.data() + h_woff[j] + h_lens[2 * i + 1];
        for (int q = 0; q < k; q++) key = (key << 2) | (uint64_t)(s[q] == 'A' ? 0 : s[q] == 'C' ? 1 : s[q] == 'G' ? 2 : 3);
        res->seeds[base_n + j] = key;
        res->off.push_back(res->off.back() + L);
      }
      res->seq += chunk;
    }
    cursor = P < n ? h_pos[P] : h_out[1];
    S += n;
    if (P == n) B = std::min(Bcap, B * 2);
    else if (P < n / 8) B = std::max(Bmin, B / 2);
  }
  return TK_OK;
}

}  // namespace

}  // namespace tk

extern "C" {

int tk_inchworm(tk_ctx* c, tk_table* t, uint32_t min_seed_cov, const uint8_t* seed_keep, tk_contigs** out) {
  using namespace tk;
  if (!c || !t || !out) return fail(TK_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (hipSetDevice(c->device) != hipSuccess) return fail(TK_ERR_HIP, "hipSetDevice");
  int rc = table_finalize(c, t);
  if (rc) return rc;
  tk_contigs* r = new tk_contigs();
  r->k = t->k;
  rc = inchworm_run(c, t, min_seed_cov, seed_keep, r);
  (void)hipStreamSynchronize(c->stream);
  if (rc) {
    delete r;
    return rc;
  }
  *out = r;
  return TK_OK;
}

int tk_contigs_info(const tk_contigs* r, uint64_t* n_contigs, uint64_t* total_len, uint64_t* rounds,
                    uint64_t* walks) {
  if (!r) return tk::fail(TK_ERR_ARG, "NULL argument");
  if (n_contigs) *n_contigs = r->off.size() - 1;
  if (total_len) *total_len = r->seq.size();
  if (rounds) *rounds = r->rounds;
  if (walks) *walks = r->walks;
  return TK_OK;
}

int tk_contigs_get(const tk_contigs* r, char* seq, uint64_t* offsets, uint64_t* sums, uint64_t* seed_keys) {
  if (!r) return tk::fail(TK_ERR_ARG, "NULL argument");
  const size_t n = r->off.size() - 1;
  if (seq && !r->seq.empty()) memcpy(seq, r->seq.data(), r->seq.size());
  if (offsets) memcpy(offsets, r->off.data(), (n + 1) * 8);
  if (sums && n) memcpy(sums, r->sums.data(), n * 8);
  if (seed_keys && n) memcpy(seed_keys, r->seeds.data(), n * 8);
  return TK_OK;
}

int tk_contigs_free(tk_contigs* r) {
  delete r;
  return TK_OK;
}

}  // extern "C"
