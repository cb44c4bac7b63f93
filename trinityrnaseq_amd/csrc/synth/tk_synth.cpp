// tk_synth.cpp -- seeded synthetic transcriptomes and Illumina-like reads.
//
// Bench/test input generator (host C++, multithreaded).  Every read is a pure
// function of (seed, read index), so any slice of a large read set can be
// regenerated on its own (the CPU baseline scores such a slice).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1p-53; }  // [0,1)
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

uint64_t mix(uint64_t a, uint64_t b) {
  SplitMix m(a * 0xD1B54A32D192ED03ULL ^ (b + 0x8CB92BA72F3D8DD7ULL));
  m.next();
  return m.next();
}

struct Tx {
  std::vector<uint8_t> codes;
  std::vector<uint64_t> starts, lengths;
  std::vector<double> weights, cdf;
};

const uint8_t kAscii[5] = {'A', 'C', 'G', 'T', 'N'};

// Geometric gap to the next event of probability p (>= 0).
inline uint64_t geom_gap(SplitMix& r, double lp1) {
  if (lp1 >= 0) return UINT64_MAX;
  double u = r.uniform();
  if (u < 1e-300) u = 1e-300;
  double g = floor(log(u) / lp1);
  return g > 1e18 ? UINT64_MAX : (uint64_t)g;
}

void mutate(SplitMix& r, uint8_t* s, int L, double sub_rate, double n_rate) {
  const double ls = sub_rate > 0 ? log1p(-sub_rate) : 0.0;
  const double ln = n_rate > 0 ? log1p(-n_rate) : 0.0;
  for (uint64_t p = geom_gap(r, ls); p < (uint64_t)L; p += 1 + geom_gap(r, ls))
    s[p] = (uint8_t)((s[p] + 1 + r.below(3)) & 3);
  for (uint64_t p = geom_gap(r, ln); p < (uint64_t)L; p += 1 + geom_gap(r, ln)) s[p] = 4;
}

// Copy read_len bases starting at signed offset s of a transcript; bases
// outside [0, TL) become N.
void copy_window(uint8_t* dst, const uint8_t* ts, int64_t TL, int64_t s, int read_len) {
  for (int j = 0; j < read_len; j++) {
    const int64_t p = s + j;
    dst[j] = (p >= 0 && p < TL) ? ts[p] : 4;
  }
}

void revcomp(uint8_t* s, int L) {
  std::reverse(s, s + L);
  for (int i = 0; i < L; i++) s[i] = s[i] < 4 ? (uint8_t)(3 - s[i]) : s[i];
}

}  // namespace

extern "C" {

typedef struct tks_tx tks_tx;

// len_model 0: uniform integer [p1, p2]; 1: lognormal(log(p1), p2) rounded and
// clipped to [lo, hi].  share_frac of transcripts are paired and copy a
// segment of [share_lo, share_hi] nt from their partner.
tks_tx* tks_transcriptome(uint64_t seed, uint64_t n, int len_model, double p1, double p2,
                          double lo, double hi, double share_frac, int share_lo, int share_hi,
                          double expr_mu, double expr_sigma) {
  Tx* t = new Tx();
  SplitMix r(mix(seed, 0x7478));
  t->lengths.resize(n);
  for (uint64_t i = 0; i < n; i++) {
    double L;
    if (len_model == 0) {
      L = p1 + (double)r.below((uint64_t)(p2 - p1 + 1));
    } else {
      L = std::round(exp(log(p1) + p2 * r.normal()));
      L = std::min(std::max(L, lo), hi);
    }
    t->lengths[i] = (uint64_t)L;
  }
  t->starts.resize(n);
  uint64_t tot = 0;
  for (uint64_t i = 0; i < n; i++) { t->starts[i] = tot; tot += t->lengths[i]; }
  t->codes.resize(tot);
  for (uint64_t i = 0; i < tot; i += 32) {
    uint64_t w = r.next();
    for (uint64_t j = i; j < std::min(tot, i + 32); j++) { t->codes[j] = w & 3; w >>= 2; }
  }
  uint64_t m = (uint64_t)(n * share_frac) / 2 * 2;
  if (m) {
    std::vector<uint64_t> perm(n);
    for (uint64_t i = 0; i < n; i++) perm[i] = i;
    for (uint64_t i = 0; i < m; i++) std::swap(perm[i], perm[i + r.below(n - i)]);
    for (uint64_t i = 0; i < m; i += 2) {
      uint64_t a = perm[i], b = perm[i + 1];
      uint64_t L = share_lo + r.below((uint64_t)(share_hi - share_lo + 1));
      L = std::min(L, std::min(t->lengths[a], t->lengths[b]));
      uint64_t sa = r.below(t->lengths[a] - L + 1), sb = r.below(t->lengths[b] - L + 1);
      memcpy(&t->codes[t->starts[b] + sb], &t->codes[t->starts[a] + sa], L);
    }
  }
  t->weights.resize(n);
  t->cdf.resize(n);
  double acc = 0;
  for (uint64_t i = 0; i < n; i++) {
    t->weights[i] = exp(expr_mu + expr_sigma * r.normal());
    acc += t->weights[i] * (double)t->lengths[i];
    t->cdf[i] = acc;
  }
  for (auto& c : t->cdf) c /= acc;
  return reinterpret_cast<tks_tx*>(t);
}

void tks_tx_info(const tks_tx* h, uint64_t* n, uint64_t* total) {
  const Tx* t = reinterpret_cast<const Tx*>(h);
  *n = t->lengths.size();
  *total = t->codes.size();
}

// ASCII of transcript i into out (length lengths[i]); returns its length.
uint64_t tks_tx_seq(const tks_tx* h, uint64_t i, char* out) {
  const Tx* t = reinterpret_cast<const Tx*>(h);
  for (uint64_t j = 0; j < t->lengths[i]; j++) out[j] = (char)kAscii[t->codes[t->starts[i] + j]];
  return t->lengths[i];
}

uint64_t tks_tx_length(const tks_tx* h, uint64_t i) {
  return reinterpret_cast<const Tx*>(h)->lengths[i];
}

void tks_tx_free(tks_tx* h) { delete reinterpret_cast<Tx*>(h); }

// Reads [first, first+n) of the stream (seed).  paired: fragment ~ N(mu,sd)
// clipped to [read_len, transcript length]; left = fragment head, right =
// reverse complement of the fragment tail; with probability rc_frac the
// fragment comes from the reverse strand (left/right swap roles).  Single
// end: a read_len window, reverse-complemented with probability rc_frac.
// Transcripts are drawn proportionally to length x expression.
void tks_reads(const tks_tx* h, uint64_t seed, int paired, uint64_t first, uint64_t n,
               int read_len, double frag_mu, double frag_sd, double sub_rate, double n_rate,
               double rc_frac, int threads, char* left, char* right) {
  const Tx* t = reinterpret_cast<const Tx*>(h);
  if (threads < 1) threads = 1;
  auto work = [&](uint64_t a, uint64_t b) {
    std::vector<uint8_t> A(read_len), B(read_len);
    for (uint64_t i = a; i < b; i++) {
      SplitMix r(mix(seed, first + i));
      const double u = r.uniform();
      uint64_t ti = (uint64_t)(std::upper_bound(t->cdf.begin(), t->cdf.end(), u) - t->cdf.begin());
      if (ti >= t->cdf.size()) ti = t->cdf.size() - 1;
      const uint64_t TL = t->lengths[ti];
      const uint8_t* ts = &t->codes[t->starts[ti]];
      const bool flip = r.uniform() < rc_frac;
      if (paired) {
        double f = std::round(frag_mu + frag_sd * r.normal());
        uint64_t F = (uint64_t)std::max(f, (double)read_len);
        if (F > TL) F = TL;
        const uint64_t s = r.below(TL - F + 1);
        copy_window(A.data(), ts, (int64_t)TL, (int64_t)s, read_len);
        copy_window(B.data(), ts, (int64_t)TL, (int64_t)(s + F) - read_len, read_len);
        revcomp(B.data(), read_len);
        uint8_t* L_ = flip ? B.data() : A.data();
        uint8_t* R_ = flip ? A.data() : B.data();
        mutate(r, L_, read_len, sub_rate, n_rate);
        mutate(r, R_, read_len, sub_rate, n_rate);
        for (int j = 0; j < read_len; j++) {
          left[i * read_len + j] = (char)kAscii[L_[j]];
          right[i * read_len + j] = (char)kAscii[R_[j]];
        }
      } else {
        const uint64_t s = r.below(TL >= (uint64_t)read_len ? TL - read_len + 1 : 1);
        copy_window(A.data(), ts, (int64_t)TL, (int64_t)s, read_len);
        if (flip) revcomp(A.data(), read_len);
        mutate(r, A.data(), read_len, sub_rate, n_rate);
        for (int j = 0; j < read_len; j++) left[i * read_len + j] = (char)kAscii[A[j]];
      }
    }
  };
  std::vector<std::thread> th;
  uint64_t per = (n + threads - 1) / threads;
  for (int k = 0; k < threads; k++) {
    uint64_t a = k * per, b = std::min(n, a + per);
    if (a < b) th.emplace_back(work, a, b);
  }
  for (auto& x : th) x.join();
}

}  // extern "C"
