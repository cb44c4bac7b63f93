// GraphFromFasta -- weld support between Inchworm contigs (Chrysalis step 1).
//
//   GraphFromFasta -i inchworm.fa -r both.fa -min_contig_length 200 -min_glue 2
//                  -glue_factor 0.05 -min_iso_ratio 0.05 -t CPU -k 24 -kk 48 [-strand]
//       Trinity:2180-2196 (k = K-1, kk = 2(K-1): Trinity:1025-1026); the output
//       is sorted by column 9 (general numeric, descending: Trinity:2199-2201)
//       and read by annotate_chrysalis_welds_with_iworm_names.pl as
//       "idx_A ptr idx_B rest..." with idx = contig index in file order.
//
// Chrysalis' own source is an empty submodule in the reference tree; this is
// the build's reading of the weld step, with the weld k-mer count -- the
// data-parallel part, every read window against every candidate weld -- on
// the GPU (tk_kset_count):
//   1. every k-mer (k = -k, all ACGT) of every contig of length >=
//      -min_contig_length, canonical unless -strand;
//   2. for each k-mer found in two different contigs A and B (in the
//      orientation where both read it forward), two candidate welds of
//      length kk: A's (kk-k)/2 bases before it + the k-mer + B's (kk-k)/2
//      bases after it, and B's before + k-mer + A's after (flanks inside the
//      contigs and all ACGT);
//   3. the candidates' occurrences in the reads (-r), canonical unless -strand;
//   4. A and B are welded when one of their candidates occurs in >= -min_glue
//      reads' windows.  One line per welded pair (A < B), with the pair's best
//      weld:
//        A -> B weld: <kk-mer> k-mer: <k-mer> support: <count>
//      column 9 = the support.
// -glue_factor and -min_iso_ratio (coverage ratios between the glued
// contigs) are accepted and not applied; -scaffolding and -debug_weld_all are
// accepted and ignored.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <unordered_map>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

namespace {

int code_of(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

std::string rc_str(const std::string& s) {
  std::string r(s.rbegin(), s.rend());
  for (char& c : r) {
    switch (c) {
      case 'A': c = 'T'; break;
      case 'C': c = 'G'; break;
      case 'G': c = 'C'; break;
      case 'T': c = 'A'; break;
      default: break;
    }
  }
  return r;
}

bool all_acgt(const std::string& s, size_t a, size_t n) {
  if (a + n > s.size()) return false;
  for (size_t i = a; i < a + n; i++)
    if (code_of(s[i]) < 0) return false;
  return true;
}

unsigned __int128 value_of(const std::string& s) {
  unsigned __int128 v = 0;
  for (char c : s) v = (v << 2) | (unsigned __int128)code_of(c);
  return v;
}

struct Occ {
  uint32_t contig;
  uint32_t pos;   // start of the k-mer in the oriented contig
  bool rc;        // the contig is read reverse-complemented to see the canonical k-mer forward
};

}  // namespace

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string iworm, reads_path;
  if (!a.value({"-i"}, iworm) || !a.value({"-r"}, reads_path))
    die("usage: GraphFromFasta -i <inchworm.fa> -r <reads.fa> [-k 24] [-kk 48] [-min_glue 2] [-strand] ...");
  const int k = (int)a.integer({"-k"}, 24);
  const int kk = (int)a.integer({"-kk"}, 2 * k);
  const long long min_glue = a.integer({"-min_glue"}, 2);
  const long long min_len = a.integer({"-min_contig_length"}, 1);
  const bool strand = a.flag({"-strand"});
  a.integer({"-t"}, 1);
  a.real({"-glue_factor"}, 0.05);
  a.real({"-min_iso_ratio"}, 0.05);
  std::string ignored;
  a.value({"-scaffolding"}, ignored);
  a.flag({"-debug_weld_all"});
  a.flag({"-report_welds"});
  if (!a.unknown_options().empty()) die("GraphFromFasta: unknown option %s", a.unknown_options()[0].c_str());
  if (k < 1 || k > 32 || kk <= k || kk > 63 || (kk - k) % 2) die("GraphFromFasta: need 1 <= k <= 32, k < kk <= 63, kk - k even");
  const int flank = (kk - k) / 2;
  try {
    SeqSet contigs = read_all(iworm, false);
    const uint64_t nc = contigs.size();
    std::vector<std::string> fwd(nc), rev(nc);
    for (uint64_t i = 0; i < nc; i++) {
      fwd[i] = contigs.seq.substr(contigs.off[i], contigs.off[i + 1] - contigs.off[i]);
      for (char& c : fwd[i]) c = (char)toupper((unsigned char)c);
      rev[i] = rc_str(fwd[i]);
    }
    // 1. k-mer -> occurrences
    std::unordered_map<uint64_t, std::vector<Occ>> occ;
    const uint64_t kmask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    for (uint64_t i = 0; i < nc; i++) {
      const std::string& s = fwd[i];
      if ((long long)s.size() < min_len) continue;
      uint64_t f = 0, r = 0;
      int run = 0;
      for (size_t p = 0; p < s.size(); p++) {
        const int c = code_of(s[p]);
        if (c < 0) {
          run = 0;
          continue;
        }
        run++;
        f = ((f << 2) | (uint64_t)c) & kmask;
        r = (r >> 2) | ((uint64_t)(3 - c) << (2 * k - 2));
        if (run < k) continue;
        const uint32_t start = (uint32_t)(p + 1 - k);
        if (strand || f <= r) occ[f].push_back({(uint32_t)i, start, false});
        else occ[r].push_back({(uint32_t)i, (uint32_t)(s.size() - start - k), true});
      }
    }
    // 2. candidate welds of contig pairs sharing a k-mer
    struct Cand {
      uint32_t a, b;
      std::string weld, kmer;
    };
    std::vector<Cand> cands;
    for (auto& kv : occ) {
      const auto& v = kv.second;
      if (v.size() < 2) continue;
      for (size_t x = 0; x < v.size(); x++)
        for (size_t y = 0; y < v.size(); y++) {
          if (v[x].contig == v[y].contig) continue;
          const std::string& sa = v[x].rc ? rev[v[x].contig] : fwd[v[x].contig];
          const std::string& sb = v[y].rc ? rev[v[y].contig] : fwd[v[y].contig];
          const uint32_t pa = v[x].pos, pb = v[y].pos;
          if (pa < (uint32_t)flank || pb + k + flank > sb.size()) continue;
          if (!all_acgt(sa, pa - flank, flank) || !all_acgt(sb, pb + k, flank)) continue;
          Cand c;
          c.a = std::min(v[x].contig, v[y].contig);
          c.b = std::max(v[x].contig, v[y].contig);
          c.kmer = sa.substr(pa, k);
          c.weld = sa.substr(pa - flank, flank) + c.kmer + sb.substr(pb + k, flank);
          if (!strand) {
            const std::string w2 = rc_str(c.weld);
            if (w2 < c.weld) c.weld = w2;
          }
          cands.push_back(std::move(c));
        }
    }
    // 3. distinct weld k-mers, counted in the reads on the device
    std::vector<std::string> welds;
    welds.reserve(cands.size());
    for (auto& c : cands) welds.push_back(c.weld);
    std::sort(welds.begin(), welds.end());
    welds.erase(std::unique(welds.begin(), welds.end()), welds.end());
    std::vector<uint64_t> keys(2 * welds.size());
    for (size_t i = 0; i < welds.size(); i++) {
      const unsigned __int128 v = value_of(welds[i]);
      keys[2 * i] = (uint64_t)v;
      keys[2 * i + 1] = (uint64_t)(v >> 64);
    }
    std::vector<uint64_t> counts(welds.size());
    tk_ctx* ctx = open_ctx();
    tk_kset* ks = nullptr;
    check(tk_kset_create(ctx, keys.data(), welds.size(), &ks), "tk_kset_create");
    SeqReader rr(reads_path, false);
    SeqSet batch;
    std::vector<uint64_t> part(welds.size());
    while (true) {
      batch.clear();
      if (!rr.next_batch(batch, 50000000, (uint64_t)4 << 30)) break;
      tk_reads* r = nullptr;
      check(tk_reads_from_host(ctx, batch.seq.data(), batch.off.data(), batch.size(), &r), "tk_reads_from_host");
      check(tk_kset_count(ctx, ks, r, kk, strand ? 0 : 1, nullptr, part.data()), "tk_kset_count");
      tk_reads_free(r);
      for (size_t i = 0; i < part.size(); i++) counts[i] += part[i];
    }
    tk_kset_free(ks);
    tk_ctx_free(ctx);
    // 4. welded pairs: best weld per pair (support, then weld, then k-mer)
    struct Best {
      uint64_t support = 0;
      std::string weld, kmer;
    };
    std::unordered_map<uint64_t, Best> best;
    for (auto& c : cands) {
      const size_t w = std::lower_bound(welds.begin(), welds.end(), c.weld) - welds.begin();
      const uint64_t sup = counts[w];
      if ((long long)sup < min_glue) continue;
      Best& b = best[((uint64_t)c.a << 32) | c.b];
      if (sup > b.support || (sup == b.support && (b.weld.empty() || c.weld < b.weld ||
                                                    (c.weld == b.weld && c.kmer < b.kmer)))) {
        b.support = sup;
        b.weld = c.weld;
        b.kmer = c.kmer;
      }
    }
    std::vector<uint64_t> pairs;
    for (auto& kv : best) pairs.push_back(kv.first);
    std::sort(pairs.begin(), pairs.end());
    Out w(stdout);
    char buf[64];
    for (uint64_t p : pairs) {
      const Best& b = best[p];
      int n = snprintf(buf, sizeof buf, "%u -> %u weld: ", (unsigned)(p >> 32), (unsigned)(p & 0xFFFFFFFFu));
      w.put(buf, (size_t)n);
      w.put(b.weld);
      w.put(" k-mer: ");
      w.put(b.kmer);
      n = snprintf(buf, sizeof buf, " support: %llu\n", (unsigned long long)b.support);
      w.put(buf, (size_t)n);
    }
    w.flush();
    fprintf(stderr, "-GraphFromFasta: %llu contigs, %zu candidate welds (%zu distinct), %zu welded pairs\n",
            (unsigned long long)nc, cands.size(), welds.size(), pairs.size());
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
