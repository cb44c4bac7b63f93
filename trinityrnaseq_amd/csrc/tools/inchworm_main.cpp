// inchworm -- k-mer dictionary load and prune of Inchworm, on the GPU.
//
//   inchworm --kmers <jellyfish dump> --run_inchworm -K 25 --monitor 1 --DS
//            --num_threads N --PARALLEL_IWORM --min_any_entropy 1.0 -L 25
//            --no_prune_error_kmers                      (Trinity:2654,2676-2711,1050)
//   inchworm --reads <fasta> --run_inchworm -K 25 [--minKmerCount m] --DS ...
//                                                       (Trinity:2669-2672)
//
// This front end builds the dictionary (from the dump, or by counting the
// reads), drops k-mers with count < --minKmerCount or base-composition
// entropy < --min_any_entropy, and writes inchworm.kmer_count (the number
// of k-mers kept; read back at Trinity:3189).  --dump_pruned FILE writes the
// kept dictionary as a jellyfish-style FASTA dump.  With --run_inchworm the
// greedy contigs (tk_inchworm) go to stdout, those of length >= -L and
// coverage >= --min_assembly_coverage, as
//   >a<N>;<cov> K: <k> length: <len>
//   <sequence>
// N counting the reported contigs from 1 in seed order and cov = the mean
// count of the contig's k-mers rounded down (scaffold_iworm_contigs.pl:67
// and filter_iworm_by_min_length_or_cov.pl:22-26 read a/N;cov).  Seeds need
// count >= --min_seed_coverage (2) and composition entropy >=
// --min_seed_entropy (1.5).  Inchworm's error-k-mer pruning (the default
// when --no_prune_error_kmers is absent) is not part of this build: the tool
// refuses to assemble without --no_prune_error_kmers.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

// keep[(a*(k+1)+c)*(k+1)+g]: entropy (bits) of the composition >= min_entropy.
static std::vector<uint8_t> entropy_keep(int k, double min_entropy) {
  const int K1 = k + 1;
  std::vector<uint8_t> keep((size_t)K1 * K1 * K1, 0);
  for (int a = 0; a <= k; a++)
    for (int c = 0; a + c <= k; c++)
      for (int g = 0; a + c + g <= k; g++) {
        const int n[4] = {a, c, g, k - a - c - g};
        double h = 0;
        for (int i = 0; i < 4; i++)
          if (n[i] > 0) {
            const double p = (double)n[i] / k;
            h -= p * log2(p);
          }
        keep[((size_t)a * K1 + c) * K1 + g] = h >= min_entropy;
      }
  return keep;
}

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string kmers, reads, dump_out;
  a.value({"--kmers"}, kmers);
  a.value({"--reads"}, reads);
  a.value({"--dump_pruned"}, dump_out);
  const int k = (int)a.integer({"-K"}, 25);
  const bool ds = a.flag({"--DS"});
  const double min_entropy = a.real({"--min_any_entropy"}, 0.0);
  const long long min_count = a.integer({"--minKmerCount"}, 1);
  const bool run = a.flag({"--run_inchworm"});
  const long long min_len = a.integer({"-L"}, 25);
  const long long min_asm_cov = a.integer({"--min_assembly_coverage"}, 1);
  const double min_seed_entropy = a.real({"--min_seed_entropy"}, 1.5);
  const long long min_seed_cov = a.integer({"--min_seed_coverage"}, 2);
  const bool no_error_prune = a.flag({"--no_prune_error_kmers"});
  // accepted for command-line compatibility (threads, logging, scratch files)
  a.integer({"--num_threads"}, 1);
  a.integer({"--monitor"}, 0);
  a.flag({"--PARALLEL_IWORM"});
  a.flag({"--keep_tmp_files"});
  if (a.flag({"--make_super_reads"})) die("inchworm: --make_super_reads is not part of this build");
  auto unk = a.unknown_options();
  if (!unk.empty()) die("inchworm: unknown option %s", unk[0].c_str());
  if (kmers.empty() == reads.empty()) die("inchworm: need exactly one of --kmers / --reads");
  if (k < 1 || k > 31) die("inchworm: -K must be in [1,31]");
  if (run && !no_error_prune)
    die("inchworm: error-k-mer pruning is not part of this build; run with --no_prune_error_kmers");

  try {
    tk_ctx* ctx = open_ctx();
    tk_table* table = nullptr;
    if (!kmers.empty()) {
      std::vector<uint64_t> keys;
      std::vector<uint32_t> counts;
      read_kmer_dump(kmers, k, ds, keys, counts);
      check(tk_table_create(ctx, k, ds ? 1 : 0, keys.size() + 64, 0.5, &table), "tk_table_create");
      check(tk_table_load(ctx, table, keys.data(), counts.data(), keys.size()), "tk_table_load");
    } else {
      SeqSet s = read_all(reads, false);
      tk_reads* r = nullptr;
      check(tk_reads_from_host(ctx, s.seq.data(), s.off.data(), s.size(), &r), "tk_reads_from_host");
      double est = 0;
      check(tk_estimate_distinct(ctx, r, k, ds ? 1 : 0, &est), "tk_estimate_distinct");
      check(tk_table_create(ctx, k, ds ? 1 : 0, (uint64_t)(est * 1.05) + 4096, 0.5, &table), "tk_table_create");
      check(tk_count(ctx, table, r), "tk_count");
      tk_reads_free(r);
    }
    std::vector<uint8_t> keep = entropy_keep(k, min_entropy);
    uint64_t kept = 0;
    check(tk_table_prune(ctx, table, min_count < 1 ? 1 : (uint32_t)min_count,
                         min_entropy > 0 ? keep.data() : nullptr, &kept),
          "tk_table_prune");
    FILE* f = fopen("inchworm.kmer_count", "w");
    if (!f) die("cannot write inchworm.kmer_count");
    fprintf(f, "%llu\n", (unsigned long long)kept);
    fclose(f);
    fprintf(stderr, "-kmers kept after pruning: %llu\n", (unsigned long long)kept);
    if (!dump_out.empty()) {
      std::vector<uint64_t> keys(kept);
      std::vector<uint32_t> counts(kept);
      uint64_t n = 0;
      check(tk_table_dump(ctx, table, 1, 0xFFFFFFFFu, keys.data(), counts.data(), kept, &n), "tk_table_dump");
      std::string tmp = dump_out + ".tmp";
      FILE* o = fopen(tmp.c_str(), "w");
      if (!o) die("cannot write %s", tmp.c_str());
      {
        Out w(o);
        std::vector<char> km(k);
        for (uint64_t i = 0; i < n; i++) {
          w.putc('>');
          w.put_u64(counts[i]);
          w.putc('\n');
          kmer_to_text(keys[i], k, km.data());
          w.put(km.data(), k);
          w.putc('\n');
        }
      }
      fclose(o);
      if (rename(tmp.c_str(), dump_out.c_str()) != 0) die("cannot rename %s", tmp.c_str());
    }
    if (run) {
      std::vector<uint8_t> seed_keep = entropy_keep(k, min_seed_entropy);
      tk_contigs* cg = nullptr;
      check(tk_inchworm(ctx, table, min_seed_cov < 1 ? 1 : (uint32_t)min_seed_cov,
                        min_seed_entropy > 0 ? seed_keep.data() : nullptr, &cg),
            "tk_inchworm");
      uint64_t n = 0, total = 0, rounds = 0, walks = 0;
      check(tk_contigs_info(cg, &n, &total, &rounds, &walks), "tk_contigs_info");
      std::string seq(total, '\0');
      std::vector<uint64_t> off(n + 1), sums(n), seeds(n);
      check(tk_contigs_get(cg, &seq[0], off.data(), sums.data(), seeds.data()), "tk_contigs_get");
      tk_contigs_free(cg);
      Out w(stdout);
      uint64_t reported = 0;
      char hdr[128];
      for (uint64_t i = 0; i < n; i++) {
        const uint64_t L = off[i + 1] - off[i];
        const uint64_t cov = sums[i] / (L - (uint64_t)k + 1);
        if ((long long)L < min_len || (long long)cov < min_asm_cov) continue;
        reported++;
        const int hl = snprintf(hdr, sizeof hdr, ">a%llu;%llu K: %d length: %llu\n", (unsigned long long)reported,
                                (unsigned long long)cov, k, (unsigned long long)L);
        w.put(hdr, (size_t)hl);
        w.put(seq.data() + off[i], L);
        w.putc('\n');
      }
      w.flush();
      fprintf(stderr, "-inchworm: %llu contigs (%llu reported), %llu rounds, %llu walks\n", (unsigned long long)n,
              (unsigned long long)reported, (unsigned long long)rounds, (unsigned long long)walks);
    }
    tk_table_free(table);
    tk_ctx_free(ctx);
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
