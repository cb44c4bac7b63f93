// fastaToKmerCoverageStats -- drop-in for Inchworm's per-read k-mer coverage tool.
//
//   fastaToKmerCoverageStats --reads X.fa --kmers <jellyfish dump> --kmer_size 25
//                            --num_threads N [--DS]
//       util/insilico_read_normalization.pl:837-845
//   fastaToKmerCoverageStats --kmers_from_reads Y.fa --reads X.fa --num_threads 1
//       Trinity:3930
//
// Output (stdout): a header line then one tab-separated row per read, in read
// order: acc, median_cov, mean_cov, stdev, pct_stdev_of_avg.  The consumers
// read the columns by header name (nbkc_normalize.pl:85-97,
// nbkc_merge_left_right_stats.pl:94-103) and require a fixed field count
// (DelimParser.pm:157-164).  K-mers missing from the dictionary count 1
// (Changelog.txt:537).  Conventions the reference tree does not fix are
// selectable: TK_STATS_MEDIAN=upper|lower|mean_floor (even-length median),
// TK_STATS_NWIN=missing|skip (windows holding N).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <thread>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

static int env_policy(const char* name, const char* const* opts, int n, int dflt) {
  const char* e = getenv(name);
  if (!e) return dflt;
  for (int i = 0; i < n; i++)
    if (!strcmp(e, opts[i])) return i;
  die("%s=%s is not one of the accepted values", name, e);
}

// One stats row from the exact integer aggregates: mean = sum/n, population
// stdev = sqrt(n*sumsq - sum^2)/n, printed like C++ ostream defaults (%g).
static void format_row(std::string& o, const std::string& acc, const tk_read_stat& st) {
  double mean = 0, sd = 0, pct = 0;
  if (st.nwin) {
    const unsigned __int128 sq = ((unsigned __int128)st.sumsq_hi << 64) | st.sumsq_lo;
    const unsigned __int128 num = (unsigned __int128)st.nwin * sq - (unsigned __int128)st.sum * st.sum;
    mean = (double)st.sum / (double)st.nwin;
    sd = sqrt((double)num) / (double)st.nwin;
    pct = mean > 0 ? sd / mean * 100.0 : 0.0;
  }
  char buf[160];
  int n = snprintf(buf, sizeof buf, "\t%u\t%g\t%g\t%g\n", st.median, mean, sd, pct);
  o += acc;
  o.append(buf, n);
}

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  if (a.flag({"-h", "--help"})) {
    fprintf(stderr,
            "usage: fastaToKmerCoverageStats --reads <fasta> (--kmers <dump> | --kmers_from_reads <fasta>)\n"
            "       [--kmer_size 25] [--num_threads N] [--DS] [--capture_coverage_info]\n");
    return 0;
  }
  std::string reads_path, kmers_path, kfr_path;
  a.value({"--reads"}, reads_path);
  a.value({"--kmers"}, kmers_path);
  a.value({"--kmers_from_reads"}, kfr_path);
  const int k = (int)a.integer({"--kmer_size", "-K"}, 25);
  int threads = (int)a.integer({"--num_threads"}, 1);
  const bool ds = a.flag({"--DS"});
  const bool capture = a.flag({"--capture_coverage_info"});
  auto unk = a.unknown_options();
  if (!unk.empty()) die("unknown option %s", unk[0].c_str());
  if (reads_path.empty() || (kmers_path.empty() == kfr_path.empty()))
    die("need --reads and exactly one of --kmers / --kmers_from_reads");
  if (capture) die("--capture_coverage_info (per-window coverage listing) is not provided by this build");
  if (k < 1 || k > 31) die("--kmer_size must be in [1,31]");
  if (threads < 1) threads = 1;

  static const char* const kMed[] = {"upper", "lower", "mean_floor"};
  static const char* const kNwin[] = {"missing", "skip"};
  tk_stats_opts opts;
  opts.min_count = 1;
  opts.missing_value = 1;
  opts.median_policy = env_policy("TK_STATS_MEDIAN", kMed, 3, TK_MEDIAN_UPPER);
  opts.nwin_policy = env_policy("TK_STATS_NWIN", kNwin, 2, TK_NWIN_AS_MISSING);

  try {
    tk_ctx* ctx = open_ctx();
    tk_table* table = nullptr;
    if (!kmers_path.empty()) {
      std::vector<uint64_t> keys;
      std::vector<uint32_t> counts;
      read_kmer_dump(kmers_path, k, ds, keys, counts);
      check(tk_table_create(ctx, k, ds ? 1 : 0, keys.size() + 64, 0.5, &table), "tk_table_create");
      check(tk_table_load(ctx, table, keys.data(), counts.data(), keys.size()), "tk_table_load");
    } else {
      SeqSet s = read_all(kfr_path, false);
      tk_reads* r = nullptr;
      check(tk_reads_from_host(ctx, s.seq.data(), s.off.data(), s.size(), &r), "tk_reads_from_host");
      double est = 0;
      check(tk_estimate_distinct(ctx, r, k, ds ? 1 : 0, &est), "tk_estimate_distinct");
      check(tk_table_create(ctx, k, ds ? 1 : 0, (uint64_t)(est * 1.05) + 4096, 0.5, &table), "tk_table_create");
      check(tk_count(ctx, table, r), "tk_count");
      tk_reads_free(r);
    }

    Out out(stdout);
    out.put(std::string("acc\tmedian_cov\tmean_cov\tstdev\tpct_stdev_of_avg\n"));
    SeqReader rr(reads_path, true);
    SeqSet s;
    std::vector<tk_read_stat> st;
    std::vector<std::string> parts(threads);
    while (true) {
      s.clear();
      if (!rr.next_batch(s, 1ull << 24, 4ull << 30)) break;
      tk_reads* r = nullptr;
      check(tk_reads_from_host(ctx, s.seq.data(), s.off.data(), s.size(), &r), "tk_reads_from_host");
      st.resize(s.size());
      check(tk_read_cov_stats(ctx, table, r, &opts, st.data()), "tk_read_cov_stats");
      tk_reads_free(r);
      const uint64_t n = s.size();
      std::vector<std::thread> th;
      for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t] {
          std::string& o = parts[t];
          o.clear();
          const uint64_t a0 = n * t / threads, a1 = n * (t + 1) / threads;
          for (uint64_t i = a0; i < a1; i++) format_row(o, s.name(i), st[i]);
        });
      }
      for (auto& x : th) x.join();
      for (auto& p : parts) out.put(p);
    }
    out.flush();
    tk_table_free(table);
    tk_ctx_free(ctx);
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
