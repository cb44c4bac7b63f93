// jellyfish -- drop-in for the jellyfish 2.x subcommands Trinity runs.
//
//   jellyfish count -t N -m K -s S [-o db] [--canonical|-C] reads.fa...
//        Trinity:2612-2619, util/insilico_read_normalization.pl:617-622
//   jellyfish dump [-L min] [-U max] [-c] [-t] [-o out] db
//        Trinity:2625, util/insilico_read_normalization.pl:641
//   jellyfish histo [-t N] [-l low] [-h high] [-i inc] [-f] [-o out] db
//        Trinity:2630, util/insilico_read_normalization.pl:636
//   jellyfish --version   (Trinity:4070-4074 requires "name 2.x")
//
// Counting runs on the GPU (tk_count); the database file holds the sorted
// (k-mer, count) pairs, so dump and histo are host passes over it.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

static int usage() {
  fprintf(stderr,
          "Usage: jellyfish <cmd> [options] arg...\n"
          "  count -m K [-s size] [-t threads] [-o db] [-C|--canonical] file...\n"
          "  dump  [-L min] [-U max] [-c] [-t] [-o out] db\n"
          "  histo [-l low] [-h high] [-i inc] [-f] [-t threads] [-o out] db\n"
          "  --version\n");
  return 1;
}

static void reject_unknown(const Args& a, const char* cmd) {
  auto u = a.unknown_options();
  if (!u.empty()) die("jellyfish %s: unknown option %s", cmd, u[0].c_str());
}

static int cmd_count(int argc, char** argv) {
  Args a(argc, argv, 2);
  const int k = (int)a.integer({"-m", "--mer-len"}, 0);
  const long long s_hint = a.integer({"-s", "--size"}, 0);
  a.integer({"-t", "--threads"}, 1);
  std::string out = "mer_counts.jf";
  a.value({"-o", "--output"}, out);
  const bool canonical = a.flag({"-C", "--canonical"});
  a.flag({"--disk"});
  reject_unknown(a, "count");
  auto files = a.positional();
  if (k < 1 || k > 31) die("jellyfish count: -m must be in [1,31] (got %d)", k);
  if (files.empty()) die("jellyfish count: no input file");

  tk_ctx* ctx = open_ctx();
  tk_table* table = nullptr;
  const uint64_t kBatchBases = 8ull << 30;
  for (const auto& f : files) {
    SeqReader r(f, false);
    SeqSet s;
    while (true) {
      s.clear();
      if (!r.next_batch(s, UINT64_MAX, kBatchBases)) break;
      tk_reads* reads = nullptr;
      check(tk_reads_from_host(ctx, s.seq.data(), s.off.data(), s.size(), &reads), "tk_reads_from_host");
      if (!table) {
        double est = 0;
        check(tk_estimate_distinct(ctx, reads, k, canonical ? 1 : 0, &est), "tk_estimate_distinct");
        // scale the first batch's estimate by the input still to come; the
        // table grows by rehash if this falls short
        double scale = r.file_size() ? std::max(1.0, (double)r.file_size() / (2.0 * s.seq.size() + 1)) : 1.0;
        uint64_t expected = (uint64_t)(est * scale * 1.05) + 4096;
        if (s_hint > 0) expected = std::min<uint64_t>(expected, (uint64_t)s_hint * 4);
        check(tk_table_create(ctx, k, canonical ? 1 : 0, expected, 0.5, &table), "tk_table_create");
      }
      check(tk_count(ctx, table, reads), "tk_count");
      tk_reads_free(reads);
    }
  }
  if (!table) check(tk_table_create(ctx, k, canonical ? 1 : 0, 64, 0.5, &table), "tk_table_create");
  CountDb db;
  db.k = k;
  db.canonical = canonical ? 1 : 0;
  uint64_t n = 0;
  check(tk_table_dump_size(ctx, table, 1, 0xFFFFFFFFu, &n), "tk_table_dump_size");
  db.keys.resize(n);
  db.counts.resize(n);
  check(tk_table_dump(ctx, table, 1, 0xFFFFFFFFu, db.keys.data(), db.counts.data(), n, &n), "tk_table_dump");
  tk_table_free(table);
  tk_ctx_free(ctx);
  write_db(out, db);
  return 0;
}

static int cmd_dump(int argc, char** argv) {
  Args a(argc, argv, 2);
  const long long lo = a.integer({"-L", "--lower-count"}, 0);
  const long long hi = a.integer({"-U", "--upper-count"}, 0xFFFFFFFFll);
  const bool column = a.flag({"-c", "--column"});
  const bool tab = a.flag({"-t", "--tab"});
  std::string outp;
  a.value({"-o", "--output"}, outp);
  reject_unknown(a, "dump");
  auto pos = a.positional();
  if (pos.size() != 1) die("jellyfish dump: expected one database");
  CountDb db = read_db(pos[0]);
  FILE* f = outp.empty() ? stdout : fopen(outp.c_str(), "w");
  if (!f) die("cannot write %s", outp.c_str());
  Out o(f);
  std::vector<char> km(db.k);
  const char sep = tab ? '\t' : ' ';
  for (size_t i = 0; i < db.keys.size(); i++) {
    const uint32_t c = db.counts[i];
    if ((long long)c < lo || (long long)c > hi) continue;
    kmer_to_text(db.keys[i], db.k, km.data());
    if (column) {
      o.put(km.data(), db.k);
      o.putc(sep);
      o.put_u64(c);
      o.putc('\n');
    } else {
      o.putc('>');
      o.put_u64(c);
      o.putc('\n');
      o.put(km.data(), db.k);
      o.putc('\n');
    }
  }
  o.flush();
  if (f != stdout) fclose(f);
  return 0;
}

static int cmd_histo(int argc, char** argv) {
  Args a(argc, argv, 2);
  const long long low = a.integer({"-l", "--low"}, 1);
  const long long high = a.integer({"-h", "--high"}, 10000);
  const long long inc = a.integer({"-i", "--increment"}, 1);
  const bool full = a.flag({"-f", "--full"});
  a.integer({"-t", "--threads"}, 1);
  std::string outp;
  a.value({"-o", "--output"}, outp);
  reject_unknown(a, "histo");
  auto pos = a.positional();
  if (pos.size() != 1) die("jellyfish histo: expected one database");
  if (inc < 1 || high < low || low < 0) die("jellyfish histo: bad range");
  CountDb db = read_db(pos[0]);
  const uint64_t base = low > 1 ? ((uint64_t)inc >= (uint64_t)low ? 1 : low - inc) : 1;
  const uint64_t ceil_ = high + inc;
  const uint64_t nb = (ceil_ + inc - base) / inc;
  std::vector<uint64_t> bins(nb, 0);
  for (uint32_t c : db.counts) {
    if (c < base) bins[0]++;
    else if (c >= ceil_) bins[nb - 1]++;
    else bins[(c - base) / inc]++;
  }
  FILE* f = outp.empty() ? stdout : fopen(outp.c_str(), "w");
  if (!f) die("cannot write %s", outp.c_str());
  for (uint64_t i = 0; i < nb; i++)
    if (full || bins[i]) fprintf(f, "%llu %llu\n", (unsigned long long)(base + i * inc), (unsigned long long)bins[i]);
  if (f != stdout) fclose(f);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  const std::string cmd = argv[1];
  try {
    if (cmd == "--version" || cmd == "-V") {
      printf("jellyfish 2.3.0\n");
      return 0;
    }
    if (cmd == "count") return cmd_count(argc, argv);
    if (cmd == "dump") return cmd_dump(argc, argv);
    if (cmd == "histo") return cmd_histo(argc, argv);
    if (cmd == "--help" || cmd == "-h") { usage(); return 0; }
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  fprintf(stderr, "Unknown command '%s'\n", cmd.c_str());
  return usage();
}
