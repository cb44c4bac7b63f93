// ReadsToTranscripts -- Chrysalis' read-to-component assignment, on the GPU.
//
//   ReadsToTranscripts -i both.fa -f bundled_iworm_contigs.fasta
//        -o readsToComponents.out -t CPU -max_mem_reads 50000000 [-strand] [-p P]
//                                                       (Trinity:2238-2250)
//
// Bundles are ">s_<comp>" records whose sequence is the component's iworm
// contigs joined by 'X' (add_LR_reads_to_iworm_bundle.pl:26,34-42); a window
// holding X (or any non-ACGT base) is skipped.  Each read (upper-cased,
// Changelog.txt:467) votes with its k-mers; the component with most hits
// wins (ties: smallest id), pct = 100*hits/(len-k+1), and the read is written
// when pct >= P (Changelog.txt:871-873) as "comp\t>acc\tpct\tSEQ"
// (Trinity:2474,2493; partition_chrysalis_graphs_n_reads.pl:103,113).  Reads
// are streamed in chunks of -max_mem_reads.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string in, bundles, outp;
  a.value({"-i"}, in);
  a.value({"-f"}, bundles);
  a.value({"-o"}, outp);
  int threads = (int)a.integer({"-t"}, 1);
  const long long max_mem_reads = a.integer({"-max_mem_reads"}, 50000000);
  const bool strand = a.flag({"-strand"});
  const long long min_pct = a.integer({"-p"}, 0);
  const int k = (int)a.integer({"-k"}, 25);
  auto unk = a.unknown_options();
  if (!unk.empty()) die("ReadsToTranscripts: unknown option %s", unk[0].c_str());
  if (in.empty() || bundles.empty() || outp.empty()) die("ReadsToTranscripts: need -i, -f and -o");
  if (k < 1 || k > 31) die("-k must be in [1,31]");
  if (threads < 1) threads = 1;
  try {
    tk_ctx* ctx = open_ctx();
    SeqSet b = read_all(bundles, true);
    std::vector<uint32_t> comp(b.size());
    for (uint64_t i = 0; i < b.size(); i++) {
      const std::string nm = b.name(i);
      char* end = nullptr;
      if (nm.compare(0, 2, "s_") != 0) die("bundle header %s: cannot decipher component id", nm.c_str());
      unsigned long v = strtoul(nm.c_str() + 2, &end, 10);
      if (end == nm.c_str() + 2 || v >= 0xFFFFFFFFul) die("bundle header %s: cannot decipher component id", nm.c_str());
      comp[i] = (uint32_t)v;
    }
    tk_reads* br = nullptr;
    check(tk_reads_from_host(ctx, b.seq.data(), b.off.data(), b.size(), &br), "tk_reads_from_host");
    tk_ctable* ct = nullptr;
    check(tk_ctable_build(ctx, k, strand ? 0 : 1, br, comp.data(), &ct), "tk_ctable_build");
    tk_reads_free(br);
    b = SeqSet();

    std::string tmp = outp + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) die("cannot write %s", tmp.c_str());
    Out out(f);
    SeqReader rr(in, true);
    SeqSet s;
    std::vector<tk_r2t_vote> v;
    std::vector<std::string> parts(threads);
    const uint64_t chunk = std::max<long long>(1, std::min<long long>(max_mem_reads, 1ll << 24));
    while (true) {
      s.clear();
      if (!rr.next_batch(s, chunk, 4ull << 30)) break;
      tk_reads* r = nullptr;
      check(tk_reads_from_host(ctx, s.seq.data(), s.off.data(), s.size(), &r), "tk_reads_from_host");
      v.resize(s.size());
      check(tk_r2t_vote_reads(ctx, ct, r, v.data()), "tk_r2t_vote_reads");
      tk_reads_free(r);
      const uint64_t n = s.size();
      std::vector<std::thread> th;
      for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t] {
          std::string& o = parts[t];
          o.clear();
          char buf[64];
          for (uint64_t i = n * t / threads; i < n * (t + 1) / threads; i++) {
            const tk_r2t_vote& x = v[i];
            if (!x.hits || !x.nwin) continue;
            if ((uint64_t)100 * x.hits < (uint64_t)min_pct * x.nwin) continue;
            int m = snprintf(buf, sizeof buf, "%u\t>", x.comp);
            o.append(buf, m);
            o += s.name(i);
            m = snprintf(buf, sizeof buf, "\t%g\t", 100.0 * x.hits / x.nwin);
            o.append(buf, m);
            for (uint64_t j = s.off[i]; j < s.off[i + 1]; j++) o.push_back((char)toupper((unsigned char)s.seq[j]));
            o.push_back('\n');
          }
        });
      }
      for (auto& x : th) x.join();
      for (auto& p : parts) out.put(p);
    }
    out.flush();
    if (fclose(f) != 0) die("write failed");
    if (rename(tmp.c_str(), outp.c_str()) != 0) die("cannot rename %s", tmp.c_str());
    tk_ctable_free(ct);
    tk_ctx_free(ctx);
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
