// QuantifyGraph -- read support of each de Bruijn node of a component.
//
//   QuantifyGraph -g c<id>.graph.tmp -i c<id>.reads.tmp -o c<id>.graph.out
//                 -max_reads N -k 24 [-strand] [-no_cleanup]   (Trinity:2348-2356)
//   QuantifyGraph --batch component_base_listing.txt -k 24 [-strand]
//       every component the listing names (partition_chrysalis_graphs_n_reads.pl:
//       "<component>\t<base>" lines, :127-140) in one pass: <base>.graph.tmp +
//       <base>.reads.tmp -> <base>.graph.out.  This replaces one process per
//       component under ParaFly (Trinity:2358-2410) with one table for all.
//
// Inputs: the graph as FastaToDeBruijn writes it ("Component <id>" then
// "<node>\t<k-mer>\t<predecessors>"), the component's reads as the partition
// script writes them ("<acc> <pct>" line, then the sequence, :103-113).
// Output: the graph with a fourth column, the number of windows of the
// component's reads (the first -max_reads of them) equal to the node's
// k-mer (canonical unless -strand; windows with a non-ACGT base do not
// count).  The count runs on the GPU through tk_kset_count with the
// component as each key's tag.  Chrysalis' own QuantifyGraph is an empty
// submodule in the reference tree: its weighting of edges and its output
// columns are not restated.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

namespace {

struct Graph {
  std::string header;                 // "Component <id>"
  std::vector<std::string> lines;     // node lines without newline
  std::vector<uint64_t> keys;         // node k-mers
};

Graph read_graph(const std::string& path, int k) {
  Graph g;
  TextSource src(path);
  const char* p = src.data();
  const uint64_t n = src.size();
  uint64_t pos = 0;
  bool first = true;
  while (pos < n) {
    const char* e = (const char*)memchr(p + pos, '\n', n - pos);
    const uint64_t end = e ? (uint64_t)(e - p) : n;
    std::string line(p + pos, end - pos);
    pos = e ? end + 1 : n;
    if (first) {
      if (line.compare(0, 10, "Component ") != 0) die("%s: no Component header", path.c_str());
      g.header = line;
      first = false;
      continue;
    }
    if (line.empty()) continue;
    const size_t t1 = line.find('\t');
    const size_t t2 = t1 == std::string::npos ? t1 : line.find('\t', t1 + 1);
    if (t1 == std::string::npos) die("%s: bad node line", path.c_str());
    const std::string km = line.substr(t1 + 1, (t2 == std::string::npos ? line.size() : t2) - t1 - 1);
    if ((int)km.size() != k) die("%s: node k-mer of length %zu, -k %d", path.c_str(), km.size(), k);
    bool ok = false;
    g.keys.push_back(kmer_from_text(km.c_str(), k, &ok));
    if (!ok) die("%s: node k-mer %s is not ACGT", path.c_str(), km.c_str());
    g.lines.push_back(line);
  }
  if (first) die("%s: empty graph file", path.c_str());
  return g;
}

// Reads of a component: sequence lines after each "<acc> <pct>" line.
void read_component_reads(const std::string& path, uint64_t max_reads, std::string& seq,
                          std::vector<uint64_t>& off, std::vector<uint32_t>& tags, uint32_t tag) {
  TextSource src(path);
  const char* p = src.data();
  const uint64_t n = src.size();
  uint64_t pos = 0, line_no = 0, got = 0;
  while (pos < n && got < max_reads) {
    const char* e = (const char*)memchr(p + pos, '\n', n - pos);
    const uint64_t end = e ? (uint64_t)(e - p) : n;
    if (line_no & 1) {
      seq.append(p + pos, end - pos);
      off.push_back(seq.size());
      tags.push_back(tag);
      got++;
    }
    line_no++;
    pos = e ? end + 1 : n;
  }
}

}  // namespace

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string gpath, rpath, opath, listing;
  a.value({"-g"}, gpath);
  a.value({"-i"}, rpath);
  a.value({"-o"}, opath);
  a.value({"--batch"}, listing);
  const int k = (int)a.integer({"-k"}, 24);
  const long long max_reads = a.integer({"-max_reads"}, -1);
  const bool strand = a.flag({"-strand"});
  a.flag({"-no_cleanup"});
  if (!a.unknown_options().empty()) die("QuantifyGraph: unknown option %s", a.unknown_options()[0].c_str());
  if (k < 1 || k > 31) die("QuantifyGraph: -k must be in [1,31]");
  struct Job {
    std::string graph, reads, out;
  };
  std::vector<Job> jobs;
  if (!listing.empty()) {
    TextSource src(listing);
    std::string all(src.data(), src.size());
    size_t pos = 0;
    while (pos < all.size()) {
      size_t e = all.find('\n', pos);
      if (e == std::string::npos) e = all.size();
      std::string line = all.substr(pos, e - pos);
      pos = e + 1;
      const size_t t = line.find('\t');
      if (t == std::string::npos) continue;
      const std::string base = line.substr(t + 1);
      jobs.push_back({base + ".graph.tmp", base + ".reads.tmp", base + ".graph.out"});
    }
  } else {
    if (gpath.empty() || rpath.empty() || opath.empty())
      die("usage: QuantifyGraph -g <graph> -i <reads> -o <out> -k 24 [-max_reads N] [-strand] | --batch <listing>");
    jobs.push_back({gpath, rpath, opath});
  }
  try {
    std::vector<Graph> graphs;
    std::vector<uint64_t> keys;
    std::string seq;
    std::vector<uint64_t> off{0};
    std::vector<uint32_t> tags;
    for (uint32_t j = 0; j < jobs.size(); j++) {
      graphs.push_back(read_graph(jobs[j].graph, k));
      for (uint64_t key : graphs.back().keys) {
        keys.push_back(key);
        keys.push_back(j);
      }
      read_component_reads(jobs[j].reads, max_reads < 0 ? UINT64_MAX : (uint64_t)max_reads, seq, off, tags, j);
    }
    const uint64_t nk = keys.size() / 2;
    std::vector<uint64_t> counts(nk);
    tk_ctx* ctx = open_ctx();
    tk_kset* ks = nullptr;
    check(tk_kset_create(ctx, keys.data(), nk, &ks), "tk_kset_create");
    tk_reads* r = nullptr;
    check(tk_reads_from_host(ctx, seq.data(), off.data(), tags.size(), &r), "tk_reads_from_host");
    check(tk_kset_count(ctx, ks, r, k, strand ? 0 : 1, tags.data(), counts.data()), "tk_kset_count");
    tk_reads_free(r);
    tk_kset_free(ks);
    tk_ctx_free(ctx);
    uint64_t at = 0;
    for (uint32_t j = 0; j < jobs.size(); j++) {
      const Graph& g = graphs[j];
      const std::string tmp = jobs[j].out + ".tmp";
      FILE* f = fopen(tmp.c_str(), "w");
      if (!f) die("cannot write %s", tmp.c_str());
      {
        Out w(f);
        w.put(g.header);
        w.putc('\n');
        for (size_t i = 0; i < g.lines.size(); i++) {
          w.put(g.lines[i]);
          w.putc('\t');
          w.put_u64(counts[at + i]);
          w.putc('\n');
        }
      }
      fclose(f);
      if (rename(tmp.c_str(), jobs[j].out.c_str()) != 0) die("cannot rename %s", tmp.c_str());
      at += g.lines.size();
    }
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
