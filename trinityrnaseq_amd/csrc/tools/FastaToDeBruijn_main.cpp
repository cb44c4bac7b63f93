// FastaToDeBruijn -- per-component de Bruijn graphs of the Inchworm bundles.
//
//   FastaToDeBruijn --fasta bundled_iworm_contigs.fasta -K 24 --graph_per_record
//                   --threads CPU [--SS]                    (Trinity:2296-2303)
//
// Output (stdout), per record (">s_<component>", its Inchworm contigs joined
// by 'X': add_LR_reads_to_iworm_bundle.pl:26,34-42):
//   Component <component>
//   <node>\t<k-mer>\t<predecessors>
// one line per distinct K-mer of the record (canonical unless --SS), nodes
// numbered from 0 in order of first appearance; <predecessors> lists, comma
// separated and ascending, the nodes that precede the k-mer somewhere in the
// record's sequences (-1: none).  partition_chrysalis_graphs_n_reads.pl:55-75
// reads "Component (\d+)" and counts one line per k-mer.  Inchworm's own
// FastaToDeBruijn is an empty submodule in the reference tree: the columns
// after the header are this build's layout.  The graph is host work (no
// table pass); QuantifyGraph weights its nodes on the GPU.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string fasta;
  if (!a.value({"--fasta"}, fasta)) die("usage: FastaToDeBruijn --fasta <bundles.fasta> -K 24 --graph_per_record [--SS]");
  const int k = (int)a.integer({"-K"}, 24);
  const bool ss = a.flag({"--SS"});
  const bool per_record = a.flag({"--graph_per_record"});
  a.integer({"--threads"}, 1);
  if (!a.unknown_options().empty()) die("FastaToDeBruijn: unknown option %s", a.unknown_options()[0].c_str());
  if (!per_record) die("FastaToDeBruijn: only --graph_per_record is supported");
  if (k < 1 || k > 31) die("FastaToDeBruijn: -K must be in [1,31]");
  try {
    SeqSet recs = read_all(fasta, true);
    Out w(stdout);
    const uint64_t kmask = (1ull << (2 * k)) - 1;
    std::vector<char> txt(k);
    for (uint64_t r = 0; r < recs.size(); r++) {
      const std::string name = recs.name(r);
      std::string comp = name;
      if (name.size() > 2 && name[0] == 's' && name[1] == '_') comp = name.substr(2);
      std::unordered_map<uint64_t, uint32_t> id;
      std::vector<uint64_t> nodes;
      std::vector<std::set<uint32_t>> preds;
      uint64_t f = 0, rv = 0;
      int run = 0;
      int64_t prev = -1;
      for (uint64_t p = recs.off[r]; p < recs.off[r + 1]; p++) {
        int c;
        switch (recs.seq[p]) {
          case 'A': case 'a': c = 0; break;
          case 'C': case 'c': c = 1; break;
          case 'G': case 'g': c = 2; break;
          case 'T': case 't': c = 3; break;
          default: c = -1;
        }
        if (c < 0) {
          run = 0;
          prev = -1;
          continue;
        }
        run++;
        f = ((f << 2) | (uint64_t)c) & kmask;
        rv = (rv >> 2) | ((uint64_t)(3 - c) << (2 * k - 2));
        if (run < k) continue;
        const uint64_t key = (!ss && rv < f) ? rv : f;
        auto it = id.find(key);
        uint32_t n;
        if (it == id.end()) {
          n = (uint32_t)nodes.size();
          id.emplace(key, n);
          nodes.push_back(key);
          preds.emplace_back();
        } else {
          n = it->second;
        }
        if (prev >= 0) preds[n].insert((uint32_t)prev);
        prev = n;
      }
      w.put("Component ");
      w.put(comp);
      w.putc('\n');
      for (uint32_t n = 0; n < nodes.size(); n++) {
        w.put_u64(n);
        w.putc('\t');
        kmer_to_text(nodes[n], k, txt.data());
        w.put(txt.data(), k);
        w.putc('\t');
        if (preds[n].empty()) {
          w.put("-1");
        } else {
          bool first = true;
          for (uint32_t q : preds[n]) {
            if (!first) w.putc(',');
            w.put_u64(q);
            first = false;
          }
        }
        w.putc('\n');
      }
    }
    w.flush();
  } catch (const std::exception& e) {
    die("%s", e.what());
  }
  return 0;
}
