// extract_normalized_reads -- write the reads the normalization selected.
//
//   extract_normalized_reads --seqType fq|fa --selected <accs> --output <out>
//                            <source reads>...
//       = make_normalized_reads_file (util/insilico_read_normalization.pl:505-583)
//         with build_selected_index (:470-502), run once per output file
//         (left.norm.fq from the left sources, right.norm.fq from the right).
//
// The selected file holds one core accession per line (nbkc_normalize's
// output); each line is chomped, blank lines skipped and a trailing "/x" (one
// word character) removed.  Sources are read in the order given, records in
// file order, and every record whose core name is selected is written:
//   fq: the 4 lines as read (Fastq_reader.pm:47-70), the core name being the
//       first whitespace field of the '@' line without '@' and without a
//       trailing /1 or /2 (:96-113), then with its first "_forward" and first
//       "_reverse" removed (insilico_read_normalization.pl:531-533);
//   fa: ">header\nsequence\n" (Fasta_reader.pm:45-73,139-153, no line
//       wrapping), the accession being the header's first whitespace field
//       without a trailing /1 or /2 and blanks.
// A selected accession that no record matched is an error: the list goes to
// <out>.missing_accs (sorted) and the tool exits non-zero.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "extract.h"
#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string type, selected, output;
  if (!a.value({"--seqType"}, type) || !a.value({"--selected"}, selected) || !a.value({"--output", "-o"}, output))
    die("usage: extract_normalized_reads --seqType fq|fa --selected <accs> --output <out> <reads>...");
  if (type != "fq" && type != "fa") die("Error, do not recognize format: %s", type.c_str());
  if (!a.unknown_options().empty()) die("unknown option %s", a.unknown_options()[0].c_str());
  std::vector<std::string> sources = a.positional();
  if (sources.empty()) die("no source reads given");
  SelectedIndex idx;
  try {
    idx = build_selected_index(selected);
  } catch (const std::exception& ex) {
    die("failed to read selected_entries file %s: %s", selected.c_str(), ex.what());
  }
  FILE* f = fopen(output.c_str(), "w");
  if (!f) die("cannot write to %s", output.c_str());
  {
    Out out(f);
    for (const std::string& s : sources) {
      fprintf(stderr, "-capturing normalized reads from: %s\n", s.c_str());
      try {
        extract_records(type, s, idx, out);
      } catch (const std::exception& ex) {
        die("%s", ex.what());
      }
    }
  }
  fclose(f);
  const std::vector<std::string> missing = missing_accessions(idx);
  if (!missing.empty()) {
    const std::string mp = output + ".missing_accs";
    FILE* m = fopen(mp.c_str(), "w");
    if (!m) die("cannot write to file %s", mp.c_str());
    for (const std::string& s : missing) fprintf(m, "%s\n", s.c_str());
    fclose(m);
    std::string all;
    for (const std::string& s : sources) all += (all.empty() ? "" : " ") + s;
    die("not all specified records have been retrieved (missing %zu) from %s, see file: %s for list of "
        "missing entries", missing.size(), all.c_str(), mp.c_str());
  }
  return 0;
}
