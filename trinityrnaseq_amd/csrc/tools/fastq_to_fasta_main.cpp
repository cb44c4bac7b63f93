// FASTQ -> FASTA read ingest, the first step of Trinity's read preparation.
//
//   fastQ_to_fastA -I in1.fq[,in2.fq.gz,...] [--rev] [-a N] [--ignore_dirty]
//       = util/support_scripts/fastQ_to_fastA.pl (Trinity:2772-2781, the
//         --no_seqtk path of prep_seqs)
//   seqtk-trinity seq -A [-r] -R N <in.fq|->
//       = the seqtk-trinity call of prep_seqs (Trinity:2783-2806): read type
//         N (1 = left, 2 = right) goes on names that carry none
//
// One binary, dispatched on its name.  Record rules (fastQ_to_fastA.pl:66-117):
// a line starting with '@' opens a 4-line record; every \r of that header line
// is deleted; the record is kept when header, sequence, '+' line and a
// non-blank quality line of the sequence's length are all there, else it is
// "dirty" (fatal unless --ignore_dirty).  The name is the first
// whitespace-separated field of the header after '@'; when more fields follow,
// the name does not end in /1 or /2 and the next field starts with "1:" or
// "2:" (Casava 1.8), "/1" or "/2" is appended; -a N appends "/N"; --rev
// reverse-complements the sequence with Nuc_translator.pm:286-292's table.
// Output: ">name\nsequence\n" per kept record.
//
// seqtk-trinity's own source is not in the reference tree: its -R rule for a
// name with neither a /1,/2 suffix nor a Casava field (append "/N"), its
// reverse-complement table (revcomp_fasta.pl's IUPAC table here) and its
// handling of dirty records (skipped, counted on stderr) are this tool's
// reading of the role, not pinned by the reference.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "ingest.h"
#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

namespace {

inline bool perl_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

using Opts = ConvertOpts;
using Tally = ConvertTally;

// One input file; a malformed record ends the tool with status 255, as the
// Perl script's exit does.
void convert(const std::string& path, const Opts& o, Out& out, Tally& t) {
  try {
    fastq_to_fasta(path, o, out, t);
  } catch (const DirtyRecord& d) {
    fputs(d.what(), stderr);
    exit(255);
  }
}

[[noreturn]] void usage_fq2fa() {
  fprintf(stderr,
          "\n##########################################################\n#\n"
          "#  -I <string>     input.fq  or \"input1.fq,input2.fq,input3.fq\"\n#\n"
          "#  --ignoreDirty   ignores poorly formed entries\n#\n"
          "#  -a <int>        append \"/num\" to the accession name.\n#\n"
          "#  --rev           reverse complement nucleotide sequence.\n#\n"
          "###########################################################\n\n");
  exit(255);
}

int main_fq2fa(int argc, char** argv) {
  Args a(argc, argv, 1);
  Opts o;
  std::string inputs;
  if (!a.value({"-I"}, inputs)) usage_fq2fa();
  o.ignore_dirty = a.flag({"--ignore_dirty", "-ignore_dirty"});
  o.rev = a.flag({"--rev", "-rev"});
  long long an = a.integer({"-a"}, -1);
  if (an >= 0) o.append = std::to_string(an);
  if (!a.unknown_options().empty()) die("unknown option %s", a.unknown_options()[0].c_str());
  Out out(stdout);
  Tally t;
  size_t s = 0;
  while (s <= inputs.size()) {
    size_t e = inputs.find(',', s);
    if (e == std::string::npos) e = inputs.size();
    std::string f;
    for (size_t i = s; i < e; i++)
      if (!perl_space(inputs[i])) f.push_back(inputs[i]);
    bool word = false;
    for (char c : f)
      if (isalnum((unsigned char)c) || c == '_') word = true;
    if (word) {
      try {
        convert(f, o, out, t);
      } catch (const std::exception& ex) {
        die("cannot open file %s (%s)", f.c_str(), ex.what());
      }
    }
    s = e + 1;
  }
  out.flush();
  return 0;
}

int main_seqtk(int argc, char** argv) {
  if (argc < 2 || strcmp(argv[1], "seq") != 0)
    die("usage: seqtk-trinity seq -A [-r] [-R 1|2] <in.fq|->");
  Opts o;
  o.seqtk = true;
  o.ignore_dirty = true;
  bool fasta = false;
  std::string in;
  for (int i = 2; i < argc; i++) {
    const std::string s = argv[i];
    if (s == "-A") fasta = true;
    else if (s == "-r") o.rev = true;
    else if (s == "-R") {
      if (i + 1 >= argc) die("-R needs a value");
      o.append = argv[++i];
      if (o.append != "1" && o.append != "2") die("-R takes 1 or 2");
    } else if (s == "-" || s[0] != '-') {
      if (!in.empty()) die("one input only");
      in = s;
    } else die("unknown option %s", s.c_str());
  }
  if (!fasta) die("only FASTA output (-A) is supported");
  if (in.empty()) in = "-";
  Out out(stdout);
  Tally t;
  try {
    convert(in, o, out, t);
  } catch (const std::exception& ex) {
    die("%s", ex.what());
  }
  out.flush();
  if (t.dirty) fprintf(stderr, "seqtk-trinity: %llu malformed record(s) skipped\n", (unsigned long long)t.dirty);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const char* base = strrchr(argv[0], '/');
  base = base ? base + 1 : argv[0];
  if (strstr(base, "seqtk")) return main_seqtk(argc, argv);
  return main_fq2fa(argc, argv);
}
