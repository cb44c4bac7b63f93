// ingest.h -- read preparation shared by fastQ_to_fastA / seqtk-trinity and
// insilico_read_normalization: FASTQ -> FASTA record conversion and the
// reverse complement of FASTA files.
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "tool_util.h"

namespace tools {

struct ConvertOpts {
  bool rev = false;           // reverse-complement each sequence
  bool ignore_dirty = false;  // skip malformed records instead of failing
  bool seqtk = false;         // seqtk-trinity's -R rule and revcomp table
  std::string append;         // -a N (fastQ_to_fastA) / -R N (seqtk-trinity)
};

struct ConvertTally {
  uint64_t clean = 0, dirty = 0;
};

// A malformed FASTQ record with ignore_dirty off; what() is the message
// fastQ_to_fastA.pl prints before exiting with status 255.
struct DirtyRecord : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Append ">name\nsequence\n" for every kept FASTQ record of `path`
// (fastQ_to_fastA.pl:66-117; the rules are in fastq_to_fasta_main.cpp).
void fastq_to_fasta(const std::string& path, const ConvertOpts& o, Out& out, ConvertTally& t);

// util/support_scripts/revcomp_fasta.pl: records split at '>', each header
// kept, the rest of the record reversed, complemented with the IUPAC table
// (newlines deleted) and written in lines of 60.
void revcomp_fasta(const std::string& path, Out& out);

}  // namespace tools
