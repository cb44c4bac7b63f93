// extract.h -- the record capture of make_normalized_reads_file
// (util/insilico_read_normalization.pl:470-583), shared by
// extract_normalized_reads and insilico_read_normalization.
#pragma once
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "tool_util.h"

namespace tools {

// core accession -> number of records written for it
using SelectedIndex = std::unordered_map<std::string, uint64_t>;

// build_selected_index (:470-502): one accession per line, each line chomped,
// blank lines skipped, a trailing "/x" (one word character) removed.
SelectedIndex build_selected_index(const std::string& path);
SelectedIndex build_selected_index_text(const std::string& text);

// Write every record of `path` whose core name is in idx, in file order.
//   fq: the 4 lines as read (Fastq_reader.pm:47-70); the core name is the
//       first whitespace field of the '@' line without '@' and a trailing
//       /1 or /2 (:96-113), then without its first "_forward" and first
//       "_reverse" (insilico_read_normalization.pl:531-533);
//   fa: ">header\nsequence\n" (Fasta_reader.pm:45-73,139-153, no line
//       wrapping); the accession is the header's first whitespace field
//       without a trailing /1 or /2 and blanks.
// seq_type is "fq" or "fa".
void extract_records(const std::string& seq_type, const std::string& path, SelectedIndex& idx, Out& out);

// Accessions no record matched, sorted.
std::vector<std::string> missing_accessions(const SelectedIndex& idx);

}  // namespace tools
