// seq_io.h -- FASTA/FASTQ ingest for the CLI tools.
//
// FASTA records follow PerlLib/Fasta_reader.pm:45-73: a record runs from '>'
// to the next line starting with '>', the accession is the header's first
// whitespace-delimited token (Sequence::new, :110), the sequence is the
// remaining lines joined with all whitespace and control characters removed.
// FASTQ (4-line records, '@' header) is accepted as jellyfish does.
// Every reader takes a plain file (memory-mapped), a gzip file (name ending
// in .gz, inflated into memory) or "-" for standard input.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace tools {

struct SeqSet {
  std::string seq;                 // concatenated sequences
  std::vector<uint64_t> off;       // n+1 offsets into seq
  std::string names;               // concatenated accessions
  std::vector<uint64_t> name_off;  // n+1 offsets into names
  uint64_t size() const { return off.empty() ? 0 : off.size() - 1; }
  std::string name(uint64_t i) const {
    return names.substr(name_off[i], name_off[i + 1] - name_off[i]);
  }
  void clear() {
    seq.clear();
    off.assign(1, 0);
    names.clear();
    name_off.assign(1, 0);
  }
};

// The bytes of one input file: mapped for plain files, inflated for .gz,
// read whole for "-" (stdin).
class TextSource {
 public:
  explicit TextSource(const std::string& path);
  // A view of bytes the caller keeps alive.
  TextSource(const char* data, uint64_t size) : data_(data), size_(size) {}
  ~TextSource();
  TextSource(const TextSource&) = delete;
  TextSource& operator=(const TextSource&) = delete;
  const char* data() const { return data_; }
  uint64_t size() const { return size_; }

 private:
  const char* data_ = nullptr;
  uint64_t size_ = 0;
  bool mapped_ = false;
  std::string owned_;
  int fd_ = -1;
};

// Streaming reader over one file.  next_batch appends up to max_records
// records (and stops once max_bases is exceeded); returns false at EOF with
// nothing appended.
class SeqReader {
 public:
  explicit SeqReader(const std::string& path, bool keep_names = true);
  // Over text in memory (kept alive by the caller); `label` names it in errors.
  SeqReader(const char* data, uint64_t size, bool keep_names, const std::string& label);
  ~SeqReader();
  bool next_batch(SeqSet& out, uint64_t max_records, uint64_t max_bases);
  uint64_t file_size() const { return size_; }

 private:
  TextSource src_;
  const char* data_ = nullptr;
  uint64_t size_ = 0;
  uint64_t pos_ = 0;
  bool fastq_ = false;
  bool keep_names_ = true;
  void start(const std::string& label);
};

// Read every record of a file.
SeqSet read_all(const std::string& path, bool keep_names = true);
// Every record of FASTA/FASTQ text in memory.
SeqSet parse_all(const std::string& text, bool keep_names, const std::string& label);

}  // namespace tools
