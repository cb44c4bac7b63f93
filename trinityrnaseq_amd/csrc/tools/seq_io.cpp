// seq_io.cpp -- FASTA/FASTQ ingest (see seq_io.h).
#include "seq_io.h"

#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <zlib.h>

#include <stdexcept>

namespace tools {

static bool ends_with(const std::string& s, const char* suf) {
  const size_t n = strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

TextSource::TextSource(const std::string& path) {
  if (path == "-") {
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, stdin)) > 0) owned_.append(buf, n);
    data_ = owned_.data();
    size_ = owned_.size();
    return;
  }
  if (ends_with(path, ".gz")) {
    gzFile g = gzopen(path.c_str(), "rb");
    if (!g) throw std::runtime_error("cannot open " + path);
    gzbuffer(g, 1 << 20);
    std::vector<char> buf(1 << 20);
    int n;
    while ((n = gzread(g, buf.data(), (unsigned)buf.size())) > 0) owned_.append(buf.data(), (size_t)n);
    int err = 0;
    const char* msg = gzerror(g, &err);
    const bool bad = n < 0 || (err != Z_OK && err != Z_STREAM_END);
    const std::string why = bad ? std::string(msg ? msg : "") : std::string();
    gzclose(g);
    if (bad) throw std::runtime_error(path + ": gzip error: " + why);
    data_ = owned_.data();
    size_ = owned_.size();
    return;
  }
  fd_ = open(path.c_str(), O_RDONLY);
  if (fd_ < 0) throw std::runtime_error("cannot open " + path);
  struct stat st;
  if (fstat(fd_, &st) != 0) throw std::runtime_error("cannot stat " + path);
  size_ = (uint64_t)st.st_size;
  if (size_) {
    void* p = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
    if (p == MAP_FAILED) throw std::runtime_error("cannot mmap " + path);
    madvise(p, size_, MADV_SEQUENTIAL);
    data_ = (const char*)p;
    mapped_ = true;
  }
}

TextSource::~TextSource() {
  if (mapped_) munmap((void*)data_, size_);
  if (fd_ >= 0) close(fd_);
}

SeqReader::SeqReader(const std::string& path, bool keep_names)
    : src_(path), data_(src_.data()), size_(src_.size()), keep_names_(keep_names) {
  start(path);
}

SeqReader::SeqReader(const char* data, uint64_t size, bool keep_names, const std::string& label)
    : src_(data, size), data_(src_.data()), size_(src_.size()), keep_names_(keep_names) {
  start(label);
}

void SeqReader::start(const std::string& path) {
  // skip leading whitespace (Fasta_reader.pm:55-59)
  while (pos_ < size_ && (data_[pos_] == '\n' || data_[pos_] == '\r' || data_[pos_] == ' ' ||
                          data_[pos_] == '\t'))
    pos_++;
  if (pos_ < size_) {
    if (data_[pos_] == '@') fastq_ = true;
    else if (data_[pos_] != '>')
      throw std::runtime_error(path + ": not FASTA or FASTQ (first record does not start with > or @)");
  }
}

SeqReader::~SeqReader() {}

static inline bool keep_seq_char(unsigned char c) {
  // drop whitespace and control characters (\000-\037, \177-\377)
  return c > 32 && c < 127;
}

bool SeqReader::next_batch(SeqSet& out, uint64_t max_records, uint64_t max_bases) {
  if (out.off.empty()) out.off.push_back(0);
  if (out.name_off.empty()) out.name_off.push_back(0);
  uint64_t got = 0;
  const uint64_t base0 = out.seq.size();
  while (pos_ < size_ && got < max_records && out.seq.size() - base0 < max_bases) {
    // header line
    uint64_t h = pos_ + 1;
    uint64_t e = h;
    while (e < size_ && data_[e] != '\n') e++;
    if (keep_names_) {
      // control bytes are deleted first, then the header splits on blanks
      for (uint64_t a = h; a < e; a++) {
        const unsigned char c = (unsigned char)data_[a];
        if (c == ' ' || c == '\t') break;
        if (c < 32 || c >= 127) continue;
        out.names.push_back((char)c);
      }
      out.name_off.push_back(out.names.size());
    }
    pos_ = e < size_ ? e + 1 : e;
    if (fastq_) {
      uint64_t s = pos_;
      while (pos_ < size_ && data_[pos_] != '\n') pos_++;
      for (uint64_t i = s; i < pos_; i++)
        if (keep_seq_char((unsigned char)data_[i])) out.seq.push_back(data_[i]);
      if (pos_ < size_) pos_++;
      // '+' line and quality line
      while (pos_ < size_ && data_[pos_] != '\n') pos_++;
      if (pos_ < size_) pos_++;
      while (pos_ < size_ && data_[pos_] != '\n') pos_++;
      if (pos_ < size_) pos_++;
      while (pos_ < size_ && data_[pos_] != '@' && (data_[pos_] == '\n' || data_[pos_] == '\r'))
        pos_++;
    } else {
      // sequence lines until a line starting with '>'
      while (pos_ < size_ && data_[pos_] != '>') {
        uint64_t s = pos_;
        while (pos_ < size_ && data_[pos_] != '\n') pos_++;
        for (uint64_t i = s; i < pos_; i++)
          if (keep_seq_char((unsigned char)data_[i])) out.seq.push_back(data_[i]);
        if (pos_ < size_) pos_++;
      }
    }
    out.off.push_back(out.seq.size());
    got++;
  }
  return got > 0;
}

SeqSet read_all(const std::string& path, bool keep_names) {
  SeqSet s;
  s.clear();
  SeqReader r(path, keep_names);
  s.seq.reserve(r.file_size());
  while (r.next_batch(s, UINT64_MAX, UINT64_MAX)) {
  }
  return s;
}

SeqSet parse_all(const std::string& text, bool keep_names, const std::string& label) {
  SeqSet s;
  s.clear();
  SeqReader r(text.data(), text.size(), keep_names, label);
  s.seq.reserve(text.size());
  while (r.next_batch(s, UINT64_MAX, UINT64_MAX)) {
  }
  return s;
}

}  // namespace tools
