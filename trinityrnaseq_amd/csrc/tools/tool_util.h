// tool_util.h -- shared helpers of the CLI tools (argument parsing, ABI error
// handling, k-mer text, the count database file).
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "trinity_gpu.h"

namespace tools {

[[noreturn]] void die(const char* fmt, ...);
// While set on a thread, die() there throws std::runtime_error(message)
// instead of exiting: a resident process runs many jobs, and one job's error
// must end that job only.
void die_throws(bool on);
// Exit 1 with the library's message when rc != TK_OK (Pipeliner treats any
// non-zero exit as failure: PerlLib/Pipeliner.pm:156-195).
void check(int rc, const char* what);

// Device index from TRINITY_GPU_DEVICE (default 0).
int device_from_env();
tk_ctx* open_ctx();

// Minimal option scanner: options may be "-x v", "--x v", "--x=v" or flags.
class Args {
 public:
  Args(int argc, char** argv, int first);
  bool flag(std::initializer_list<const char*> names);
  bool value(std::initializer_list<const char*> names, std::string& out);
  long long integer(std::initializer_list<const char*> names, long long dflt);
  double real(std::initializer_list<const char*> names, double dflt);
  std::vector<std::string> positional() const;   // arguments not consumed
  std::vector<std::string> unknown_options() const;

 private:
  std::vector<std::string> a_;
  std::vector<bool> used_;
  int find(std::initializer_list<const char*> names, bool needs_value, std::string* val);
};

uint64_t kmer_from_text(const char* s, int k, bool* ok);
void kmer_to_text(uint64_t v, int k, char* out);  // writes k chars
uint64_t revcomp(uint64_t v, int k);

// Buffered writer for large text outputs.
class Out {
 public:
  explicit Out(FILE* f) : f_(f) { buf_.reserve(1 << 22); }
  // Appends to *sink instead of a file.
  explicit Out(std::string* sink) : sink_(sink) { buf_.reserve(1 << 22); }
  ~Out() {
    try {
      flush();
    } catch (...) {  // die() in a job thread: the job already failed
    }
  }
  void put(const char* s, size_t n) {
    buf_.append(s, n);
    if (buf_.size() >= (1u << 22)) flush();
  }
  void put(const std::string& s) { put(s.data(), s.size()); }
  void putc(char c) { buf_.push_back(c); }
  void put_u64(uint64_t v);
  void flush();

 private:
  FILE* f_ = nullptr;
  std::string* sink_ = nullptr;
  std::string buf_;
};

// Count database written by `jellyfish count` and read by dump/histo/inchworm:
//   "TKJF1\n" u32 k, u32 canonical, u64 n, then n u64 keys (ascending), n u32 counts.
struct CountDb {
  int k = 25;
  int canonical = 1;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> counts;
};
void write_db(const std::string& path, const CountDb& db);
CountDb read_db(const std::string& path);

// Parse a jellyfish FASTA dump (">count\nKMER") into keys/counts.  With
// canonical set, each key is replaced by min(kmer, revcomp) and duplicates add.
void read_kmer_dump(const std::string& path, int k, bool canonical, std::vector<uint64_t>& keys,
                    std::vector<uint32_t>& counts);

}  // namespace tools
