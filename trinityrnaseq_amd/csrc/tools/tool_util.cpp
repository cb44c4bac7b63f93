// tool_util.cpp -- shared helpers of the CLI tools.
#include "tool_util.h"

#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>

#include "seq_io.h"

namespace tools {

static thread_local bool t_die_throws = false;

void die_throws(bool on) { t_die_throws = on; }

void die(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (t_die_throws) {
    char buf[1024];
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw std::runtime_error(buf);
  }
  if (strncmp(fmt, "Error", 5) != 0) fputs("Error: ", stderr);  // messages in the scripts' own wording print as is
  vfprintf(stderr, fmt, ap);
  fputc('\n', stderr);
  va_end(ap);
  exit(1);
}

void check(int rc, const char* what) {
  if (rc != TK_OK) die("%s failed (%d): %s", what, rc, tk_last_error());
}

int device_from_env() {
  const char* e = getenv("TRINITY_GPU_DEVICE");
  return e ? atoi(e) : 0;
}

tk_ctx* open_ctx() {
  tk_ctx* c = nullptr;
  check(tk_init(device_from_env(), &c), "tk_init");
  return c;
}

Args::Args(int argc, char** argv, int first) {
  for (int i = first; i < argc; i++) a_.push_back(argv[i]);
  used_.assign(a_.size(), false);
}

int Args::find(std::initializer_list<const char*> names, bool needs_value, std::string* val) {
  for (size_t i = 0; i < a_.size(); i++) {
    if (used_[i]) continue;
    for (const char* n : names) {
      const std::string& s = a_[i];
      const size_t nl = strlen(n);
      if (s == n) {
        used_[i] = true;
        if (needs_value) {
          if (i + 1 >= a_.size()) die("option %s needs a value", n);
          used_[i + 1] = true;
          *val = a_[i + 1];
        }
        return (int)i;
      }
      if (needs_value && s.size() > nl + 1 && s.compare(0, nl, n) == 0 && s[nl] == '=') {
        used_[i] = true;
        *val = s.substr(nl + 1);
        return (int)i;
      }
    }
  }
  return -1;
}

bool Args::flag(std::initializer_list<const char*> names) { return find(names, false, nullptr) >= 0; }

bool Args::value(std::initializer_list<const char*> names, std::string& out) {
  return find(names, true, &out) >= 0;
}

long long Args::integer(std::initializer_list<const char*> names, long long dflt) {
  std::string v;
  if (!value(names, v)) return dflt;
  char* end = nullptr;
  double d = strtod(v.c_str(), &end);  // jellyfish accepts -s 100e6 style sizes
  if (!end || *end) die("bad integer value '%s'", v.c_str());
  return (long long)d;
}

double Args::real(std::initializer_list<const char*> names, double dflt) {
  std::string v;
  if (!value(names, v)) return dflt;
  char* end = nullptr;
  double d = strtod(v.c_str(), &end);
  if (!end || *end) die("bad numeric value '%s'", v.c_str());
  return d;
}

std::vector<std::string> Args::positional() const {
  std::vector<std::string> p;
  for (size_t i = 0; i < a_.size(); i++)
    if (!used_[i] && (a_[i].empty() || a_[i][0] != '-' || a_[i] == "-")) p.push_back(a_[i]);
  return p;
}

std::vector<std::string> Args::unknown_options() const {
  std::vector<std::string> p;
  for (size_t i = 0; i < a_.size(); i++)
    if (!used_[i] && a_[i].size() > 1 && a_[i][0] == '-') p.push_back(a_[i]);
  return p;
}

static const int8_t* code_table() {
  static int8_t t[256];
  static bool init = false;
  if (!init) {
    memset(t, -1, sizeof t);
    t['A'] = t['a'] = 0;
    t['C'] = t['c'] = 1;
    t['G'] = t['g'] = 2;
    t['T'] = t['t'] = 3;
    init = true;
  }
  return t;
}

uint64_t kmer_from_text(const char* s, int k, bool* ok) {
  const int8_t* t = code_table();
  uint64_t v = 0;
  *ok = true;
  for (int i = 0; i < k; i++) {
    int c = t[(unsigned char)s[i]];
    if (c < 0) { *ok = false; return 0; }
    v = (v << 2) | (uint64_t)c;
  }
  return v;
}

void kmer_to_text(uint64_t v, int k, char* out) {
  static const char B[4] = {'A', 'C', 'G', 'T'};
  for (int i = k - 1; i >= 0; i--) {
    out[i] = B[v & 3];
    v >>= 2;
  }
}

uint64_t revcomp(uint64_t v, int k) {
  uint64_t r = 0;
  for (int i = 0; i < k; i++) {
    r = (r << 2) | (3 - (v & 3));
    v >>= 2;
  }
  return r;
}

void Out::put_u64(uint64_t v) {
  char tmp[24];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) buf_.push_back(tmp[--n]);
  if (buf_.size() >= (1u << 22)) flush();
}

void Out::flush() {
  if (sink_) {
    sink_->append(buf_);
    buf_.clear();
    return;
  }
  if (!buf_.empty()) {
    if (fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) die("write failed");
    buf_.clear();
  }
}

static const char kMagic[6] = {'T', 'K', 'J', 'F', '1', '\n'};

void write_db(const std::string& path, const CountDb& db) {
  std::string tmp = path + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) die("cannot write %s", tmp.c_str());
  uint32_t hk = (uint32_t)db.k, hc = (uint32_t)db.canonical;
  uint64_t n = db.keys.size();
  bool ok = fwrite(kMagic, 1, 6, f) == 6 && fwrite(&hk, 4, 1, f) == 1 && fwrite(&hc, 4, 1, f) == 1 &&
            fwrite(&n, 8, 1, f) == 1 && (n == 0 || fwrite(db.keys.data(), 8, n, f) == n) &&
            (n == 0 || fwrite(db.counts.data(), 4, n, f) == n);
  if (fclose(f) != 0 || !ok) die("short write to %s", tmp.c_str());
  if (rename(tmp.c_str(), path.c_str()) != 0) die("cannot rename %s", tmp.c_str());
}

CountDb read_db(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) die("cannot open %s", path.c_str());
  char m[6];
  uint32_t hk = 0, hc = 0;
  uint64_t n = 0;
  if (fread(m, 1, 6, f) != 6 || memcmp(m, kMagic, 6) != 0) die("%s is not a k-mer count database", path.c_str());
  if (fread(&hk, 4, 1, f) != 1 || fread(&hc, 4, 1, f) != 1 || fread(&n, 8, 1, f) != 1)
    die("truncated header in %s", path.c_str());
  CountDb db;
  db.k = (int)hk;
  db.canonical = (int)hc;
  db.keys.resize(n);
  db.counts.resize(n);
  if (n && (fread(db.keys.data(), 8, n, f) != n || fread(db.counts.data(), 4, n, f) != n))
    die("truncated database %s", path.c_str());
  fclose(f);
  return db;
}

void read_kmer_dump(const std::string& path, int k, bool canonical, std::vector<uint64_t>& keys,
                    std::vector<uint32_t>& counts) {
  SeqReader r(path, true);
  SeqSet s;
  s.clear();
  keys.clear();
  counts.clear();
  while (true) {
    s.clear();
    if (!r.next_batch(s, 1 << 22, UINT64_MAX)) break;
    for (uint64_t i = 0; i < s.size(); i++) {
      const uint64_t len = s.off[i + 1] - s.off[i];
      if ((int)len != k) die("%s: record %s has length %llu, expected k=%d", path.c_str(),
                             s.name(i).c_str(), (unsigned long long)len, k);
      bool ok = false;
      uint64_t v = kmer_from_text(s.seq.data() + s.off[i], k, &ok);
      if (!ok) die("%s: non-ACGT k-mer in record %llu", path.c_str(), (unsigned long long)i);
      if (canonical) {
        uint64_t rc = revcomp(v, k);
        if (rc < v) v = rc;
      }
      const std::string nm = s.name(i);
      char* end = nullptr;
      unsigned long long c = strtoull(nm.c_str(), &end, 10);
      if (nm.empty() || !end || *end) die("%s: header '%s' is not a count", path.c_str(), nm.c_str());
      keys.push_back(v);
      counts.push_back(c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c);
    }
  }
}

}  // namespace tools
