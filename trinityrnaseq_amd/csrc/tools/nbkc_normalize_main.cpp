// nbkc_normalize -- pair merge and coverage-normalization selection.
//
//   nbkc_normalize --left L.stats.sort --right R.stats.sort [--sorted]
//                  --max_cov C [--min_cov 1] [--max_CV 10000] [--merged_out pairs.K25.stats]
//       = nbkc_merge_left_right_stats.pl --sorted (:80-145) followed by
//         nbkc_normalize.pl (:67-127), as run by
//         util/insilico_read_normalization.pl:951-976
//   nbkc_normalize --stats_file pairs.K25.stats --max_cov C [--min_cov 1] [--max_CV 10000]
//       = nbkc_normalize.pl alone (also the single-end path, :917-944)
//
// Selected core accessions go to stdout, one per line, in file order.  The
// stream of rand(1) values is Perl's drand48 after srand(12345); every
// numeric field is taken from the stats text exactly as Perl numifies it.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

#include <memory>
#include <string>
#include <vector>

#include "tool_util.h"

using namespace tools;

namespace {

// DelimParser::Reader (PerlLib/DelimParser.pm:121-173) over a tab-separated file.
class DelimReader {
 public:
  explicit DelimReader(const std::string& path) : path_(path) {
    f_ = fopen(path.c_str(), "r");
    if (!f_) die("cannot open %s", path.c_str());
    std::string h;
    if (!getline(h) || h.empty()) die("Error, no header row read. (%s)", path.c_str());
    cols_ = split(h, true);
  }
  ~DelimReader() {
    if (f_) fclose(f_);
  }
  // Next row as fields in header order; false at EOF.
  bool next(std::vector<std::string>& fields) {
    std::string line;
    if (!getline(line)) return false;
    fields = split(line, false);
    if (fields.size() != cols_.size())
      die("Error, line: [%s] is lacking %zu fields (%s)", line.c_str(), cols_.size(), path_.c_str());
    return true;
  }
  int col(const char* name) const {
    for (size_t i = 0; i < cols_.size(); i++)
      if (cols_[i] == name) return (int)i;
    die("%s: no column '%s'", path_.c_str(), name);
  }

 private:
  bool getline(std::string& out) {
    out.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f_)) != EOF) {
      any = true;
      if (c == '\n') break;
      out.push_back((char)c);
    }
    return any;
  }
  // The header is chomped before split(/\t/) (trailing empty fields drop);
  // a row is split with its newline attached and chomped after, so its
  // trailing empty fields stay (DelimParser.pm:128-131,150-152).
  static std::vector<std::string> split(const std::string& s, bool drop_trailing) {
    std::vector<std::string> v;
    size_t a = 0;
    while (true) {
      size_t b = s.find('\t', a);
      if (b == std::string::npos) {
        v.push_back(s.substr(a));
        break;
      }
      v.push_back(s.substr(a, b - a));
      a = b + 1;
    }
    while (drop_trailing && !v.empty() && v.back().empty()) v.pop_back();
    return v;
  }
  std::string path_;
  FILE* f_ = nullptr;
  std::vector<std::string> cols_;
};

// Perl numification of a field (leading number, 0 when there is none).
double num(const std::string& s) { return strtod(s.c_str(), nullptr); }

std::string fmt1(double v) {
  char b[64];
  snprintf(b, sizeof b, "%.1f", v);
  return b;
}

struct Drand48 {
  uint64_t x;
  explicit Drand48(uint32_t seed) : x((((uint64_t)seed << 16) | 0x330E) & ((1ull << 48) - 1)) {}
  double operator()() {
    x = (0x5DEECE66Dull * x + 0xB) & ((1ull << 48) - 1);
    return (double)x * 0x1p-48;
  }
};

// "/1" style suffix removal of nbkc_merge_left_right_stats.pl: ^(\S+)/\d$
std::string core_of(const std::string& acc) {
  const size_t n = acc.size();
  if (n >= 3 && acc[n - 2] == '/' && isdigit((unsigned char)acc[n - 1])) {
    bool ok = true;
    for (size_t i = 0; i + 2 < n; i++)
      if (isspace((unsigned char)acc[i])) ok = false;
    if (ok) return acc.substr(0, n - 2);
  }
  return acc;
}

struct Selector {
  int max_cov, min_cov, max_cv;
  Drand48 rnd{12345};
  uint64_t total = 0, selected = 0, aberrant = 0, below = 0;
  Out* out;
  // nbkc_normalize.pl:78-116 on one merged row
  void row(std::string acc, const std::string& med_s, const std::string& mean_s, const std::string& sd_s) {
    total++;
    if (acc.size() >= 2 && acc[acc.size() - 2] == '/' && (acc.back() == '1' || acc.back() == '2'))
      acc.resize(acc.size() - 2);
    const double med = num(med_s);
    if (med < min_cov) { below++; return; }
    const double sd = num(sd_s), u = num(mean_s);
    if (u <= 0) { aberrant++; return; }
    if (sd / u > max_cv) { aberrant++; return; }
    if (rnd() <= (double)max_cov / med) {
      out->put(acc);
      out->putc('\n');
      selected++;
    }
  }
};

}  // namespace

int main(int argc, char** argv) {
  Args a(argc, argv, 1);
  std::string left, right, stats, merged_out;
  a.value({"--left"}, left);
  a.value({"--right"}, right);
  a.value({"--stats_file"}, stats);
  a.value({"--merged_out"}, merged_out);
  const bool sorted = a.flag({"--sorted"});
  Selector sel;
  sel.max_cov = (int)a.integer({"--max_cov"}, 0);
  sel.min_cov = (int)a.integer({"--min_cov"}, 1);
  sel.max_cv = (int)a.integer({"--max_CV"}, -1);
  auto unk = a.unknown_options();
  if (!unk.empty()) die("unknown option %s", unk[0].c_str());
  if (!sel.max_cov || sel.max_cv < 0) die("need --max_cov and --max_CV");
  if (stats.empty() == (left.empty() || right.empty())) die("need --stats_file, or --left and --right");
  Out out(stdout);
  sel.out = &out;
  if (!stats.empty()) {
    DelimReader r(stats);
    const int ca = r.col("acc"), cm = r.col("median_cov"), cu = r.col("mean_cov"), cs = r.col("stdev");
    std::vector<std::string> f;
    while (r.next(f)) sel.row(f[ca], f[cm], f[cu], f[cs]);
  } else {
    DelimReader L(left), R(right);
    const int la = L.col("acc"), lm = L.col("median_cov"), lu = L.col("mean_cov"), ls = L.col("stdev");
    const int ra = R.col("acc"), rm = R.col("median_cov"), ru = R.col("mean_cov"), rs = R.col("stdev");
    std::unique_ptr<Out> mo;
    FILE* mf = nullptr;
    if (!merged_out.empty()) {
      mf = fopen(merged_out.c_str(), "w");
      if (!mf) die("cannot write %s", merged_out.c_str());
      mo.reset(new Out(mf));
      mo->put(std::string("acc\tleft_acc\tleft_median_cov\tleft_mean_cov\tleft_stdev\t"
                          "right_acc\tright_median_cov\tright_mean_cov\tright_stdev\t"
                          "median_cov\tmean_cov\tstdev\n"));
    }
    std::vector<std::string> l, r;
    bool hl = L.next(l), hr = R.next(r);
    while (hl && hr) {
      const std::string lc = core_of(l[la]), rc = core_of(r[ra]);
      if (lc != rc) {
        if (!sorted) die("Error, core accs are not equivalent: [%s] vs. [%s] reads, and --sorted flag wasn't used here.",
                         lc.c_str(), rc.c_str());
        if (l[la] < r[ra]) hl = L.next(l);
        else hr = R.next(r);
        continue;
      }
      const std::string med = fmt1((num(l[lm]) + num(r[rm])) / 2);
      const std::string mean = fmt1((num(l[lu]) + num(r[ru])) / 2);
      const std::string sd = fmt1((num(l[ls]) + num(r[rs])) / 2);
      if (mo) {
        const std::string row = lc + "\t" + l[la] + "\t" + l[lm] + "\t" + l[lu] + "\t" + l[ls] + "\t" + r[ra] +
                                "\t" + r[rm] + "\t" + r[ru] + "\t" + r[rs] + "\t" + med + "\t" + mean + "\t" +
                                sd + "\n";
        mo->put(row);
      }
      sel.row(lc, med, mean, sd);
      hl = L.next(l);
      hr = R.next(r);
    }
    if (mo) {
      mo->flush();
      fclose(mf);
    }
  }
  out.flush();
  if (!sel.total) die("Error, no reads made it to the normalization process...  ");
  fprintf(stderr, "%llu / %llu = %.2f%% reads selected during normalization.\n",
          (unsigned long long)sel.selected, (unsigned long long)sel.total, 100.0 * sel.selected / sel.total);
  fprintf(stderr, "%llu / %llu = %.2f%% reads discarded as likely aberrant based on coverage profiles.\n",
          (unsigned long long)sel.aberrant, (unsigned long long)sel.total, 100.0 * sel.aberrant / sel.total);
  fprintf(stderr, "%llu / %llu = %.2f%% reads discarded as below minimum coverage threshold=%d\n",
          (unsigned long long)sel.below, (unsigned long long)sel.total, 100.0 * sel.below / sel.total, sel.min_cov);
  return 0;
}
