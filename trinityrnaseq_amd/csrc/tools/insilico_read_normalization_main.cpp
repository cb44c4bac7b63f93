// insilico_read_normalization -- one sample's in silico normalization, the
// whole job of util/insilico_read_normalization.pl in one process, with the
// k-mer work on the GPU through the resident normalization service.
//
//   insilico_read_normalization --seqType fq --JM 10G --max_cov 200 --min_cov 1
//       --CPU 4 --output dir --max_CV 10000 [--SS_lib_type RF]
//       (--left L1[,L2..] --right R1[,R2..] [--pairs_together] [--PARALLEL_STATS]
//        | --single S1[,S2..] | --left_list F --right_list F)
//       [--KMER_SIZE 25] [--no_cleanup] [--tmp_dir_name n] [--NO_SEQTK]
//     = the command Trinity's normalize() builds (Trinity:3421-3446) for
//       util/insilico_read_normalization.pl, same options and outputs:
//       <output>/<base>.normalized_K<k>_maxC<c>_minC<m>_maxCV<v>.<fq|fa> per
//       side (:386-400) and the left.norm.<t> / right.norm.<t> (single.norm.<t>)
//       links (:438-452; copies when NO_SYMLINK is set).
//
// Steps (the Perl script's, without its intermediate files):
//   1. read preparation per side (:659-713): FASTQ through the seqtk-trinity
//      rule (-A -R 1|2, -r for a reverse-complemented side) or, with
//      --NO_SEQTK, fastQ_to_fastA.pl's (--rev); FASTA as is, or through
//      revcomp_fasta.pl for an R side;
//   2. on the device (tk_server_submit_job): count of every read of both.fa
//      (canonical unless --SS_lib_type: :617-622), coverage statistics of
//      every read against the k-mers of count >= 2 (dump -L 2: :641;
//      fastaToKmerCoverageStats [--DS]: :837-845), then the selections;
//   3. selection units in the order of the sorted stats (`sort -k1,1` in
//      the C locale: :874, COMMON.pm:7): --pairs_together merges left and
//      right like nbkc_merge_left_right_stats.pl --sorted (lockstep over both
//      sorted lists; when the core names differ the side whose accession is
//      smaller advances: :88-142) and selects pairs; otherwise each side (or
//      the single reads) is selected on its own (:917-944);
//   4. extraction of the selected records from the original files
//      (make_normalized_reads_file: :505-583), one output per side.
// Accessions are sorted by their bytes; two reads with the same accession
// keep their input order (sort would order them by the rest of their stats
// line).
//
// Resident service (the reference starts one process per sample, Trinity:
// 3371-3395, and per component under ParaFly; here each process would pay
// its own device start-up):
//   insilico_read_normalization --serve <socket> [--slots 4] [--hbm_budget_gb 96]
//       keeps one tk_server (its slots overlap the jobs' device work) and runs
//       every job a client sends over the Unix socket, in a thread per job;
//   insilico_read_normalization --server <socket> <options as above>
//       (or TRINITY_GPU_NORM_SERVER=<socket> in the environment) sends its
//       options and working directory there and exits with the job's status,
//       printing the job's log; with only the environment variable set and no
//       server listening, the job runs in this process;
//   insilico_read_normalization --server <socket> --shutdown
//       stops the service once its running jobs end.
//   --plan_only prints the selection units (accessions) and exits before any
//   device work.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <sys/un.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "extract.h"
#include "ingest.h"
#include "seq_io.h"
#include "tool_util.h"

using namespace tools;

namespace {

inline bool perl_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

// ------------------------------------------------------------ options

struct NormOpts {
  std::string seq_type, jm, ss, output, tmp_dir_name = "tmp_normalized_reads";
  std::string left_list, right_list;
  std::vector<std::string> left, right, single;
  long long max_cov = 0, min_cov = 0, max_cv = 10000, k = 25, cpu = 2;
  bool pairs_together = false, parallel_stats = false, no_cleanup = false, no_seqtk = false;
  bool plan_only = false, help = false, devel_cov = false;
  std::string cwd;          // relative paths resolve here
  bool no_symlink = false;  // NO_SYMLINK in the caller's environment
  bool validated = false;
};

std::string usage() {
  return "usage: insilico_read_normalization --seqType fq|fa --JM <n>G --max_cov <int>\n"
         "         (--left <files> --right <files> [--pairs_together] | --single <files>\n"
         "          | --left_list <file> --right_list <file>)\n"
         "         [--SS_lib_type RF|FR|F|R] [--output <dir>] [--CPU n] [--PARALLEL_STATS]\n"
         "         [--KMER_SIZE 25] [--max_CV 10000] [--min_cov 0] [--no_cleanup]\n"
         "         [--tmp_dir_name name] [--NO_SEQTK] [--plan_only]\n"
         "       insilico_read_normalization --serve <socket> [--slots 4] [--hbm_budget_gb 96]\n"
         "       insilico_read_normalization --server <socket> (<options> | --shutdown)\n";
}

bool is_option(const std::string& s) { return s.size() > 1 && s[0] == '-'; }

// Getopt::Long over the script's option set: "-x" and "--x", "--x=v", and
// for left/right/single ("=s{,}") every following argument up to the next
// option.
NormOpts parse_norm_args(const std::vector<std::string>& argv) {
  NormOpts o;
  std::vector<std::string> unknown;
  for (size_t i = 0; i < argv.size(); i++) {
    std::string a = argv[i], val;
    bool has_eq = false;
    if (!is_option(a)) {
      unknown.push_back(a);
      continue;
    }
    std::string name = a.substr(a[1] == '-' ? 2 : 1);
    const size_t eq = name.find('=');
    if (eq != std::string::npos) {
      val = name.substr(eq + 1);
      name = name.substr(0, eq);
      has_eq = true;
    }
    auto take = [&](std::string& out) {
      if (has_eq) {
        out = val;
      } else {
        if (i + 1 >= argv.size()) die("Option %s requires an argument", name.c_str());
        out = argv[++i];
      }
    };
    auto take_int = [&](long long& out) {
      std::string v;
      take(v);
      char* end = nullptr;
      errno = 0;
      const long long x = strtoll(v.c_str(), &end, 10);
      if (v.empty() || *end || errno) die("Value \"%s\" invalid for option %s (number expected)", v.c_str(), name.c_str());
      out = x;
    };
    auto take_list = [&](std::vector<std::string>& out) {
      if (has_eq) {
        out.push_back(val);
        return;
      }
      if (i + 1 >= argv.size()) die("Option %s requires an argument", name.c_str());
      out.push_back(argv[++i]);
      while (i + 1 < argv.size() && !is_option(argv[i + 1])) out.push_back(argv[++i]);
    };
    if (name == "h" || name == "help") o.help = true;
    else if (name == "seqType") take(o.seq_type);
    else if (name == "left") take_list(o.left);
    else if (name == "right") take_list(o.right);
    else if (name == "single") take_list(o.single);
    else if (name == "left_list") take(o.left_list);
    else if (name == "right_list") take(o.right_list);
    else if (name == "SS_lib_type") take(o.ss);
    else if (name == "max_cov") take_int(o.max_cov);
    else if (name == "min_cov") take_int(o.min_cov);
    else if (name == "output") take(o.output);
    else if (name == "JM") take(o.jm);
    else if (name == "KMER_SIZE" || name == "kmer_size") take_int(o.k);
    else if (name == "CPU") take_int(o.cpu);
    else if (name == "PARALLEL_STATS") o.parallel_stats = true;
    else if (name == "max_CV") take_int(o.max_cv);
    else if (name == "pairs_together") o.pairs_together = true;
    else if (name == "no_cleanup") o.no_cleanup = true;
    else if (name == "__devel_report_kmer_cov_stats") o.devel_cov = true;
    else if (name == "jelly_s") {
      long long ignored = 0;
      take_int(ignored);  // jellyfish's hash size: the device table sizes itself
    } else if (name == "tmp_dir_name") take(o.tmp_dir_name);
    else if (name == "NO_SEQTK") o.no_seqtk = true;
    else if (name == "plan_only") o.plan_only = true;
    else unknown.push_back(a);
  }
  if (o.help) die("%s", usage().c_str());
  if (!unknown.empty()) {
    std::string all;
    for (const std::string& u : unknown) all += (all.empty() ? "" : " ") + u;
    die("Error, do not understand options: %s", all.c_str());
  }
  return o;
}

std::string full_path(const std::string& f, const std::string& cwd) {
  if (!f.empty() && f[0] == '/') return f;
  return cwd + "/" + f;
}

std::string basename_of(const std::string& p) {
  std::string s = p;
  while (s.size() > 1 && s.back() == '/') s.pop_back();
  const size_t i = s.rfind('/');
  return i == std::string::npos ? s : s.substr(i + 1);
}

// Perl's split(",", join(",", @list)): empty trailing fields dropped.
std::vector<std::string> split_commas(const std::vector<std::string>& in) {
  std::string all;
  for (size_t i = 0; i < in.size(); i++) all += (i ? "," : "") + in[i];
  std::vector<std::string> out;
  size_t s = 0;
  while (s <= all.size()) {
    size_t e = all.find(',', s);
    if (e == std::string::npos) e = all.size();
    out.push_back(all.substr(s, e - s));
    s = e + 1;
  }
  while (!out.empty() && out.back().empty()) out.pop_back();
  return out;
}

// read_list_file (:578-596): one path per line, lines without a non-blank skipped.
std::vector<std::string> read_list_file(const std::string& path) {
  std::vector<std::string> files;
  TextSource src(path);
  std::string text(src.data(), src.size());
  size_t s = 0;
  while (s < text.size()) {
    size_t e = text.find('\n', s);
    if (e == std::string::npos) e = text.size();
    const std::string line = text.substr(s, e - s);
    s = e + 1;
    bool nonblank = false;
    for (char c : line)
      if (!perl_space(c)) nonblank = true;
    if (nonblank) files.push_back(line);
  }
  return files;
}

// The script's checks, in its order (:184-256); lists resolved, paths made
// absolute.
void validate(NormOpts& o) {
  if (o.validated) return;
  o.validated = true;
  if (o.seq_type != "fq" && o.seq_type != "fa") die("Error, set --seqType to 'fq' or 'fa'");
  if (o.jm.empty() || !isdigit((unsigned char)o.jm[0]) || o.jm.find('G') == std::string::npos)
    die("Error, must set --JM to number of G of RAM (ie. 10G)  ");
  {
    size_t i = 0;
    while (i < o.jm.size() && isdigit((unsigned char)o.jm[i])) i++;
    if (i == 0 || i >= o.jm.size() || o.jm[i] != 'G')
      die("Error, must set --JM to number of G of RAM (ie. 10G)  ");
    bool ok = !o.jm.empty() && o.jm.back() == 'G';
    for (size_t j = 0; j + 1 < o.jm.size(); j++)
      if (!isdigit((unsigned char)o.jm[j]) && o.jm[j] != '.') ok = false;
    if (!ok)
      die("Error, cannot parse max_memory value of %s.  Set it to 'xG' where x is a numerical value", o.jm.c_str());
  }
  if (!o.ss.empty()) {
    if (o.ss != "R" && o.ss != "F" && o.ss != "RF" && o.ss != "FR")
      die("Error, unrecognized SS_lib_type value of %s. Should be: F, R, RF, or FR", o.ss.c_str());
    if (o.ss == "R") o.ss = "F";  // single-end R: counted and scored as F, not reverse-complemented
  }
  if (!o.left_list.empty()) o.left = read_list_file(full_path(o.left_list, o.cwd));
  if (!o.right_list.empty()) o.right = read_list_file(full_path(o.right_list, o.cwd));
  if (!(!o.single.empty() || (!o.left.empty() && !o.right.empty())))
    die("Error, need either options 'left' and 'right' or option 'single'");
  o.left = split_commas(o.left);
  o.right = split_commas(o.right);
  o.single = split_commas(o.single);
  if (!(o.max_cov && o.max_cov >= 2)) die("Error, need to set --max_cov at least 2");
  if (o.pairs_together && !(!o.left.empty() && !o.right.empty()))
    die("Error, if setting --pairs_together, must use the --left and --right parameters.");
  if (o.k < 1 || o.k > 31) die("Error, --KMER_SIZE must be in [1,31] for this build");
  if (o.devel_cov) die("Error, --__devel_report_kmer_cov_stats is not provided by this build");
  if (o.max_cov > 0x7FFFFFFF || o.max_cv > 0x7FFFFFFF || o.min_cov > 0x7FFFFFFF || o.min_cov < -0x7FFFFFFF)
    die("Error, --max_cov / --min_cov / --max_CV out of range");
  for (auto* v : {&o.left, &o.right, &o.single})
    for (std::string& f : *v) f = full_path(f, o.cwd);
  o.output = full_path(o.output.empty() ? o.cwd : o.output, o.cwd);
}

// ------------------------------------------------------------ one job

// Converted reads of one side (the bytes prep_seqs would append to
// <side>.fa), parsed as Fasta_reader reads them.
SeqSet prep_side(const NormOpts& o, const std::vector<std::string>& files, const char* read_type, bool rev,
                 std::string& log) {
  std::string text;
  for (const std::string& f : files) {
    const bool fifo_kind = f.size() > 3 && (f.compare(f.size() - 3, 3, ".xz") == 0 ||
                                            (f.size() > 4 && f.compare(f.size() - 4, 4, ".bz2") == 0));
    if (fifo_kind) die("Error, %s: .xz / .bz2 inputs are not read by this build (plain or .gz)", f.c_str());
    Out out(&text);
    if (o.seq_type == "fq") {
      ConvertOpts co;
      co.rev = rev;
      if (o.no_seqtk) {
        co.ignore_dirty = false;  // fastQ_to_fastA.pl without --ignore_dirty
      } else {
        co.seqtk = true;
        co.ignore_dirty = true;
        co.append = read_type;
      }
      ConvertTally t;
      try {
        fastq_to_fasta(f, co, out, t);
      } catch (const DirtyRecord& d) {
        die("%s", d.what());
      }
      if (t.dirty && co.seqtk) log += "-" + f + ": " + std::to_string(t.dirty) + " malformed record(s) skipped\n";
    } else if (rev) {
      revcomp_fasta(f, out);
    } else {
      TextSource src(f);
      out.put(src.data(), src.size());
    }
    out.flush();
  }
  return parse_all(text, true, files.empty() ? std::string("<none>") : files[0]);
}

// Read indices [begin, end) of `s` in the order of the sorted stats file
// (accession bytes; equal accessions keep their input order).
std::vector<uint64_t> sorted_order(const SeqSet& s, uint64_t begin, uint64_t end) {
  std::vector<uint64_t> idx(end - begin);
  std::iota(idx.begin(), idx.end(), begin);
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
    const uint64_t la = s.name_off[a + 1] - s.name_off[a], lb = s.name_off[b + 1] - s.name_off[b];
    const int c = memcmp(s.names.data() + s.name_off[a], s.names.data() + s.name_off[b], std::min(la, lb));
    return c < 0 || (c == 0 && la < lb);
  });
  return idx;
}

// /^(\S+)\/\d$/ of the merge script: the core of an accession.
std::string core_of(const std::string& acc) {
  const size_t n = acc.size();
  if (n >= 3 && acc[n - 2] == '/' && isdigit((unsigned char)acc[n - 1])) {
    for (size_t i = 0; i + 2 < n; i++)
      if (perl_space(acc[i])) return acc;
    return acc.substr(0, n - 2);
  }
  return acc;
}

struct Plan {
  std::vector<uint64_t> a, b;  // units: reads a[u] (and b[u] for pairs)
};

// nbkc_merge_left_right_stats.pl --sorted over the sorted left and right
// reads (indices into `both`).
Plan merge_pairs(const SeqSet& both, const std::vector<uint64_t>& L, const std::vector<uint64_t>& R) {
  Plan p;
  size_t i = 0, j = 0;
  while (i < L.size() && j < R.size()) {
    const std::string la = both.name(L[i]), ra = both.name(R[j]);
    if (core_of(la) != core_of(ra)) {
      if (la < ra) i++;  // Perl's `lt` on byte strings
      else j++;
      continue;
    }
    p.a.push_back(L[i++]);
    p.b.push_back(R[j++]);
  }
  return p;
}

void append(SeqSet& dst, const SeqSet& src) {
  const uint64_t base = dst.seq.size(), nbase = dst.names.size();
  dst.seq += src.seq;
  dst.names += src.names;
  for (size_t i = 1; i < src.off.size(); i++) dst.off.push_back(base + src.off[i]);
  for (size_t i = 1; i < src.name_off.size(); i++) dst.name_off.push_back(nbase + src.name_off[i]);
}

bool file_nonempty(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && st.st_size > 0;
}

void touch(const std::string& p) {
  FILE* f = fopen(p.c_str(), "a");
  if (!f) die("Error, cannot touch %s", p.c_str());
  fclose(f);
}

void copy_file(const std::string& from, const std::string& to) {
  TextSource src(from);
  const std::string tmp = to + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) die("Error, cannot write %s", tmp.c_str());
  const bool ok = src.size() == 0 || fwrite(src.data(), 1, src.size(), f) == src.size();
  if (fclose(f) != 0 || !ok) die("Error, cannot write %s", tmp.c_str());
  if (rename(tmp.c_str(), to.c_str()) != 0) die("Error, cannot rename %s", tmp.c_str());
}

// `ln -sf target link` (or cp with NO_SYMLINK).
void link_output(const std::string& target, const std::string& link, bool no_symlink) {
  if (no_symlink) {
    copy_file(target, link);
    return;
  }
  unlink(link.c_str());
  if (symlink(target.c_str(), link.c_str()) != 0) die("Error, cannot link %s to %s", link.c_str(), target.c_str());
}

void write_text(const std::string& path, const std::string& text) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) die("Error, cannot write %s", path.c_str());
  const bool ok = text.empty() || fwrite(text.data(), 1, text.size(), f) == text.size();
  if (fclose(f) != 0 || !ok) die("Error, cannot write %s", path.c_str());
}

// make_normalized_reads_file (:505-583) for one side.
void extract_side(const NormOpts& o, const std::vector<std::string>& sources, const std::string& accs,
                  const std::string& outfile, std::string& log) {
  SelectedIndex idx = build_selected_index_text(accs);
  const std::string tmp = outfile;  // the script writes the output in place, then checks
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) die("Error, cannot write to %s", outfile.c_str());
  {
    Out out(f);
    for (const std::string& s : sources) {
      log += "-capturing normalized reads from: " + s + "\n";
      extract_records(o.seq_type, s, idx, out);
    }
    out.flush();
  }
  if (fclose(f) != 0) die("Error, cannot write to %s", outfile.c_str());
  const std::vector<std::string> missing = missing_accessions(idx);
  if (!missing.empty()) {
    std::string m;
    for (const std::string& s : missing) m += s + "\n";
    write_text(outfile + ".missing_accs", m);
    std::string all;
    for (const std::string& s : sources) all += (all.empty() ? "" : " ") + s;
    die("Error, not all specified records have been retrieved (missing %zu) from %s, see file: %s.missing_accs "
        "for list of missing entries",
        missing.size(), all.c_str(), outfile.c_str());
  }
}

// The whole job; throws (die) on any failure.  `srv` runs the device part.
void run_normalization(NormOpts o, tk_server* srv, std::string& log, std::string* plan_out) {
  validate(o);
  const bool paired = !o.left.empty() && !o.right.empty();
  std::string lt, rt;  // per-side strand types
  if (!o.ss.empty()) {
    lt = o.ss.substr(0, 1);
    rt = o.ss.size() > 1 ? o.ss.substr(1, 1) : std::string();
  }
  // 1. read preparation
  SeqSet both;
  both.clear();
  uint64_t n_left = 0;
  if (paired) {
    log += "-prepping seqs\n";
    SeqSet ls, rs;
    std::string llog, rlog;
    std::exception_ptr lerr, rerr;
    // both directions in parallel (:293-313)
    std::thread tl([&] {
      die_throws(true);
      try {
        ls = prep_side(o, o.left, "1", lt == "R", llog);
      } catch (...) {
        lerr = std::current_exception();
      }
    });
    try {
      rs = prep_side(o, o.right, "2", rt == "R", rlog);
    } catch (...) {
      rerr = std::current_exception();
    }
    tl.join();
    if (lerr) std::rethrow_exception(lerr);
    if (rerr) std::rethrow_exception(rerr);
    log += llog + rlog;
    n_left = ls.size();
    append(both, ls);
    append(both, rs);
  } else {
    log += "-prepping seqs\n";
    SeqSet ss = prep_side(o, o.single, "1", o.ss == "R", log);
    n_left = ss.size();
    append(both, ss);
  }
  // 3. selection units, in the sorted stats' order
  tk_server_job job{};
  std::vector<Plan> plans;
  if (paired) {
    const std::vector<uint64_t> L = sorted_order(both, 0, n_left);
    const std::vector<uint64_t> R = sorted_order(both, n_left, both.size());
    if (o.pairs_together) {
      // an empty merge still leaves the header line, so the script's
      // emptiness check (:962-964) does not fire: no pair, nothing selected
      plans.push_back(merge_pairs(both, L, R));
    } else {
      plans.push_back(Plan{L, {}});
      plans.push_back(Plan{R, {}});
    }
  } else {
    plans.push_back(Plan{sorted_order(both, 0, both.size()), {}});
  }
  if (plan_out) {
    for (size_t q = 0; q < plans.size(); q++) {
      *plan_out += "# selection " + std::to_string(q) + "\n";
      for (size_t u = 0; u < plans[q].a.size(); u++) {
        *plan_out += both.name(plans[q].a[u]);
        if (!plans[q].b.empty()) *plan_out += "\t" + both.name(plans[q].b[u]);
        *plan_out += "\n";
      }
    }
    return;
  }
  // output directory and the script's tmp directory
  if (mkdir(o.output.c_str(), 0777) != 0 && errno != EEXIST) die("Error, cannot mkdir %s", o.output.c_str());
  const std::string tmpdir = o.output + "/" + o.tmp_dir_name;
  bool created_tmp = false;
  if (mkdir(tmpdir.c_str(), 0777) == 0) created_tmp = true;
  else if (errno != EEXIST) die("Error, cannot mkdir %s", tmpdir.c_str());
  // 2. the device job
  log += "-kmer counting, stats and selection on the GPU (" + std::to_string(both.size()) + " reads)\n";
  std::vector<std::vector<uint8_t>> selected(plans.size());
  job.seq = both.seq.data();
  job.offsets = both.off.data();
  job.n_reads = both.size();
  job.k = (int32_t)o.k;
  job.canonical = o.ss.empty() ? 1 : 0;
  job.max_cov = (int32_t)o.max_cov;
  job.min_cov = (int32_t)o.min_cov;
  job.max_cv = (int32_t)o.max_cv;
  job.n_selections = (int32_t)plans.size();
  for (size_t q = 0; q < plans.size(); q++) {
    selected[q].assign(plans[q].a.size(), 0);
    job.sel[q].left = plans[q].a.data();
    job.sel[q].right = plans[q].b.empty() ? nullptr : plans[q].b.data();
    job.sel[q].n_units = plans[q].a.size();
    job.sel[q].selected = selected[q].data();
  }
  uint64_t id = 0;
  check(tk_server_submit_job(srv, &job, &id), "tk_server_submit_job");
  tk_job_result res{};
  check(tk_server_wait(srv, id, &res), "tk_server_wait");
  if (res.status != TK_OK) die("normalization job failed (%d): %s", res.status, res.message);
  char buf[256];
  snprintf(buf, sizeof buf, "-device job: %.3f s (queued %.3f s), %llu units selected\n", res.run_s, res.wait_s,
           (unsigned long long)res.n_selected);
  log += buf;
  // the selected accessions as nbkc_normalize.pl prints them: the merged
  // rows' core names, or each stats row's accession
  std::vector<std::string> accs(plans.size());
  for (size_t q = 0; q < plans.size(); q++)
    for (size_t u = 0; u < plans[q].a.size(); u++)
      if (selected[q][u]) {
        const std::string a = both.name(plans[q].a[u]);
        accs[q] += (plans[q].b.empty() ? a : core_of(a)) + "\n";
      }
  if (o.no_cleanup) {
    const std::string tag = "maxC" + std::to_string(o.max_cov) + ".minC" + std::to_string(o.min_cov) + ".maxCV" +
                            std::to_string(o.max_cv) + ".accs";
    if (o.pairs_together) {
      write_text(tmpdir + "/pairs.K" + std::to_string(o.k) + ".stats.C" + std::to_string(o.max_cov) + ".maxCV" +
                     std::to_string(o.max_cv) + ".accs",
                 accs[0]);
    } else if (paired) {
      write_text(tmpdir + "/left.fa.K" + std::to_string(o.k) + ".stats.sort." + tag, accs[0]);
      write_text(tmpdir + "/right.fa.K" + std::to_string(o.k) + ".stats.sort." + tag, accs[1]);
    } else {
      write_text(tmpdir + "/single.fa.K" + std::to_string(o.k) + ".stats.sort." + tag, accs[0]);
    }
  }
  // 4. extraction, one output per side (in parallel, as the script's threads)
  log += "-search and capture.\n";
  struct Side {
    const std::vector<std::string>* sources;
    const std::string* accs;
    std::string outfile, log;
    std::exception_ptr err;
  };
  std::vector<Side> sides;
  auto outname = [&](const std::vector<std::string>& src) {
    const std::string base = src.size() == 1 ? basename_of(src[0]) : basename_of(src[0]) + "_ext_all_reads";
    return o.output + "/" + base + ".normalized_K" + std::to_string(o.k) + "_maxC" + std::to_string(o.max_cov) +
           "_minC" + std::to_string(o.min_cov) + "_maxCV" + std::to_string(o.max_cv) + "." + o.seq_type;
  };
  if (paired) {
    sides.push_back({&o.left, &accs[0], outname(o.left), "", nullptr});
    sides.push_back({&o.right, &accs[o.pairs_together ? 0 : 1], outname(o.right), "", nullptr});
  } else {
    sides.push_back({&o.single, &accs[0], outname(o.single), "", nullptr});
  }
  std::vector<std::thread> th;
  for (Side& s : sides) {
    if (access((s.outfile + ".ok").c_str(), F_OK) == 0) continue;  // done by an earlier run (:405-412)
    th.emplace_back([&o, &s] {
      die_throws(true);
      try {
        extract_side(o, *s.sources, *s.accs, s.outfile, s.log);
      } catch (...) {
        s.err = std::current_exception();
      }
    });
  }
  for (auto& t : th) t.join();
  for (Side& s : sides) {
    log += s.log;
    if (s.err) std::rethrow_exception(s.err);
  }
  for (Side& s : sides)
    if (file_nonempty(s.outfile) && access((s.outfile + ".ok").c_str(), F_OK) != 0) touch(s.outfile + ".ok");
  if (paired) {
    link_output(sides[0].outfile, o.output + "/left.norm." + o.seq_type, o.no_symlink);
    link_output(sides[1].outfile, o.output + "/right.norm." + o.seq_type, o.no_symlink);
  } else {
    link_output(sides[0].outfile, o.output + "/single.norm." + o.seq_type, o.no_symlink);
  }
  if (!o.no_cleanup && created_tmp) {
    log += "-removing tmp dir " + tmpdir + "\n";
    rmdir(tmpdir.c_str());
  }
  log += "\n\nNormalization complete. See outputs: \n";
  for (Side& s : sides) log += "\t" + s.outfile + "\n";
}

tk_server* make_server(int slots, double budget_gb) {
  tk_server_opts so{};
  so.device = device_from_env();
  so.slots = slots;
  so.k = 25;
  so.min_count = 2;  // dump -L 2 (MIN_KMER_COV_CONST, :45)
  so.max_cov = 50;   // the per-job value is used
  so.max_cv = 10000;
  so.max_load = 0.5;
  so.hbm_budget_bytes = (uint64_t)(budget_gb * (double)(1ull << 30));
  tk_server* s = nullptr;
  check(tk_server_create(&so, &s), "tk_server_create");
  return s;
}

// ------------------------------------------------------------ the socket protocol
// request:  "TKNR1\n", u32 n, n x (u32 len, bytes): cwd, NO_SYMLINK ("0"/"1"), args...
// response: i32 status, u32 len, log bytes

// (send with MSG_NOSIGNAL: a peer that went away is an error here, not a
// SIGPIPE that would end the service)
bool write_all(int fd, const void* p, size_t n) {
  const char* c = (const char*)p;
  while (n) {
    const ssize_t w = send(fd, c, n, MSG_NOSIGNAL);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return false;
    c += w;
    n -= (size_t)w;
  }
  return true;
}

bool read_all_fd(int fd, void* p, size_t n) {
  char* c = (char*)p;
  while (n) {
    const ssize_t r = read(fd, c, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    c += r;
    n -= (size_t)r;
  }
  return true;
}

bool send_strings(int fd, const std::vector<std::string>& v) {
  if (!write_all(fd, "TKNR1\n", 6)) return false;
  const uint32_t n = (uint32_t)v.size();
  if (!write_all(fd, &n, 4)) return false;
  for (const std::string& s : v) {
    const uint32_t l = (uint32_t)s.size();
    if (!write_all(fd, &l, 4) || !write_all(fd, s.data(), s.size())) return false;
  }
  return true;
}

bool recv_strings(int fd, std::vector<std::string>& v) {
  char magic[6];
  if (!read_all_fd(fd, magic, 6) || memcmp(magic, "TKNR1\n", 6) != 0) return false;
  uint32_t n = 0;
  if (!read_all_fd(fd, &n, 4) || n > 1000000) return false;
  v.resize(n);
  for (auto& s : v) {
    uint32_t l = 0;
    if (!read_all_fd(fd, &l, 4) || l > (64u << 20)) return false;
    s.resize(l);
    if (l && !read_all_fd(fd, &s[0], l)) return false;
  }
  return true;
}

int connect_socket(const std::string& path) {
  sockaddr_un a{};
  if (path.size() >= sizeof(a.sun_path)) return -1;
  a.sun_family = AF_UNIX;
  memcpy(a.sun_path, path.c_str(), path.size() + 1);
  const int fd = socket(AF_UNIX, SOCK_STREAM, 0);
  if (fd < 0) return -1;
  if (connect(fd, (sockaddr*)&a, sizeof a) != 0) {
    close(fd);
    return -1;
  }
  return fd;
}

// $ENV{NO_SYMLINK} true in Perl's sense: set, not "" and not "0".
bool no_symlink_env() {
  const char* e = getenv("NO_SYMLINK");
  return e && *e && strcmp(e, "0") != 0;
}

std::string current_dir() {
  char buf[8192];
  if (!getcwd(buf, sizeof buf)) die("Error, cannot read the working directory");
  return buf;
}

// Client: the job runs in the service; its log is printed here.  -1 when the
// service cannot be reached.
int run_client(const std::string& sock, const std::vector<std::string>& args) {
  const int fd = connect_socket(sock);
  if (fd < 0) return -1;
  std::vector<std::string> req{current_dir(), no_symlink_env() ? "1" : "0"};
  req.insert(req.end(), args.begin(), args.end());
  int32_t status = 1;
  uint32_t len = 0;
  std::string log;
  const bool ok = send_strings(fd, req) && read_all_fd(fd, &status, 4) && read_all_fd(fd, &len, 4);
  if (ok) {
    log.resize(len);
    if (len && !read_all_fd(fd, &log[0], len)) status = 1;
  } else {
    log = "Error: the normalization service closed the connection\n";
    status = 1;
  }
  close(fd);
  fputs(log.c_str(), stderr);
  return status;
}

struct Service {
  tk_server* srv = nullptr;
  int listen_fd = -1;
  std::string path;
  std::mutex mu;
  std::condition_variable cv;
  int active = 0;
  std::atomic<bool> stop{false};
};

void serve_one(Service* sv, int fd) {
  die_throws(true);
  // a client that connects and sends nothing must not hold the service
  // (shutdown waits for every job thread): reads give up after 60 s
  timeval tv{60, 0};
  setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &tv, sizeof tv);
  std::vector<std::string> req;
  int32_t status = 0;
  std::string log;
  if (!recv_strings(fd, req) || req.size() < 2) {
    status = 2;
    log = "Error: malformed request\n";
  } else if (req.size() == 3 && req[2] == "--shutdown") {
    sv->stop = true;
    shutdown(sv->listen_fd, SHUT_RDWR);
    log = "-normalization service stopping\n";
  } else {
    try {
      NormOpts o = parse_norm_args(std::vector<std::string>(req.begin() + 2, req.end()));
      o.cwd = req[0];
      o.no_symlink = req[1] == "1";
      std::string plan;
      run_normalization(o, sv->srv, log, o.plan_only ? &plan : nullptr);
      log += plan;
    } catch (const std::exception& e) {
      log += (strncmp(e.what(), "Error", 5) == 0 ? std::string() : std::string("Error: ")) + e.what() + "\n";
      status = 1;
    }
  }
  const uint32_t len = (uint32_t)log.size();
  (void)(write_all(fd, &status, 4) && write_all(fd, &len, 4) && write_all(fd, log.data(), log.size()));
  close(fd);
  // notified under the lock: once it is released the service may end, and
  // this thread touches nothing of it after that
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->active--;
  sv->cv.notify_all();
}

int run_service(const std::string& path, int slots, double budget_gb) {
  Service sv;
  sv.path = path;
  sockaddr_un a{};
  if (path.size() >= sizeof(a.sun_path)) die("socket path too long: %s", path.c_str());
  struct stat st;
  if (lstat(path.c_str(), &st) == 0) {
    if (!S_ISSOCK(st.st_mode)) die("%s exists and is not a socket", path.c_str());
    if (connect_socket(path) >= 0) die("a service already listens on %s", path.c_str());
    unlink(path.c_str());
  }
  a.sun_family = AF_UNIX;
  memcpy(a.sun_path, path.c_str(), path.size() + 1);
  sv.listen_fd = socket(AF_UNIX, SOCK_STREAM, 0);
  if (sv.listen_fd < 0 || bind(sv.listen_fd, (sockaddr*)&a, sizeof a) != 0 || listen(sv.listen_fd, 64) != 0)
    die("cannot listen on %s: %s", path.c_str(), strerror(errno));
  sv.srv = make_server(slots, budget_gb);
  fprintf(stderr, "-normalization service on %s (%d slots, %.0f GB device budget)\n", path.c_str(), slots,
          budget_gb);
  fflush(stderr);
  while (!sv.stop) {
    const int fd = accept(sv.listen_fd, nullptr, nullptr);
    if (fd < 0) {
      if (errno == EINTR) continue;
      break;  // shut down
    }
    {
      std::lock_guard<std::mutex> lk(sv.mu);
      sv.active++;
    }
    std::thread(serve_one, &sv, fd).detach();
  }
  {
    std::unique_lock<std::mutex> lk(sv.mu);
    sv.cv.wait(lk, [&] { return sv.active == 0; });
  }
  close(sv.listen_fd);
  unlink(path.c_str());
  tk_server_free(sv.srv);
  fprintf(stderr, "-normalization service stopped\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  std::string serve, server;
  long long slots = 4;
  double budget = 96;
  bool shutdown_req = false;
  std::vector<std::string> rest;
  for (size_t i = 0; i < args.size(); i++) {
    const std::string& a = args[i];
    auto val = [&](const char* what) -> std::string {
      if (i + 1 >= args.size()) die("option %s needs a value", what);
      return args[++i];
    };
    if (a == "--serve") serve = val("--serve");
    else if (a == "--server") server = val("--server");
    else if (a == "--slots") slots = atoll(val("--slots").c_str());
    else if (a == "--hbm_budget_gb") budget = atof(val("--hbm_budget_gb").c_str());
    else if (a == "--shutdown") shutdown_req = true;
    else rest.push_back(a);
  }
  if (args.empty()) {
    fputs(usage().c_str(), stderr);
    return 1;
  }
  if (!serve.empty()) {
    if (slots < 1 || slots > 16) die("--slots must be in [1,16]");
    if (!(budget > 0)) die("--hbm_budget_gb must be positive");
    return run_service(serve, (int)slots, budget);
  }
  const bool explicit_server = !server.empty();
  if (server.empty() && getenv("TRINITY_GPU_NORM_SERVER")) server = getenv("TRINITY_GPU_NORM_SERVER");
  if (shutdown_req) {
    if (server.empty()) die("--shutdown needs --server <socket>");
    const int rc = run_client(server, {"--shutdown"});
    if (rc < 0) die("no normalization service at %s", server.c_str());
    return rc;
  }
  if (!server.empty()) {
    const int rc = run_client(server, rest);
    if (rc >= 0) return rc;
    if (explicit_server) die("no normalization service at %s", server.c_str());
    fprintf(stderr, "-no normalization service at %s: running the job in this process\n", server.c_str());
  }
  NormOpts o = parse_norm_args(rest);
  o.cwd = current_dir();
  o.no_symlink = no_symlink_env();
  std::string log, plan;
  tk_server* srv = nullptr;
  try {
    if (o.plan_only) {
      run_normalization(o, nullptr, log, &plan);
      fputs(log.c_str(), stderr);
      fputs(plan.c_str(), stdout);
      return 0;
    }
    die_throws(true);
    validate(o);  // option errors end the run before the device is touched
    srv = make_server(1, 0);
    run_normalization(o, srv, log, nullptr);
    die_throws(false);
  } catch (const std::exception& e) {
    die_throws(false);
    fputs(log.c_str(), stderr);
    if (srv) tk_server_free(srv);
    // messages that carry the script's own "Error, ..." wording print as is
    if (strncmp(e.what(), "Error", 5) == 0) fprintf(stderr, "%s\n", e.what());
    else fprintf(stderr, "Error: %s\n", e.what());
    return 1;
  }
  tk_server_free(srv);
  fputs(log.c_str(), stderr);
  return 0;
}
