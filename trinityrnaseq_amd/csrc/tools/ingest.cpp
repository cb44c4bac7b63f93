// ingest.cpp -- FASTQ -> FASTA conversion and FASTA reverse complement
// (ingest.h).  Record rules of the conversion (fastQ_to_fastA.pl:66-117): a
// line starting with '@' opens a 4-line record; every \r of that header line
// is deleted; the record is kept when header, sequence, '+' line and a
// non-blank quality line of the sequence's length are all there, else it is
// "dirty".  The name is the first whitespace-separated field of the header
// after '@'; when more fields follow, the name does not end in /1 or /2 and
// the next field starts with "1:" or "2:" (Casava 1.8), "/1" or "/2" is
// appended; -a N appends "/N" (seqtk-trinity's -R N only to names without a
// /1,/2 suffix); --rev reverse-complements the sequence with
// Nuc_translator.pm:286-292's table (seqtk-trinity: the IUPAC table of
// revcomp_fasta.pl).
#include "ingest.h"

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "seq_io.h"

namespace tools {

namespace {

// Perl's \s on byte strings: space, \t, \n, \v, \f, \r.
inline bool perl_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

// Perl string truthiness: false for "" and "0".
inline bool perl_true(const std::string& s) { return !(s.empty() || s == "0"); }

struct Lines {
  const char* p;
  uint64_t n, pos = 0;
  Lines(const char* d, uint64_t sz) : p(d), n(sz) {}
  // Next line with its newline (Perl's <$fh>); false at EOF.
  bool next(std::string& out) {
    if (pos >= n) return false;
    const char* e = (const char*)memchr(p + pos, '\n', n - pos);
    const uint64_t end = e ? (uint64_t)(e - p) + 1 : n;
    out.assign(p + pos, end - pos);
    pos = end;
    return true;
  }
};

void chomp(std::string& s) {
  if (!s.empty() && s.back() == '\n') s.pop_back();
}

char rc_nuc_translator(char c) {  // tr/ACGTacgtyrkmYRKMUu/TGCAtgcarymkRYMKAa/
  switch (c) {
    case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; case 'T': return 'A';
    case 'a': return 't'; case 'c': return 'g'; case 'g': return 'c'; case 't': return 'a';
    case 'y': return 'r'; case 'r': return 'y'; case 'k': return 'm'; case 'm': return 'k';
    case 'Y': return 'R'; case 'R': return 'Y'; case 'K': return 'M'; case 'M': return 'K';
    case 'U': return 'A'; case 'u': return 'a';
    default: return c;
  }
}

char rc_iupac(char c) {  // tr/ACGTUMRWSYKVHDBNacgtumrwsykvhdbn/TGCAAKYWSRMBDHVNtgcaakywsrmbdhvn/
  static const char* from = "ACGTUMRWSYKVHDBNacgtumrwsykvhdbn";
  static const char* to = "TGCAAKYWSRMBDHVNtgcaakywsrmbdhvn";
  const char* f = strchr(from, c);
  return (f && c) ? to[f - from] : c;
}

// Split the header (after '@') like Perl's split(/\s+/, $h): a leading empty
// field when it starts with whitespace, trailing empty fields dropped.
std::vector<std::string> perl_split_ws(const std::string& h) {
  std::vector<std::string> f;
  std::string cur;
  for (size_t i = 0; i < h.size();) {
    if (perl_space(h[i])) {
      f.push_back(cur);
      cur.clear();
      while (i < h.size() && perl_space(h[i])) i++;
    } else {
      cur.push_back(h[i++]);
    }
  }
  f.push_back(cur);
  while (!f.empty() && f.back().empty()) f.pop_back();
  return f;
}

bool ends_pair_suffix(const std::string& s) {
  return s.size() >= 2 && s[s.size() - 2] == '/' && (s.back() == '1' || s.back() == '2');
}

}  // namespace

void fastq_to_fasta(const std::string& path, const ConvertOpts& o, Out& out, ConvertTally& t) {
  TextSource src(path);
  Lines in(src.data(), src.size());
  std::string line, header, seq, qh, ql;
  while (in.next(line)) {
    std::string l;
    l.reserve(line.size());
    for (char c : line)
      if (c != '\r') l.push_back(c);
    if (l.empty() || l[0] != '@') continue;
    header = l;
    const bool hs = in.next(seq), hq = in.next(qh), hl = in.next(ql);
    chomp(header);
    if (hs) chomp(seq);
    if (hq) chomp(qh);
    if (hl) chomp(ql);
    bool ql_nonblank = false;
    if (hl)
      for (char c : ql)
        if (!perl_space(c)) { ql_nonblank = true; break; }
    const bool ok = perl_true(header) && hs && perl_true(seq) && hq && perl_true(qh) && ql_nonblank &&
                    qh[0] == '+' && seq.size() == ql.size();
    if (!ok) {
      t.dirty++;
      if (!o.ignore_dirty) {
        out.flush();
        throw DirtyRecord("Error, improperly formatted entry:\n\n" + header + "\n" + (hs ? seq : std::string()) +
                          "\n" + (hq ? qh : std::string()) + "\n" + (hl ? ql : std::string()) + "\n");
      }
      continue;
    }
    std::vector<std::string> parts = perl_split_ws(header.substr(1));
    std::string name = parts.empty() ? std::string() : parts[0];
    if (parts.size() > 1 && !ends_pair_suffix(name) && parts[1].size() >= 2 &&
        (parts[1][0] == '1' || parts[1][0] == '2') && parts[1][1] == ':') {
      name += '/';
      name += parts[1][0];
    }
    if (!o.append.empty() && !(o.seqtk && ends_pair_suffix(name))) {
      name += '/';
      name += o.append;
    }
    if (o.rev) {
      std::string r(seq.rbegin(), seq.rend());
      for (char& c : r) c = o.seqtk ? rc_iupac(c) : rc_nuc_translator(c);
      seq.swap(r);
    }
    out.putc('>');
    out.put(name);
    out.putc('\n');
    out.put(seq);
    out.putc('\n');
    t.clean++;
  }
}


void revcomp_fasta(const std::string& path, Out& out) {
  TextSource src(path);
  const char* p = src.data();
  const uint64_t n = src.size();
  uint64_t pos = 0;
  std::string rec, seq;
  while (pos < n) {
    // `local $/ = ">"; chomp`: the record up to and without the next '>'
    const char* e = (const char*)memchr(p + pos, '>', n - pos);
    const uint64_t end = e ? (uint64_t)(e - p) : n;
    rec.assign(p + pos, end - pos);
    pos = e ? end + 1 : n;
    const size_t nl = rec.find('\n');
    // substr($_, 0, index(...)): index -1 (no newline) takes all but the last byte
    const std::string header = nl == std::string::npos ? rec.substr(0, rec.empty() ? 0 : rec.size() - 1)
                                                       : rec.substr(0, nl);
    if (header.empty() || header == "0") continue;
    // substr($_, -1) when there is no newline: the last byte
    const std::string body = nl == std::string::npos ? rec.substr(rec.size() - 1) : rec.substr(nl);
    seq.clear();
    for (auto it = body.rbegin(); it != body.rend(); ++it) {
      const char c = *it;
      if (c == '\n') continue;
      seq.push_back(rc_iupac(c));
    }
    out.putc('>');
    out.put(header);
    out.putc('\n');
    for (size_t i = 0; i < seq.size(); i += 60) {
      out.put(seq.data() + i, std::min<size_t>(60, seq.size() - i));
      out.putc('\n');
    }
  }
}

}  // namespace tools
