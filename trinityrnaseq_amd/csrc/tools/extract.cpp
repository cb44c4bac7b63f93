// extract.cpp -- record capture of make_normalized_reads_file (extract.h).
#include "extract.h"

#include <stdio.h>
#include <string.h>
#include <ctype.h>

#include <algorithm>
#include <string>
#include <vector>

#include "seq_io.h"

namespace tools {

namespace {

inline bool perl_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
inline bool word_char(char c) { return isalnum((unsigned char)c) || c == '_'; }

void erase_first(std::string& s, const char* what) {
  const size_t i = s.find(what);
  if (i != std::string::npos) s.erase(i, strlen(what));
}

// First field of Perl's split(/\s+/, $s): empty when $s starts with blanks.
std::string first_field(const std::string& s) {
  size_t i = 0;
  while (i < s.size() && !perl_space(s[i])) i++;
  return s.substr(0, i);
}

void extract_fq(const std::string& path, SelectedIndex& idx, Out& out) {
  TextSource src(path);
  const char* p = src.data();
  const uint64_t n = src.size();
  uint64_t pos = 0;
  while (pos < n) {  // `if (! eof)`: read up to four lines
    const uint64_t start = pos;
    uint64_t nl[4];
    int got = 0;
    while (got < 4 && pos < n) {
      const char* e = (const char*)memchr(p + pos, '\n', n - pos);
      pos = e ? (uint64_t)(e - p) + 1 : n;
      nl[got++] = pos;
    }
    std::string rec(p + start, pos - start);
    // split(/\n/, $text) must give 4 fields (trailing empty fields dropped)
    std::vector<std::string> f;
    {
      std::string cur;
      for (char c : rec) {
        if (c == '\n') {
          f.push_back(cur);
          cur.clear();
        } else {
          cur.push_back(c);
        }
      }
      f.push_back(cur);
      while (!f.empty() && f.back().empty()) f.pop_back();
    }
    (void)nl;
    if (f.size() != 4) die("fastQ entry doesn't have 4 lines: %s (%s)", rec.c_str(), path.c_str());
    if (f[0].empty() || f[0][0] != '@')
      die("cannot identify first line as read name line: %s (%s)", rec.c_str(), path.c_str());
    std::string name = first_field(f[0]);
    if (!name.empty() && name[0] == '@') name.erase(0, 1);
    // /^(\S+)\/([12])$/
    if (name.size() >= 3 && name[name.size() - 2] == '/' && (name.back() == '1' || name.back() == '2')) {
      bool nonspace = true;
      for (size_t i = 0; i + 2 < name.size(); i++)
        if (perl_space(name[i])) nonspace = false;
      if (nonspace) name.resize(name.size() - 2);
    }
    erase_first(name, "_forward");
    erase_first(name, "_reverse");
    auto it = idx.find(name);
    if (it != idx.end()) {
      it->second++;
      out.put(rec);
    }
  }
}

void extract_fa(const std::string& path, SelectedIndex& idx, Out& out) {
  // Fasta_reader semantics: records separated by "\n>", control bytes other
  // than \t and \n deleted, header = first line, sequence = the other lines
  // joined with all whitespace removed.
  TextSource src(path);
  const char* p = src.data();
  const uint64_t n = src.size();
  uint64_t pos = 0;
  // one "\n>"-terminated chunk, as `$/ = "\n>"; <$fh>` returns it
  auto next_chunk = [&](std::string& rec) -> bool {
    if (pos >= n) return false;
    const uint64_t start = pos;
    uint64_t end = n;
    for (uint64_t i = pos; i + 1 < n; i++)
      if (p[i] == '\n' && p[i + 1] == '>') {
        end = i + 2;
        break;
      }
    pos = end;
    rec.assign(p + start, end - start);
    return true;
  };
  auto has_word = [](const std::string& r) {
    for (char c : r)
      if (word_char(c)) return true;
    return false;
  };
  std::string rec;
  while (next_chunk(rec)) {
    // a chunk with no word character (blank lines before a record) is
    // replaced by the next one; at EOF the reader returns nothing
    if (!has_word(rec) && !next_chunk(rec)) break;
    // s/^>|>$//g
    if (!rec.empty() && rec[0] == '>') rec.erase(0, 1);
    if (!rec.empty() && rec.back() == '>') rec.pop_back();
    std::string clean;
    clean.reserve(rec.size());
    for (char c : rec) {
      const unsigned char u = (unsigned char)c;
      if (c == '\t' || c == '\n' || (u > 31 && u < 127)) clean.push_back(c);
    }
    size_t h = clean.find('\n');
    std::string header = clean.substr(0, h);
    std::string seq;
    if (h != std::string::npos)
      for (size_t i = h + 1; i < clean.size(); i++)
        if (!perl_space(clean[i])) seq.push_back(clean[i]);
    std::string acc = first_field(header);
    // s|/[12]\s*$||
    size_t e = acc.size();
    while (e > 0 && perl_space(acc[e - 1])) e--;
    if (e >= 2 && acc[e - 2] == '/' && (acc[e - 1] == '1' || acc[e - 1] == '2')) acc.resize(e - 2);
    auto it = idx.find(acc);
    if (it != idx.end()) {
      it->second++;
      out.putc('>');
      out.put(header);
      out.putc('\n');
      out.put(seq);
      out.putc('\n');
    }
  }
}

}  // namespace

SelectedIndex build_selected_index_text(const std::string& text) {
  SelectedIndex idx;
  const char* p = text.data();
  const uint64_t n = text.size();
  uint64_t pos = 0;
  while (pos < n) {
    const char* e = (const char*)memchr(p + pos, '\n', n - pos);
    const uint64_t end = e ? (uint64_t)(e - p) : n;
    std::string line(p + pos, end - pos);
    pos = e ? end + 1 : n;
    bool nonblank = false;
    for (char c : line)
      if (!perl_space(c)) nonblank = true;
    if (!nonblank) continue;
    // s|/\w$||: $ also matches before a final newline, which chomp removed
    if (line.size() >= 2 && line[line.size() - 2] == '/' && word_char(line.back())) line.resize(line.size() - 2);
    idx.emplace(std::move(line), 0);
  }
  return idx;
}

SelectedIndex build_selected_index(const std::string& path) {
  TextSource src(path);
  return build_selected_index_text(std::string(src.data(), src.size()));
}

void extract_records(const std::string& seq_type, const std::string& path, SelectedIndex& idx, Out& out) {
  if (seq_type == "fq") extract_fq(path, idx, out);
  else if (seq_type == "fa") extract_fa(path, idx, out);
  else die("Error, do not recognize format: %s", seq_type.c_str());
}

std::vector<std::string> missing_accessions(const SelectedIndex& idx) {
  std::vector<std::string> m;
  for (const auto& kv : idx)
    if (kv.second == 0) m.push_back(kv.first);
  std::sort(m.begin(), m.end());
  return m;
}

}  // namespace tools
