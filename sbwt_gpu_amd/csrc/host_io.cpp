// Host-side query ingest and result text of libsbwt_gpu_amd.
//
// sbwt_parse_lines replaces KmerSearch::read_file
// (include/epicseq/kmer_search.hpp:319-342): one row per line, lines starting
// with '>' skipped, empty lines skipped. Two behaviours differ on purpose: a
// trailing '\r' is dropped (the reference keeps it as a character), and a
// last line without '\n' keeps its last character (the reference drops it,
// because getline's count then includes no delimiter).
//
// sbwt_parse_records reads FASTA and FASTQ (no reference counterpart; the
// reference reads one sequence per line, so a multi-line record loses the
// k-mers that span its line breaks): one row per record.
//
// sbwt_format_results / sbwt_write_results replace
// ResultsPersister::save_results_to_file (include/epicseq/results_persister.hpp:82-96):
// every answer as a signed decimal (~0 prints as -1) followed by a space, one
// line per row. Rows are formatted in parallel chunks and written in order.
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host_io.hpp"

namespace {

using u64 = uint64_t;

// Formats rows [r0, r1) into out.
void format_rows(const u64* res, const u64* off, u64 r0, u64 r1, std::string& out) {
  char tmp[24];
  for (u64 r = r0; r < r1; r++) {
    for (u64 j = off[r]; j < off[r + 1]; j++) {
      auto p = std::to_chars(tmp, tmp + sizeof tmp, static_cast<long long>(static_cast<int64_t>(res[j])));
      out.append(tmp, p.ptr);
      out.push_back(' ');
    }
    out.push_back('\n');
  }
}

// Splits rows into chunks of roughly equal k-mer counts.
std::vector<u64> chunk_rows(const u64* off, u64 rows, unsigned parts) {
  std::vector<u64> cut{0};
  const u64 total = off[rows] - off[0];
  for (unsigned p = 1; p < parts; p++) {
    const u64 target = off[0] + total * p / parts;
    const u64 r = static_cast<u64>(std::lower_bound(off, off + rows, target) - off);
    if (r > cut.back() && r < rows) cut.push_back(r);
  }
  cut.push_back(rows);
  return cut;
}

std::vector<std::string> format_all(const u64* res, const u64* off, u64 rows) {
  unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if (off[rows] - off[0] < (1u << 16)) hw = 1;
  const auto cut = chunk_rows(off, rows, hw);
  std::vector<std::string> parts(cut.size() - 1);
  std::vector<std::thread> th;
  for (size_t i = 0; i + 1 < cut.size(); i++)
    th.emplace_back([&, i] { format_rows(res, off, cut[i], cut[i + 1], parts[i]); });
  for (auto& t : th) t.join();
  return parts;
}

}  // namespace

namespace sbwt {

namespace {

// Line [b, e) starting at p ('\n' or "\r\n" ends it); returns the next line's start.
u64 line_at(const char* text, u64 len, u64 p, u64& b, u64& e) {
  const char* nl = static_cast<const char*>(memchr(text + p, '\n', len - p));
  const u64 end = nl ? static_cast<u64>(nl - text) : len;
  b = p;
  e = end > p && text[end - 1] == '\r' ? end - 1 : end;
  return end + 1;
}

// One FASTQ record from the record start p (after empty lines): its end, or
// ~0 if it is malformed; seq_out gets its sequence length.
u64 fastq_record(const char* text, u64 len, u64 p, u64* seq_out, const char** why) {
  u64 b = 0, e = 0;
  u64 nx = line_at(text, len, p, b, e);
  while (e == b && nx <= len) {  // empty lines between records
    p = nx;
    if (p >= len) return len;
    nx = line_at(text, len, p, b, e);
  }
  if (text[b] != '@') {
    *why = "does not start with '@'";
    return ~0ull;
  }
  p = nx;
  u64 seq = 0;
  bool plus = false;
  while (p < len) {
    p = line_at(text, len, p, b, e);
    if (e > b && text[b] == '+') {
      plus = true;
      break;
    }
    seq += e - b;
  }
  u64 qual = 0;
  while (plus && qual < seq && p < len) {
    p = line_at(text, len, p, b, e);
    qual += e - b;
  }
  if (!plus || qual != seq) {
    *why = plus ? "quality length differs from the sequence length" : "no '+' line";
    return ~0ull;
  }
  if (seq_out) *seq_out = seq;
  return p < len ? p : len;
}

}  // namespace

TextKind text_kind(const char* text, u64 len) {
  u64 b = 0, e = 0;
  for (u64 q = 0; q < len;) {
    const u64 nx = line_at(text, len, q, b, e);
    if (e > b) return text[b] == '>' ? kTextFasta : (text[b] == '@' ? kTextFastq : kTextLines);
    q = nx;
  }
  return kTextLines;
}

int parse_text(TextKind kind, const char* text, u64 len, uint8_t* chars, u64* row_starts, u64 row_cap, u64* rows_out) {
  if (kind == kTextLines) return sbwt_parse_lines(text, len, chars, row_starts, row_cap, rows_out);
  u64 rows = 0, nch = 0, p = 0, b = 0, e = 0;
  auto open_row = [&]() -> bool {
    if (rows + 1 >= row_cap) return false;
    row_starts[rows++] = nch;
    return true;
  };
  auto full = [&]() {
    return fail(SBWT_ERR_ARG, "sbwt_parse_records: more than %llu records", static_cast<unsigned long long>(row_cap - 1));
  };
  if (kind == kTextFasta) {
    bool in_record = false;
    while (p < len) {
      const u64 nx = line_at(text, len, p, b, e);
      if (e > b && text[b] == '>') {
        if (!open_row()) return full();
        in_record = true;
      } else if (e > b && in_record) {
        memcpy(chars + nch, text + b, e - b);  // sequence lines of the record, joined
        nch += e - b;
      }
      p = nx;
    }
  } else {
    while (p < len) {
      const char* why = "";
      u64 seq = 0;
      const u64 end = fastq_record(text, len, p, &seq, &why);
      if (end == ~0ull)
        return fail(SBWT_ERR_FORMAT, "sbwt_parse_records: FASTQ record %llu %s", static_cast<unsigned long long>(rows + 1),
                    why);
      // empty lines at the end of the text are no record
      u64 q = p;
      while (q < end && (text[q] == '\n' || text[q] == '\r')) q++;
      if (q >= end) break;
      if (!open_row()) return full();
      // the sequence lines: from the header's next line up to the '+' line
      u64 s = line_at(text, len, q, b, e);
      while (s < end) {
        s = line_at(text, len, s, b, e);
        if (e > b && text[b] == '+') break;
        memcpy(chars + nch, text + b, e - b);
        nch += e - b;
      }
      p = end;
    }
  }
  row_starts[rows] = nch;
  *rows_out = rows;
  return SBWT_OK;
}

u64 text_cut(TextKind kind, const char* text, u64 len, u64 from, u64 target, int* err) {
  *err = SBWT_OK;
  if (target >= len) return len;
  u64 b = 0, e = 0;
  if (kind == kTextFastq) {  // whole records, walked from the last cut
    u64 p = from;
    while (p < target && p < len) {
      const char* why = "";
      const u64 end = fastq_record(text, len, p, nullptr, &why);
      if (end == ~0ull) {
        *err = fail(SBWT_ERR_FORMAT, "sbwt_search_text: malformed FASTQ record at byte %llu: %s",
                    static_cast<unsigned long long>(p), why);
        return len;
      }
      p = end;
    }
    return p < len ? p : len;
  }
  // the first line start at or after target
  u64 p = target;
  if (text[p - 1] != '\n') {
    const char* nl = static_cast<const char*>(memchr(text + p, '\n', len - p));
    p = nl ? static_cast<u64>(nl - text) + 1 : len;
  }
  if (kind == kTextFasta)  // ... that starts a record
    while (p < len && text[p] != '>') p = line_at(text, len, p, b, e);
  return p;
}

// Host threads for parsing: the machine's, at most 16.
unsigned parse_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

namespace {

// The rows of the lines in [p, end) (sbwt_parse_lines' rules). With out
// chars / starts NULL it only counts; otherwise row i starts at
// base + (chars written before it).
void lines_range(const char* text, u64 p, u64 end, uint8_t* chars, u64* starts, u64 base, u64* rows_out,
                 u64* chars_out) {
  u64 rows = 0, nch = 0;
  while (p < end) {
    const char* nl = static_cast<const char*>(memchr(text + p, '\n', end - p));
    const u64 e = nl ? static_cast<u64>(nl - text) : end;
    u64 le = e;
    if (le > p && text[le - 1] == '\r') le--;
    if (le > p && text[p] != '>') {
      if (chars) {
        starts[rows] = base + nch;
        memcpy(chars + nch, text + p, le - p);
      }
      rows++;
      nch += le - p;
    }
    p = e + 1;
  }
  *rows_out = rows;
  *chars_out = nch;
}

}  // namespace

int parse_lines_parallel(const char* text, u64 len, uint8_t* chars, u64* row_starts, u64 row_cap, u64* rows_out) {
  const unsigned nt = len < (4u << 20) ? 1u : parse_threads();
  if (row_cap == 0) return fail(SBWT_ERR_ARG, "sbwt_parse_lines: row_cap is 0");
  if (nt == 1) {
    // (checks the room for each row as it opens it, so a small row_cap stops early)
    u64 rows = 0, nch = 0, p = 0;
    while (p < len) {
      const char* nl = static_cast<const char*>(memchr(text + p, '\n', len - p));
      const u64 e = nl ? static_cast<u64>(nl - text) : len;
      u64 le = e;
      if (le > p && text[le - 1] == '\r') le--;
      if (le > p && text[p] != '>') {
        if (rows + 1 >= row_cap)
          return fail(SBWT_ERR_ARG, "sbwt_parse_lines: more than %llu rows", static_cast<unsigned long long>(row_cap - 1));
        row_starts[rows++] = nch;
        memcpy(chars + nch, text + p, le - p);
        nch += le - p;
      }
      p = e + 1;
    }
    row_starts[rows] = nch;
    *rows_out = rows;
    return SBWT_OK;
  }
  // pieces cut at line starts; a line never straddles two pieces
  std::vector<u64> cut(nt + 1, len);
  cut[0] = 0;
  for (unsigned t = 1; t < nt; t++) {
    u64 p = std::max(cut[t - 1], len / nt * t);
    if (p > 0 && p < len && text[p - 1] != '\n') {
      const char* nl = static_cast<const char*>(memchr(text + p, '\n', len - p));
      p = nl ? static_cast<u64>(nl - text) + 1 : len;
    }
    cut[t] = std::min(p, len);
  }
  std::vector<u64> rows(nt), nch(nt);
  auto run = [&](auto&& body) {
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(body, t);
    body(0u);
    for (auto& x : th) x.join();
  };
  run([&](unsigned t) { lines_range(text, cut[t], cut[t + 1], nullptr, nullptr, 0, &rows[t], &nch[t]); });
  std::vector<u64> r0(nt + 1, 0), c0(nt + 1, 0);
  for (unsigned t = 0; t < nt; t++) {
    r0[t + 1] = r0[t] + rows[t];
    c0[t + 1] = c0[t] + nch[t];
  }
  if (r0[nt] + 1 > row_cap)
    return fail(SBWT_ERR_ARG, "sbwt_parse_lines: more than %llu rows", static_cast<unsigned long long>(row_cap - 1));
  run([&](unsigned t) {
    u64 a, b;
    lines_range(text, cut[t], cut[t + 1], chars + c0[t], row_starts + r0[t], c0[t], &a, &b);
  });
  row_starts[r0[nt]] = c0[nt];
  *rows_out = r0[nt];
  return SBWT_OK;
}

}  // namespace sbwt

extern "C" {

int sbwt_parse_lines(const char* text, uint64_t len, uint8_t* chars, uint64_t* row_starts, uint64_t row_cap,
                     uint64_t* rows_out) {
  sbwt::clear_error();
  if ((!text && len) || !chars || !row_starts || !rows_out)
    return sbwt::fail(SBWT_ERR_ARG, "sbwt_parse_lines: NULL argument");
  return sbwt::parse_lines_parallel(text, len, chars, row_starts, row_cap, rows_out);
}

int sbwt_parse_records(const char* text, uint64_t len, uint8_t* chars, uint64_t* row_starts, uint64_t row_cap,
                       uint64_t* rows_out) {
  sbwt::clear_error();
  if ((!text && len) || !chars || !row_starts || !rows_out)
    return sbwt::fail(SBWT_ERR_ARG, "sbwt_parse_records: NULL argument");
  return sbwt::parse_text(sbwt::text_kind(text, len), text, len, chars, row_starts, row_cap, rows_out);
}

int sbwt_format_results(const uint64_t* results, const uint64_t* offsets, uint64_t rows, char* buf, uint64_t cap,
                        uint64_t* size_out) {
  sbwt::clear_error();
  if (!offsets || !size_out || (rows && !results)) return sbwt::fail(SBWT_ERR_ARG, "sbwt_format_results: NULL argument");
  const auto parts = format_all(results, offsets, rows);
  u64 size = 0;
  for (const auto& s : parts) size += s.size();
  *size_out = size;
  if (!buf) return SBWT_OK;
  if (cap < size) return sbwt::fail(SBWT_ERR_ARG, "sbwt_format_results: buffer of %llu bytes, %llu needed",
                                    static_cast<unsigned long long>(cap), static_cast<unsigned long long>(size));
  u64 o = 0;
  for (const auto& s : parts) {
    memcpy(buf + o, s.data(), s.size());
    o += s.size();
  }
  return SBWT_OK;
}

int sbwt_write_results(const uint64_t* results, const uint64_t* offsets, uint64_t rows, int fd) {
  sbwt::clear_error();
  if (!offsets || (rows && !results)) return sbwt::fail(SBWT_ERR_ARG, "sbwt_write_results: NULL argument");
  const auto parts = format_all(results, offsets, rows);
  for (const auto& s : parts) {
    size_t o = 0;
    while (o < s.size()) {
      const ssize_t w = ::write(fd, s.data() + o, s.size() - o);
      if (w <= 0) return sbwt::fail(SBWT_ERR_IO, "writing the results failed");
      o += static_cast<size_t>(w);
    }
  }
  return SBWT_OK;
}

}  // extern "C"
