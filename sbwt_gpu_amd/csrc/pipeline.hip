// Pipelined query path of libsbwt_gpu_amd: query text in host memory ->
// result text, in batches with three in flight. No reference counterpart:
// the reference reads the whole query file, copies it, searches, copies the
// answers back and prints them, one stage after another
// (include/epicseq/kmer_search.hpp:97-153).
//
// Batch b runs on two slots of double-buffered staging:
//   host       parse batch b into pinned slot b % 2 (after batch b - 2's
//              answers were written, which frees the slot);
//   h2d stream copy its characters and row starts to the device;
//   compute    wait for the copy, run the batch's read search on the shared
//              plan (one stream, so batches that share the plan's scratch
//              run in order);
//   compute    ... and print the batch's result text (sbwt_format_text), or
//              make its found bitmap (sbwt_found_bitmap);
//   d2h stream wait for the search, copy the text (or the answers, or the
//              bitmap) to pinned memory;
//   host       (at batch b + 2) wait for that copy, write it.
// So the host's parsing and writing and both copies overlap the searches,
// and the printing, the reference's slowest stage, runs on the GPU.
// Every stream keeps the default priority; pinned memory comes from
// hipHostMalloc.
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "host_io.hpp"
#include "index.hpp"

namespace {

using sbwt::u64;

struct Slot {
  uint8_t* h_chars = nullptr;
  u64* h_starts = nullptr;
  u64* h_out = nullptr;    // the answers (node output)
  u64* h_bits = nullptr;   // the found bitmap (bitmap output)
  u64* h_found = nullptr;  // the batch's found k-mers (stats), counted on the device
  char* h_text = nullptr;
  u64* h_size = nullptr;
  uint8_t* d_chars = nullptr;
  u64* d_starts = nullptr;
  u64* d_out = nullptr;
  u64* d_offsets = nullptr;
  u64* d_bits = nullptr;
  u64* d_found = nullptr;
  char* d_text = nullptr;
  u64* d_size = nullptr;
  void* d_scratch = nullptr;
  hipEvent_t copied = nullptr, searched = nullptr, returned = nullptr;
  std::vector<u64> offsets;  // the batch's k-mer offsets
  u64 rows = 0, kmers = 0, text_bytes = 0;
  bool busy = false;
};

struct Pipeline {
  Slot slot[2];
  hipStream_t h2d = nullptr, comp = nullptr, d2h = nullptr;
  sbwt_reads_plan* plan = nullptr;

  ~Pipeline() {
    for (Slot& s : slot) {
      if (s.h_chars) (void)hipHostFree(s.h_chars);
      if (s.h_starts) (void)hipHostFree(s.h_starts);
      if (s.h_out) (void)hipHostFree(s.h_out);
      if (s.h_bits) (void)hipHostFree(s.h_bits);
      if (s.h_found) (void)hipHostFree(s.h_found);
      if (s.h_text) (void)hipHostFree(s.h_text);
      if (s.h_size) (void)hipHostFree(s.h_size);
      for (void* d : {static_cast<void*>(s.d_chars), static_cast<void*>(s.d_starts), static_cast<void*>(s.d_out),
                      static_cast<void*>(s.d_offsets), static_cast<void*>(s.d_bits), static_cast<void*>(s.d_found),
                      static_cast<void*>(s.d_text),
                      static_cast<void*>(s.d_size), s.d_scratch})
        if (d) (void)hipFree(d);
      for (hipEvent_t e : {s.copied, s.searched, s.returned})
        if (e) (void)hipEventDestroy(e);
    }
    for (hipStream_t st : {h2d, comp, d2h})
      if (st) (void)hipStreamDestroy(st);
    if (plan) (void)sbwt_reads_plan_free(plan);
  }
};

// Writes n bytes to fd (nothing when fd < 0).
int write_all(int fd, const void* p, sbwt::u64 n) {
  const char* c = static_cast<const char*>(p);
  for (sbwt::u64 o = 0; fd >= 0 && o < n;) {
    const ssize_t w = write(fd, c + o, n - o);
    if (w < 0) return sbwt::fail(SBWT_ERR_IO, "writing the results failed");
    o += static_cast<sbwt::u64>(w);
  }
  return SBWT_OK;
}

}  // namespace

extern "C" int sbwt_search_text(const sbwt_index* ix, const char* text, uint64_t len, uint32_t format,
                                uint32_t flags, uint64_t batch_chars, int fd, uint64_t* stats) {
  using namespace sbwt;
  clear_error();
  SBWT_REQUIRE(ix, "index is NULL");
  SBWT_REQUIRE(text || len == 0, "text is NULL");
  const u32 input = format & 0xFFu, output = format & ~0xFFu;
  SBWT_REQUIRE(input <= SBWT_TEXT_RECORDS, "format %#x: input kind %u is not SBWT_TEXT_LINES or SBWT_TEXT_RECORDS",
               format, input);
  SBWT_REQUIRE(output == SBWT_OUTPUT_TEXT || output == SBWT_OUTPUT_NODES || output == SBWT_OUTPUT_BITMAP,
               "format %#x: output kind %#x is not SBWT_OUTPUT_TEXT, _NODES or _BITMAP", format, output);
  SBWT_REQUIRE(!(flags & SBWT_COUNT), "SBWT_COUNT is not supported here");
  if (stats) stats[0] = stats[1] = stats[2] = 0;
  const bool text_out = output == SBWT_OUTPUT_TEXT;
  const u32 k = ix->k;
  if (!text_out) {
    // binary results: "SBWTRES1", u32 kind (1 nodes, 2 bitmap), u32 k
    unsigned char head[16] = {'S', 'B', 'W', 'T', 'R', 'E', 'S', '1'};
    const u32 kind_code = output == SBWT_OUTPUT_NODES ? 1u : 2u;
    for (int i = 0; i < 4; i++) {
      head[8 + i] = static_cast<unsigned char>(kind_code >> (8 * i));
      head[12 + i] = static_cast<unsigned char>(k >> (8 * i));
    }
    SBWT_TRY(write_all(fd, head, sizeof head));
  }
  if (len == 0) return SBWT_OK;
  if (batch_chars == 0) batch_chars = 64ull << 20;
  const TextKind kind = input ? text_kind(text, len) : kTextLines;

  // batch boundaries (rows never straddle a cut) and the largest batch
  std::vector<u64> cut{0};
  u64 max_bytes = 0, max_lines = 0;
  while (cut.back() < len) {
    int err = SBWT_OK;
    u64 next = text_cut(kind, text, len, cut.back(), cut.back() + batch_chars, &err);
    if (err != SBWT_OK) return err;
    if (next <= cut.back()) next = len;
    max_bytes = std::max(max_bytes, next - cut.back());
    cut.push_back(next);
  }
  {
    // newlines per batch (an upper bound of its rows), counted on host threads
    std::vector<u64> lines(cut.size() - 1, 0);
    const unsigned nt = std::min<u64>(parse_threads(), lines.size());
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
      th.emplace_back([&, t] {
        for (u64 b = t; b < lines.size(); b += nt)
          for (const char* q = text + cut[b]; (q = static_cast<const char*>(memchr(q, '\n', text + cut[b + 1] - q)));) {
            lines[b]++;
            q++;
          }
      });
    for (auto& x : th) x.join();
    for (u64 l : lines) max_lines = std::max(max_lines, l);
  }
  const u64 rows_cap = max_lines + 2, chars_cap = max_bytes;
  const u64 nb = cut.size() - 1;
  // printed answer width: a node id below n and a space, or "-1 "
  u64 width = 2;
  for (u64 x = ix->n > 1 ? ix->n - 1 : 0; x >= 10; x /= 10) width++;
  if (width < 3) width = 3;
  const u64 text_cap = text_out ? chars_cap * width + rows_cap : 0;
  const u64 scratch = text_out ? sbwt_format_text_scratch(chars_cap) : 0;
  // the bitmap is also how the found k-mers are counted (stats)
  const u64 bits_cap = output == SBWT_OUTPUT_BITMAP || stats ? chars_cap / 64 + 1 : 0;

  Pipeline pl;
  for (hipStream_t* st : {&pl.h2d, &pl.comp, &pl.d2h}) SBWT_HIP(hipStreamCreateWithFlags(st, hipStreamNonBlocking));
  for (Slot& s : pl.slot) {
    SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_chars), chars_cap, hipHostMallocDefault));
    SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_starts), rows_cap * 8, hipHostMallocDefault));
    if (output == SBWT_OUTPUT_NODES)
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_out), chars_cap * 8, hipHostMallocDefault));
    if (output == SBWT_OUTPUT_BITMAP)
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_bits), bits_cap * 8, hipHostMallocDefault));
    if (bits_cap) SBWT_HIP(hipMalloc(&s.d_bits, bits_cap * 8));
    if (stats) {
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_found), 8, hipHostMallocDefault));
      SBWT_HIP(hipMalloc(&s.d_found, 8));
    }
    if (text_out) {
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_text), text_cap, hipHostMallocDefault));
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_size), 8, hipHostMallocDefault));
      SBWT_HIP(hipMalloc(&s.d_text, text_cap));
      SBWT_HIP(hipMalloc(&s.d_size, 8));
      SBWT_HIP(hipMalloc(&s.d_scratch, scratch));
    }
    SBWT_HIP(hipMalloc(&s.d_chars, chars_cap));
    SBWT_HIP(hipMalloc(&s.d_starts, rows_cap * 8));
    SBWT_HIP(hipMalloc(&s.d_out, chars_cap * 8));
    SBWT_HIP(hipMalloc(&s.d_offsets, rows_cap * 8));
    for (hipEvent_t* e : {&s.copied, &s.searched, &s.returned})
      SBWT_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  SBWT_TRY(sbwt_reads_plan_create(rows_cap, chars_cap, ix->cfg.segment_kmers, &pl.plan));

  u64 total_rows = 0, total_kmers = 0, total_found = 0;
  std::vector<u64> block;
  // waits for a slot's results and writes them
  auto finish = [&](Slot& s) -> int {
    if (!s.busy) return SBWT_OK;
    SBWT_HIP(hipEventSynchronize(s.returned));
    s.busy = false;
    const u64 words = (s.kmers + 63) / 64;
    if (stats) total_found += *s.h_found;
    if (text_out) {
      const u64 size = *s.h_size;
      if (size > s.text_bytes) return fail(SBWT_ERR_HIP, "result text of %llu bytes exceeds its %llu-byte buffer",
                                           (unsigned long long)size, (unsigned long long)s.text_bytes);
      return write_all(fd, s.h_text, size);
    }
    // a block: rows, k-mers, each row's k-mer count, then the answers or bits
    block.resize(s.rows + 2);
    block[0] = s.rows;
    block[1] = s.kmers;
    for (u64 r = 0; r < s.rows; r++) block[r + 2] = s.offsets[r + 1] - s.offsets[r];
    SBWT_TRY(write_all(fd, block.data(), block.size() * 8));
    return s.h_bits ? write_all(fd, s.h_bits, words * 8) : write_all(fd, s.h_out, s.kmers * 8);
  };
  for (u64 b = 0; b < nb; b++) {
    Slot& s = pl.slot[b & 1];
    SBWT_TRY(finish(s));  // batch b - 2
    u64 rows = 0;
    if (kind == kTextLines)
      SBWT_TRY(parse_lines_parallel(text + cut[b], cut[b + 1] - cut[b], s.h_chars, s.h_starts, rows_cap, &rows));
    else
      SBWT_TRY(parse_text(kind, text + cut[b], cut[b + 1] - cut[b], s.h_chars, s.h_starts, rows_cap, &rows));
    const u64 n_chars = s.h_starts[rows];
    s.rows = rows;
    s.offsets.assign(rows + 1, 0);
    for (u64 r = 0; r < rows; r++) {
      const u64 l = s.h_starts[r + 1] - s.h_starts[r];
      s.offsets[r + 1] = s.offsets[r] + (l >= k ? l - k + 1 : 0);
    }
    s.kmers = s.offsets[rows];
    total_rows += rows;
    total_kmers += s.kmers;
    if (rows == 0) continue;
    if (n_chars) SBWT_HIP(hipMemcpyAsync(s.d_chars, s.h_chars, n_chars, hipMemcpyHostToDevice, pl.h2d));
    SBWT_HIP(hipMemcpyAsync(s.d_starts, s.h_starts, (rows + 1) * 8, hipMemcpyHostToDevice, pl.h2d));
    SBWT_HIP(hipEventRecord(s.copied, pl.h2d));
    SBWT_HIP(hipStreamWaitEvent(pl.comp, s.copied, 0));
    if (s.kmers)
      SBWT_TRY(sbwt_reads_plan_search(pl.plan, ix, s.d_chars, n_chars, s.d_starts, rows, s.d_out, s.d_offsets,
                                      flags, pl.comp));
    else
      SBWT_HIP(hipMemcpyAsync(s.d_offsets, s.offsets.data(), (rows + 1) * 8, hipMemcpyHostToDevice, pl.comp));
    if (text_out) {
      s.text_bytes = s.kmers * width + rows;
      SBWT_TRY(sbwt_format_text(s.d_out, s.d_offsets, rows, s.kmers, s.d_text, s.text_bytes, s.d_size,
                                s.d_scratch, scratch, pl.comp));
    }
    if (s.d_bits) SBWT_TRY(sbwt_found_bitmap(s.d_out, s.kmers, 0, s.d_bits, s.d_found, pl.comp));
    SBWT_HIP(hipEventRecord(s.searched, pl.comp));
    SBWT_HIP(hipStreamWaitEvent(pl.d2h, s.searched, 0));
    if (text_out) {
      SBWT_HIP(hipMemcpyAsync(s.h_size, s.d_size, 8, hipMemcpyDeviceToHost, pl.d2h));
      SBWT_HIP(hipMemcpyAsync(s.h_text, s.d_text, s.text_bytes, hipMemcpyDeviceToHost, pl.d2h));
    }
    if (s.h_bits && s.kmers)
      SBWT_HIP(hipMemcpyAsync(s.h_bits, s.d_bits, (s.kmers + 63) / 64 * 8, hipMemcpyDeviceToHost, pl.d2h));
    if (s.h_found) SBWT_HIP(hipMemcpyAsync(s.h_found, s.d_found, 8, hipMemcpyDeviceToHost, pl.d2h));
    if (s.h_out && s.kmers)
      SBWT_HIP(hipMemcpyAsync(s.h_out, s.d_out, s.kmers * 8, hipMemcpyDeviceToHost, pl.d2h));
    SBWT_HIP(hipEventRecord(s.returned, pl.d2h));
    s.busy = true;
  }
  // the last two batches, in order
  SBWT_TRY(finish(pl.slot[nb & 1]));
  SBWT_TRY(finish(pl.slot[(nb + 1) & 1]));
  if (stats) {
    stats[0] = total_rows;
    stats[1] = total_kmers;
    stats[2] = total_found;
  }
  return SBWT_OK;
}
