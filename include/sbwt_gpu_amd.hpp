// C++ host API of sbwt_gpu_amd over the C ABI of sbwt_gpu_amd.h.
//
// Mirrors the reference's operator interface for the search path
// (epicharri/sbwt_gpu), with the same names, argument meaning and error
// behaviour (int 0 = ok, 1 = error, message on stderr):
//   sbwt_amd::Parameters              epic::Parameters            include/parameters.hpp:39-58,79-297
//   sbwt_amd::DeviceMemory            epic::gpu::DeviceMemory     include/gpu/device_memory.hpp:14-74
//   sbwt_amd::MatrixSBWT::create      epic::MatrixSBWT::create    include/matrix_sbwt.hpp:32-200
//   sbwt_amd::KmerSearch::search      epic::KmerSearch::search    include/epicseq/kmer_search.hpp:97-153
//   sbwt_amd::ResultsPersister        epic::ResultsPersister      include/epicseq/results_persister.hpp:56-100
// The index and the query buffers live in HBM; the search itself runs in the
// library's HIP kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "sbwt_gpu_amd.h"

namespace sbwt_amd {

inline int report(int status, const char* what) {
  if (status != SBWT_OK) {
    fprintf(stderr, "%s failed: %s\n", what, sbwt_last_error());
    return 1;
  }
  return 0;
}

// Run configuration; flag names follow the reference CLI (--k, --k-presearch,
// --without-presearch, --file-a/c/g/t, --file-queries, --do-not-store-results)
// plus this build's own knobs.
struct Parameters {
  bool with_presearch = true;
  uint32_t k = 30;            // include/parameters.hpp:48
  uint32_t k_presearch = 12;  // include/parameters.hpp:49
  uint32_t layout = SBWT_LAYOUT_L16;
  uint32_t segment_kmers = 256;
  bool store_results = true;
  bool independent = false;   // search every k-mer from the presearch table
  bool n_as_a = false;        // reference packing of non-ACGT characters
  bool pipelined = false;     // sbwt_search_text: batches with copies, search and printing overlapped
  bool records = false;       // FASTA / FASTQ records (one row each); implies pipelined
  uint64_t batch_chars = 0;   // characters per pipelined batch, 0 = 64 MiB
  uint32_t output = SBWT_OUTPUT_TEXT;  // --output=nodes|bitmap: binary results; implies pipelined
  std::string filename_A, filename_C, filename_G, filename_T, filename_groups, fileQueries;
  std::string file_image;  // --file-image=: resume from an index image instead of the four files
  std::string save_image;  // --save-image=: checkpoint the index to an image after creating it

  static bool take(const char* arg, const char* key, std::string& out) {
    const size_t n = strlen(key);
    if (strncmp(arg, key, n) == 0 && arg[n] == '=') {
      out = arg + n + 1;
      return true;
    }
    return false;
  }

  // Returns 0 on success, 1 on an invalid flag (include/parameters.hpp:248-259).
  int read_arguments(int argc, char** argv) {
    for (int i = 1; i < argc; i++) {
      const char* a = argv[i];
      std::string v;
      if (take(a, "--k", v)) k = static_cast<uint32_t>(std::stoul(v));
      else if (take(a, "--k-presearch", v)) k_presearch = static_cast<uint32_t>(std::stoul(v));
      else if (take(a, "--layout", v)) layout = static_cast<uint32_t>(std::stoul(v));
      else if (take(a, "--segment", v)) segment_kmers = static_cast<uint32_t>(std::stoul(v));
      else if (take(a, "--file-a", v)) filename_A = v;
      else if (take(a, "--file-c", v)) filename_C = v;
      else if (take(a, "--file-g", v)) filename_G = v;
      else if (take(a, "--file-t", v)) filename_T = v;
      else if (take(a, "--file-groups", v)) filename_groups = v;
      else if (take(a, "--file-image", v)) file_image = v;
      else if (take(a, "--save-image", v)) save_image = v;
      else if (take(a, "--file-queries", v)) fileQueries = v;
      else if (!strcmp(a, "--without-presearch")) with_presearch = false;
      else if (!strcmp(a, "--do-not-store-results")) store_results = false;
      else if (!strcmp(a, "--independent")) independent = true;
      else if (!strcmp(a, "--n-as-a")) n_as_a = true;
      else if (!strcmp(a, "--pipelined")) pipelined = true;
      else if (!strcmp(a, "--records")) records = pipelined = true;
      else if (take(a, "--batch-chars", v)) batch_chars = std::stoull(v);
      else if (take(a, "--output", v) && (v == "text" || v == "nodes" || v == "bitmap")) {
        output = v == "text" ? SBWT_OUTPUT_TEXT : v == "nodes" ? SBWT_OUTPUT_NODES : SBWT_OUTPUT_BITMAP;
        if (output != SBWT_OUTPUT_TEXT) pipelined = true;
      }
      else {
        fprintf(stderr, "Unknown argument %s.\n", a);
        return 1;
      }
    }
    if (k < 1 || k > 32) {
      fprintf(stderr, "The value of k must be at least 1 and at most 32.\n");
      return 1;
    }
    if (!with_presearch) k_presearch = 0;
    if (k_presearch > k) k_presearch = k;
    return 0;
  }
};

// Owns the device-resident index and the stream every call runs on.
class DeviceMemory {
 public:
  sbwt_index* index = nullptr;
  hipStream_t stream = nullptr;
  DeviceMemory() { (void)hipStreamCreate(&stream); }
  ~DeviceMemory() {
    if (index) sbwt_index_free(index);
    if (stream) (void)hipStreamDestroy(stream);
  }
  DeviceMemory(const DeviceMemory&) = delete;
  DeviceMemory& operator=(const DeviceMemory&) = delete;
};

class MatrixSBWT {
 public:
  // Reads the four bit-vector files (and the optional group file), or an
  // index image (--file-image; k is the image's), builds the rank layout,
  // the C array and the presearch table in HBM, and writes an image when
  // asked (--save-image).
  int create(DeviceMemory& dm, Parameters& p) {
    sbwt_config cfg;
    sbwt_config_init(&cfg);
    cfg.presearch_depth = p.with_presearch ? p.k_presearch : 0;
    cfg.layout = p.layout;
    cfg.segment_kmers = p.segment_kmers;
    const char* paths[4] = {p.filename_A.c_str(), p.filename_C.c_str(), p.filename_G.c_str(), p.filename_T.c_str()};
    if (dm.index) sbwt_index_free(dm.index);
    dm.index = nullptr;
    if (!p.file_image.empty()) {
      if (report(sbwt_index_load(p.file_image.c_str(), &cfg, &dm.index), "MatrixSBWT::create")) return 1;
      sbwt_index_info info;
      sbwt_index_get_info(dm.index, &info);
      p.k = info.k;
    } else if (report(sbwt_index_from_files(paths, p.filename_groups.empty() ? nullptr : p.filename_groups.c_str(),
                                            p.k, &cfg, &dm.index),
                      "MatrixSBWT::create")) {
      return 1;
    }
    if (!p.save_image.empty()) return report(sbwt_index_save(dm.index, p.save_image.c_str()), "MatrixSBWT::create");
    return 0;
  }
};

class ResultsPersister {
 public:
  int save_results_to_file(const std::vector<uint64_t>& results, const std::vector<uint64_t>& offsets,
                           const Parameters& p) {
    if (!p.store_results) {
      fprintf(stderr, "The results are not stored to the file.\n");
      return 0;
    }
    return report(sbwt_write_results(results.data(), offsets.data(), offsets.size() - 1, 1),
                  "ResultsPersister::save_results_to_file");
  }
};

class KmerSearch {
 public:
  std::vector<uint64_t> results;  // one answer per k-mer, rows concatenated
  std::vector<uint64_t> offsets;  // rows + 1
  uint64_t number_of_positions = 0;
  float millis_search = 0.0f;

  KmerSearch(std::string filename, Parameters& p) : filename_(std::move(filename)), p_(p) {}

  int search(DeviceMemory& dm) {
    std::ifstream fs(filename_, std::ios::binary | std::ios::ate);
    if (!fs.is_open()) {
      fprintf(stderr, "The file %s can not be opened.\n", filename_.c_str());
      return 1;
    }
    const uint64_t size = static_cast<uint64_t>(fs.tellg());
    if (size == 0) {
      fprintf(stderr, "The size of the file %s is too small to contain any data.\n", filename_.c_str());
      return 1;
    }
    std::vector<char> text(size);
    fs.seekg(0);
    fs.read(text.data(), static_cast<std::streamsize>(size));
    if (p_.pipelined) return search_pipelined(dm, text);
    std::vector<uint8_t> chars(size);
    uint64_t cap = 2;
    for (char ch : text) cap += ch == '\n';
    std::vector<uint64_t> starts(cap);
    uint64_t rows = 0;
    if (report(sbwt_parse_lines(text.data(), size, chars.data(), starts.data(), cap, &rows), "read_file")) return 1;
    starts.resize(rows + 1);
    const uint64_t n_chars = starts[rows];
    offsets.assign(rows + 1, 0);
    for (uint64_t r = 0; r < rows; r++) {
      const uint64_t len = starts[r + 1] - starts[r];
      offsets[r + 1] = offsets[r] + (len >= p_.k ? len - p_.k + 1 : 0);
    }
    number_of_positions = offsets[rows];
    results.assign(number_of_positions, 0);
    if (rows == 0 || number_of_positions == 0) return persist();

    const uint32_t seg = p_.segment_kmers;
    const uint64_t bound = sbwt_reads_work_bound(rows, number_of_positions, seg);
    uint8_t* d_chars = nullptr;
    uint64_t *d_starts = nullptr, *d_packed = nullptr, *d_nmask = nullptr, *d_off = nullptr, *d_work = nullptr,
             *d_out = nullptr;
    uint32_t* d_items = nullptr;
    int st = 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    do {
      if (hipMalloc(&d_chars, n_chars ? n_chars : 1) || hipMalloc(&d_starts, (rows + 1) * 8) ||
          hipMalloc(&d_packed, sbwt_packed_words(n_chars) * 8) || hipMalloc(&d_nmask, sbwt_nmask_words(n_chars) * 8) ||
          hipMalloc(&d_off, (rows + 1) * 8) || hipMalloc(&d_work, (rows + 1) * 8) ||
          hipMalloc(&d_out, number_of_positions * 8) || hipMalloc(&d_items, bound * 4)) {
        fprintf(stderr, "Allocating device memory for the queries failed.\n");
        break;
      }
      if (hipMemcpyAsync(d_chars, chars.data(), n_chars, hipMemcpyHostToDevice, dm.stream) ||
          hipMemcpyAsync(d_starts, starts.data(), (rows + 1) * 8, hipMemcpyHostToDevice, dm.stream)) {
        fprintf(stderr, "Copying the queries to the device failed.\n");
        break;
      }
      void* s = dm.stream;
      if (report(sbwt_pack(d_chars, n_chars, d_packed, d_nmask, s), "pack") ||
          report(sbwt_kmer_offsets(d_starts, rows, p_.k, d_off, s), "offsets") ||
          report(sbwt_reads_work(d_off, rows, seg, d_work, d_items, s), "work"))
        break;
      const uint32_t flags = (p_.n_as_a ? SBWT_READS_N_AS_A : 0u) | (p_.independent ? SBWT_READS_INDEPENDENT : 0u);
      (void)hipEventRecord(e0, dm.stream);
      if (report(sbwt_search_reads(dm.index, d_packed, d_nmask, d_starts, d_off, d_work, d_items, rows, bound, seg,
                                   d_out, flags, s),
                 "SearchCaller::search"))
        break;
      (void)hipEventRecord(e1, dm.stream);
      if (hipMemcpyAsync(results.data(), d_out, number_of_positions * 8, hipMemcpyDeviceToHost, dm.stream) ||
          hipStreamSynchronize(dm.stream)) {
        fprintf(stderr, "Copying the results to the host failed.\n");
        break;
      }
      (void)hipEventElapsedTime(&millis_search, e0, e1);
      st = 0;
    } while (false);
    for (void* ptr : {static_cast<void*>(d_chars), static_cast<void*>(d_starts), static_cast<void*>(d_packed),
                      static_cast<void*>(d_nmask), static_cast<void*>(d_off), static_cast<void*>(d_work),
                      static_cast<void*>(d_out), static_cast<void*>(d_items)})
      if (ptr) (void)hipFree(ptr);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    fprintf(stderr, "Search in GPU: %f ms\n", millis_search);
    fprintf(stderr, "Search in GPU per query: %f ns\n",
            static_cast<double>(millis_search) * 1e6 / static_cast<double>(number_of_positions));
    return persist();
  }

 private:
  // Parsing, copies, search and printing in overlapping batches; the result
  // text goes to stdout as persist() would write it (--pipelined, --records),
  // or binary results (--output=nodes|bitmap).
  int search_pipelined(DeviceMemory& dm, const std::vector<char>& text) {
    const uint32_t flags = (p_.n_as_a ? SBWT_READS_N_AS_A : 0u) | (p_.independent ? SBWT_READS_INDEPENDENT : 0u);
    uint64_t stats[3] = {0, 0, 0};
    if (!p_.store_results) fprintf(stderr, "The results are not stored to the file.\n");
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, nullptr);
    const int st = report(sbwt_search_text(dm.index, text.data(), text.size(),
                                           (p_.records ? SBWT_TEXT_RECORDS : SBWT_TEXT_LINES) | p_.output, flags,
                                           p_.batch_chars, p_.store_results ? 1 : -1, stats),
                          "KmerSearch::search");
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&millis_search, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    number_of_positions = stats[1];
    fprintf(stderr, "Search of %llu k-mers in %llu rows, text to text: %f ms\n",
            static_cast<unsigned long long>(stats[1]), static_cast<unsigned long long>(stats[0]), millis_search);
    return 0;
  }

  int persist() {
    ResultsPersister rp;
    return rp.save_results_to_file(results, offsets, p_);
  }
  std::string filename_;
  Parameters& p_;
};

}  // namespace sbwt_amd
