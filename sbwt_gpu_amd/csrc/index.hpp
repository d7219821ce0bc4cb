// Internal definition of sbwt_index and the helpers shared by the HIP
// sources of libsbwt_gpu_amd.
#pragma once
#include "layout.hpp"

struct sbwt_index {
  int device = 0;
  sbwt::u32 k = 0;
  sbwt::u64 n = 0;
  sbwt::u64 C[5] = {0, 0, 0, 0, 0};
  // rows A,C,G,T,group as MSB-first u64 words, row r at d_rows + r * nwords
  sbwt::u64 nwords = 0;
  sbwt::u64* d_rows = nullptr;
  sbwt::u32 has_groups = 0;
  // rank chunks of the current layout, 5 rows of row_q 16-byte units
  sbwt::u32 layout = SBWT_LAYOUT_L16;
  sbwt::u64 row_q = 0;
  uint4* d_lines = nullptr;
  // presearch table of 4^m intervals
  sbwt::u32 m = 0;
  uint2* d_pre = nullptr;
  sbwt::u32 pre_alloc_m = 0;  // depth d_pre was allocated for
  // access counters of the counted search variants (lines, steps, k-mers, full searches)
  unsigned long long* d_counters = nullptr;
  // work-item counter of the persistent read search (zeroed by each launch)
  unsigned long long* d_item_ctr = nullptr;
  int num_cus = 0;
  // sampled select index (chunk of every 4096th 1-bit per row), built on demand
  sbwt::u32* d_select = nullptr;
  sbwt::u32 select_layout = 0;
  sbwt::u64 select_per_row = 0;
  sbwt_config cfg{};
  sbwt::u64 num_kmers = 0;  // |K| when built on the GPU, else 0
};

namespace sbwt {

constexpr u32 kRows = 5;    // A, C, G, T, suffix-group starts
constexpr u32 kGroupRow = 4;
constexpr u32 kMaxPresearch = 17;  // 4^17 entries x 8 B = 137 GB of HBM, beside a 3 Gbp index in 288 GB
constexpr u32 kCounters = 4;

inline u64 words_for(u64 n) { return ((n + 4095) / 4096) * 64; }

DevIndex dev_view(const sbwt_index* ix);

// Builds d_lines for ix->layout from d_rows and sets C from the device.
int build_layout(sbwt_index* ix, hipStream_t s);
// Finishes an index whose d_rows are filled: layout, C check, suffix
// groups, presearch. A missing group row is derived; a given one is checked
// against the edges (cheap, for rows this library wrote) or, with
// verify_groups, compared with the derived row word for word.
int finish_index(sbwt_index* ix, hipStream_t s, bool verify_groups = false);
// Frees everything (safe on partially built indexes).
void destroy_index(sbwt_index* ix);
// A fresh index shell with cfg applied (rows not allocated).
sbwt_index* new_index(const sbwt_config* cfg, u32 k, u64 n);
// Validates n and k against what the device layout supports.
int check_shape(u64 n, u32 k);
// Source column of every node's incoming edge (labels.hip): u32[n],
// allocated here, freed by the caller.
int build_pred(const sbwt_index* ix, u32** pred, hipStream_t s);
// Writes the suffix-group-start row derived from the four rows to dst
// (nwords words; labels.hip); the rank layout must be built.
int derive_groups(const sbwt_index* ix, u64* dst, hipStream_t s);
// SBWT_ERR_FORMAT unless the index's group row equals the derived one.
int verify_groups(const sbwt_index* ix, hipStream_t s);
// SBWT_ERR_FORMAT unless every column with an edge, and column 0, starts a
// group of the index's group row (labels.hip).
int check_groups(const sbwt_index* ix, hipStream_t s);
// Labels of all nodes as top-aligned k-mers, ~0 / 0 for padded nodes.
int node_labels(sbwt_index* ix, u64* d_labels, u8* d_full, hipStream_t s);
// Run items (runs.hip): work items {row, a, b, c} that tile every row:
// windows [a, b) hold a non-ACGT character (answered ~0 by
// run_items before the search, which skips them), [b, c) are an N-free run
// of at most seg k-mers. Needs
// k >= kRunItemsMinK, for which run_items_bound items always suffice.
constexpr u32 kRunItemsMinK = 16;
u64 run_items_bound(u64 rows, u64 chars, u32 seg);
// the items, their number in *count (device), and out[0 .. offsets[rows]) = ~0
// (the answers of the gaps; the search writes the runs' answers over it)
int run_items(const u64* nmask, const u64* row_starts, const u64* offsets, u64 rows, u32 k, u32 seg, uint4* runs,
              u64 cap, unsigned long long* count, u64* out, hipStream_t s);
// Work items of a read search over run items (null runs: segments of the
// work prefix).
struct RunList {
  const uint4* runs = nullptr;
  const unsigned long long* count = nullptr;
  u64 cap = 0;
};
// The persistent read search (stream.hip). item_ctr: the work-item counter
// of the launch (8 device bytes; null: the index's own). Two searches on one
// index may overlap only if they count in different places.
int search_reads_persistent(const sbwt_index* ix, const u64* packed, const u64* nmask, const u64* rs,
                            const u64* off, const u64* work, const u32* items, u64 rows, u64 bound, u32 seg,
                            const u8* maybe, u64* out, bool nbreak, bool stream, bool cert, bool count,
                            const RunList& runs, unsigned long long* item_ctr, hipStream_t s);
// The read search with run items, as sbwt_search_reads with flags (search.hip).
int search_reads_runs(const sbwt_index* ix, const u64* packed, const u64* row_starts, const u64* offsets, u64 rows,
                      u32 seg, const u8* maybe, u64* out, u32 flags, const RunList& runs,
                      unsigned long long* item_ctr, hipStream_t s);
// sbwt_search_reads_masked (d_maybe may be null) with the work-item counter
// of a plan (search.hip).
int search_reads_counted_on(const sbwt_index* ix, const u64* packed, const u64* nmask, const u64* row_starts,
                            const u64* offsets, const u64* work, const u32* items, u64 rows, u64 work_bound,
                            u32 seg, const u8* maybe, u64* out, u32 flags, unsigned long long* item_ctr,
                            hipStream_t s);

}  // namespace sbwt
