/*
 * sbwt_gpu_amd — SBWT k-mer membership search on one AMD Instinct MI355X.
 *
 * C ABI of libsbwt_gpu_amd.so. Plain pointers and sizes only. Every function
 * returns SBWT_OK (0) or an error code; sbwt_last_error() gives the message
 * of the calling thread's last failure. Pointers named d_* are device (HBM)
 * pointers; `stream` is a hipStream_t (NULL = the null stream). Calls on one
 * index are thread-compatible, not thread-safe: do not mutate an index
 * (presearch, relayout, free) while another thread searches it.
 *
 * Bit order of rows: bit i of a row is bit 63-(i&63) of word i>>6
 * (MSB-first), the order the reference reaches after its byte swap
 * (include/rank_data_structures.hpp:98). Node ids are SBWT column indices; an
 * absent k-mer answers ~0 (the reference's -1, include/gpu/search_kernel.hpp:98-103).
 *
 * Reference interfaces each entry point replaces are cited per function
 * (paths relative to the epicharri/sbwt_gpu tree).
 */
#ifndef SBWT_GPU_AMD_H
#define SBWT_GPU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference returns int 0/1, include/gpu/cuda.hpp:7-11) */
#define SBWT_OK 0
#define SBWT_ERR_ARG 1         /* invalid argument or size */
#define SBWT_ERR_HIP 2         /* HIP runtime error */
#define SBWT_ERR_IO 3          /* file missing / short */
#define SBWT_ERR_FORMAT 4      /* inconsistent index (unequal rows, C[4] != n) */
#define SBWT_ERR_UNSUPPORTED 5 /* configuration this build does not support */
#define SBWT_ERR_INTERNAL 6    /* a device-side consistency check failed; the
                                  message names the stage */

/* ---- rank layouts: bytes per rank chunk (u32 count + payload bits) */
#define SBWT_LAYOUT_AUTO 0
#define SBWT_LAYOUT_L16 16 /* 96 columns per 16-byte chunk  */
#define SBWT_LAYOUT_L32 32 /* 224 columns per 32-byte chunk */
#define SBWT_LAYOUT_L64 64 /* 480 columns per 64-byte chunk */

/* ---- flags of the read search */
#define SBWT_READS_N_AS_A 1u      /* code non-ACGT as A and search it (reference
                                     packing table, create_kmer_string.hpp:20);
                                     default: a window with a non-ACGT char answers ~0 */
#define SBWT_READS_INDEPENDENT 2u /* search every k-mer from the presearch table
                                     (the reference's way); default: stream
                                     consecutive k-mers through suffix groups
                                     (when the index has them) and skip k-mers
                                     that contain a substring proven absent */
#define SBWT_READS_PERSISTENT 8u  /* the persistent uniform-step kernel (each
                                     lane a state machine taking one
                                     two-rank step per iteration, work items
                                     claimed per wave); the default when
                                     streaming with certificates */
#define SBWT_READS_SEGMENT_LANES 32u /* one lane per work item, k-mers in
                                     order; the default for independent
                                     searches or without certificates */
#define SBWT_READS_NO_CERT 16u    /* no dead-substring certificates (for
                                     measurement) */
#define SBWT_READS_RUN_ITEMS 64u  /* sbwt_reads_plan_search only: work items are
                                     the rows' runs of k-mers free of non-ACGT
                                     characters (pieces of at most seg k-mers),
                                     each led by the gap of windows holding
                                     such a character before it, found by a
                                     pass over the rows' non-ACGT mask; the
                                     search kernels answer a gap in one step
                                     and never test a character. Ignored for
                                     k < 16, with SBWT_READS_N_AS_A, and for
                                     batches of 2^32 - 64 characters or more */
#define SBWT_SEARCH_COOPERATIVE 128u /* sbwt_search_kmers_flags / _masked on
                                     32- and 64-byte chunks: two or four lanes
                                     per k-mer, each loading 16 bytes of every
                                     rank chunk, ranks summed by shuffles
                                     (include/gpu/rank_kernel.hpp:298-396); the
                                     16-byte layout has one load per rank and
                                     keeps one lane per k-mer */
#define SBWT_COUNT 4u             /* counted variant: add the 128-byte lines the
                                     rank reads touch to the index's counters
                                     (sbwt_counters); also accepted by
                                     sbwt_search_kmers_flags */

typedef struct sbwt_index sbwt_index;

typedef struct sbwt_config {
  uint32_t presearch_depth; /* m; 0 = no presearch table */
  uint32_t layout;          /* SBWT_LAYOUT_* */
  uint32_t block_threads;   /* threads per block of the search kernels, 0 = auto */
  uint32_t segment_kmers;   /* k-mers per work item of the read search, 0 = auto */
  uint32_t variant;         /* search kernel variant, 0 = auto */
  uint32_t reserved[3];
} sbwt_config;

typedef struct sbwt_index_info {
  uint64_t n;               /* columns (nodes) */
  uint64_t C[5];            /* C[0]=1, C[c+1]=C[c]+ones(c) */
  uint64_t device_bytes;    /* HBM held by the index */
  uint32_t k;
  uint32_t presearch_depth;
  uint32_t layout;
  uint32_t has_groups;      /* suffix-group starts present (streaming search) */
} sbwt_index_info;

/* Defaults: presearch_depth 12 (include/parameters.hpp:50), layout L16. */
int sbwt_config_init(sbwt_config *cfg);

/* Index from four host rows (A,C,G,T; each >= ceil(n/64) words) plus the
 * optional suffix-group-start row. Replaces epic::MatrixSBWT::create
 * (include/matrix_sbwt.hpp:32-200) after the file read: rank structures, C
 * array, device upload. Fails with SBWT_ERR_FORMAT unless C[4] == n. When
 * group is NULL the row is derived on the device from the four rows (node
 * d starts a group unless its last k-1 label characters equal node d-1's),
 * so the streaming read search is available for the reference's own
 * four-file indexes; a given group row must equal the derived one
 * (SBWT_ERR_FORMAT otherwise).
 * Builds the presearch table of cfg->presearch_depth. */
int sbwt_index_from_rows(const uint64_t *const rows[4], const uint64_t *group,
                         uint64_t n, uint32_t k, const sbwt_config *cfg,
                         sbwt_index **out);

/* Same from the reference's four bit-vector files ([u64 LE n][ceil(n/8)
 * bytes], include/bit_vector.hpp:65-128); group_path may be NULL (the
 * group row is then derived, as above). Replaces MatrixSBWT::create
 * including BitVector::create/read. */
int sbwt_index_from_files(const char *const paths[4], const char *group_path,
                          uint32_t k, const sbwt_config *cfg, sbwt_index **out);

/* GPU SBWT construction from sequences (no reference counterpart: the
 * reference reads bit vectors made by an external tool). d_seq holds the
 * concatenated ASCII sequences, d_starts[nseq+1] their offsets; non-ACGT
 * characters break k-mers. Builds the suffix-group-start row too. */
int sbwt_index_build(const uint8_t *d_seq, const uint64_t *d_starts,
                     uint64_t nseq, uint32_t k, const sbwt_config *cfg,
                     void *stream, sbwt_index **out);

/* Largest item count of one sort, select or unique call inside
 * sbwt_index_build (0 = the default, 2^28); larger buckets are split by
 * further key bits. Process-wide; not for concurrent builds. It exists so
 * tests can exercise the bucket seams on small inputs. No reference
 * counterpart. */
int sbwt_build_set_sort_chunk(uint64_t max_items);

/* Frees every device buffer (include/gpu/device_memory.hpp:81-121). */
int sbwt_index_free(sbwt_index *ix);

/* Checkpoint of an index: writes an image file (a 128-byte "SBWTIMG1"
 * header with k, n, C, the configuration and a checksum, then the five rows
 * as they sit in device memory). No reference counterpart: the reference
 * rebuilds its index from the four bit-vector files at every start
 * (include/matrix_sbwt.hpp:32-200). */
int sbwt_index_save(const sbwt_index *ix, const char *path);

/* Resume from an image written by sbwt_index_save: streams the rows to the
 * device through pinned buffers, checks their checksum (SBWT_ERR_FORMAT on a
 * damaged file) and derives the rank layout, C (checked against the header)
 * and the presearch table on the device, with the saved configuration, or
 * with cfg when not NULL. */
int sbwt_index_load(const char *path, const sbwt_config *cfg, sbwt_index **out);

int sbwt_index_get_info(const sbwt_index *ix, sbwt_index_info *info);

/* Digests of the index's device memory, for checking that nothing changed
 * it between two points: out[0] = the rows, out[1] = the rank chunks,
 * out[2] = the presearch table (0 without one). Each is the sum over the
 * region's 64-bit words of splitmix64's output function applied to
 * (word + index * 2^64/phi), so a changed word or an exchange of two words
 * changes it. Synchronous (the null stream). No reference counterpart: the
 * reference holds its index in 18 separate allocations it never re-reads
 * for checking (include/gpu/device_memory.hpp:14-74). */
int sbwt_index_digest(const sbwt_index *ix, uint64_t out[3]);

/* Copies the rows back to host memory, nwords words each (>= ceil(n/64)),
 * MSB-first; group may be NULL. */
int sbwt_index_get_rows(const sbwt_index *ix, uint64_t *const rows[4],
                        uint64_t *group, uint64_t nwords);

/* Writes the rows in the reference's bit-vector file format. */
int sbwt_index_write_files(const sbwt_index *ix, const char *const paths[4],
                           const char *group_path);

/* (Re)builds the presearch table for m-mers, 0 <= m <= min(k, 17) (4^m x 8 bytes).
 * Replaces epic::PreSearch::call_presearch (include/presearch.hpp:32-130)
 * and the presearch kernel (include/gpu/presearch_kernel.hpp:19-37). */
int sbwt_presearch(sbwt_index *ix, uint32_t m, void *stream);

/* Host copies of the presearch table: 4^m intervals [left, right]. */
int sbwt_presearch_get(const sbwt_index *ix, uint64_t *left, uint64_t *right);

/* Access counters accumulated by searches run with SBWT_COUNT:
 * out[0] = 128-byte rank lines touched (a presearch gather counts one line,
 * a narrowing step one or two, a streaming step one or more), out[1] =
 * narrowing + streaming steps, out[2] = k-mers answered, out[3] = k-mers
 * that took a full search. reset != 0 zeroes them after the copy. No
 * reference counterpart (the reference times kernels only,
 * include/gpu/device_stream.hpp:52-78). */
int sbwt_counters(const sbwt_index *ix, uint64_t out[4], int reset, void *stream);

/* sbwt_search_kmers with flags (0, SBWT_COUNT, SBWT_SEARCH_COOPERATIVE). */
int sbwt_search_kmers_flags(const sbwt_index *ix, const uint64_t *d_kmers,
                            uint64_t count, uint64_t *d_out, uint32_t flags,
                            void *stream);

/* Re-lays the rank chunks (SBWT_LAYOUT_*), e.g. for the autotuner. */
int sbwt_index_set_layout(sbwt_index *ix, uint32_t layout, void *stream);

/* Searches count top-aligned k-mers (char 0 in bits 63:62, as the
 * reference's kernel holds them after extraction, search_kernel.hpp:66-68);
 * out[q] = node or ~0. */
int sbwt_search_kmers(const sbwt_index *ix, const uint64_t *d_kmers,
                      uint64_t count, uint64_t *d_out, void *stream);

/* sbwt_search_kmers_flags where k-mers with d_maybe[q] == 0 (one byte per
 * k-mer, e.g. from sbwt_prefilter) answer ~0 without a search. Equals the
 * unmasked search whenever no indexed k-mer is marked 0. No reference
 * counterpart (the prefilter's exact stage). */
int sbwt_search_kmers_masked(const sbwt_index *ix, const uint64_t *d_kmers,
                             uint64_t count, const uint8_t *d_maybe,
                             uint64_t *d_out, uint32_t flags, void *stream);

/* The reference kernel ABI (include/gpu/search_kernel.hpp:62-63, launched by
 * launch_search_kernel, include/gpu/search_and_presearch_launchers.hpp:21):
 * d_packed is the 2-bit string of sbwt_pack, d_positions holds count k-mer
 * start offsets (in characters) and is overwritten in place by the results. */
int sbwt_search_positions(const sbwt_index *ix, const uint64_t *d_packed,
                          uint64_t *d_positions, uint64_t count, void *stream);

/* Words of the packed string and of the N mask for n_chars characters. */
uint64_t sbwt_packed_words(uint64_t n_chars);
uint64_t sbwt_nmask_words(uint64_t n_chars);

/* ASCII -> 2-bit, 32 characters per word, char 0 in bits 63:62, C/c=1,
 * G/g=2, T/t=3, anything else 0 (include/gpu/create_kmer_string.hpp:22-42).
 * Characters past n_chars read as '\0'. d_nmask (optional) gets one bit per
 * character, MSB-first, set where the character is not A/C/G/T (any case). */
int sbwt_pack(const uint8_t *d_chars, uint64_t n_chars, uint64_t *d_packed,
              uint64_t *d_nmask, void *stream);

/* d_offsets[rows+1]: exclusive scan of max(0, len_r - k + 1) over rows
 * given by d_row_starts[rows+1] (rows shorter than k yield no k-mers). */
int sbwt_kmer_offsets(const uint64_t *d_row_starts, uint64_t rows, uint32_t k,
                      uint64_t *d_offsets, void *stream);

/* Start offset of every k-mer of every row (include/gpu/positions_kernel.hpp:18-36);
 * d_positions has d_offsets[rows] entries. */
int sbwt_positions(const uint64_t *d_row_starts, const uint64_t *d_offsets,
                   uint64_t rows, uint32_t k, uint64_t *d_positions,
                   void *stream);

/* sbwt_search_reads where k-mers with d_maybe[offset] == 0 (one byte per
 * k-mer in output order, e.g. from sbwt_prefilter) answer ~0 without a
 * search. The answers equal the unmasked search whenever no indexed k-mer
 * is marked 0. No reference counterpart (the prefilter's exact stage). */
int sbwt_search_reads_masked(const sbwt_index *ix, const uint64_t *d_packed,
                             const uint64_t *d_nmask, const uint64_t *d_row_starts,
                             const uint64_t *d_offsets, const uint64_t *d_work,
                             const uint32_t *d_items, uint64_t rows,
                             uint64_t work_bound, uint32_t seg,
                             const uint8_t *d_maybe, uint64_t *d_out,
                             uint32_t flags, void *stream);

/* Top-aligned k-mers starting at count character positions of the packed
 * string (the extraction of include/gpu/search_kernel.hpp:66-68). Each
 * position p needs packed words up to 2p/64 + 1. */
int sbwt_extract_kmers(const uint64_t *d_packed, const uint64_t *d_positions,
                       uint64_t count, uint64_t *d_kmers, void *stream);

/* Upper bound of the work items of sbwt_search_reads (segments of seg
 * k-mers; 0 = 256), the work-item prefix d_work[rows+1], and, when d_items
 * is not NULL, the row of every work item (work_bound u32 entries), which
 * lets the persistent kernel fetch an item with one load. */
uint64_t sbwt_reads_work_bound(uint64_t rows, uint64_t total_kmers, uint32_t seg);
int sbwt_reads_work(const uint64_t *d_offsets, uint64_t rows, uint32_t seg,
                    uint64_t *d_work, uint32_t *d_items, void *stream);

/* Searches every k-mer of every row; d_out[d_offsets[r] + j] is the answer
 * for the k-mer starting at character j of row r. Replaces the query half of
 * epic::KmerSearch::search (include/epicseq/kmer_search.hpp:97-153):
 * packing, positions and SearchCaller::search fused into one pass over the
 * rows. d_nmask may be NULL only with SBWT_READS_N_AS_A; d_items may be
 * NULL (work items are then located by binary search). work_bound is
 * sbwt_reads_work_bound of the same rows. No host synchronisation; can be
 * captured in a HIP graph. */
int sbwt_search_reads(const sbwt_index *ix, const uint64_t *d_packed,
                      const uint64_t *d_nmask, const uint64_t *d_row_starts,
                      const uint64_t *d_offsets, const uint64_t *d_work,
                      const uint32_t *d_items, uint64_t rows,
                      uint64_t work_bound, uint32_t seg, uint64_t *d_out,
                      uint32_t flags, void *stream);

/* ---- select and enumeration (no reference counterpart: the reference has
 * rank only). Needed to enumerate the index's k-mers, e.g. as training
 * positives of the prefilter. ----------------------------------------- */

/* d_out[q] = column of the d_ranks[q]-th (0-based) 1-bit of row c (0..3),
 * or ~0 when the row has fewer 1-bits. Builds a sampled select index on
 * first use (held by the index). */
int sbwt_select(sbwt_index *ix, uint32_t c, const uint64_t *d_ranks,
                uint64_t count, uint64_t *d_out, void *stream);

/* Label of every node j < n as a top-aligned k-mer in d_labels[j], with
 * d_full[j] = 1, or ~0 and d_full[j] = 0 for '$'-padded nodes. */
int sbwt_node_labels(sbwt_index *ix, uint64_t *d_labels, uint8_t *d_full,
                     void *stream);

/* Every k-mer of the index, in node (colex) order: d_kmers needs n entries;
 * *count_out gets |K|. Synchronizes the stream. */
int sbwt_enumerate_kmers(sbwt_index *ix, uint64_t *d_kmers, uint64_t *count_out,
                         void *stream);

/* ---- learned negative prefilter (no reference counterpart) ------------
 * An MLP over one-hot 2-bit k-mers, 4k inputs -> 128 -> 64 -> 1 (ReLU,
 * logistic output), bf16 weights with FP32 accumulation, trained on the GPU;
 * and a blocked Bloom filter over the k-mers of the index the model rejects,
 * so that "maybe" = logit >= threshold or Bloom hit never misses an indexed
 * k-mer. Parameters are one flat FP32 vector: W1[4k][128], b1[128],
 * W2[128][64], b2[64], w3[64], b3 (row = input index 4t + c). */
typedef struct sbwt_mlp sbwt_mlp;
typedef struct sbwt_bloom sbwt_bloom;

/* A model for k-mers of length k with He-initialised weights from `seed`,
 * and scratch for training batches of up to max_batch k-mers. */
int sbwt_mlp_create(uint32_t k, uint64_t seed, uint64_t max_batch, sbwt_mlp **out);
int sbwt_mlp_free(sbwt_mlp *nn);
uint64_t sbwt_mlp_param_count(const sbwt_mlp *nn);
/* Host copies of the FP32 master parameters (set also resets Adam). */
int sbwt_mlp_set_params(sbwt_mlp *nn, const float *params, uint64_t count);
int sbwt_mlp_get_params(const sbwt_mlp *nn, float *params, uint64_t count);
/* Gradient (device, param_count floats) and mean binary cross-entropy of a
 * batch of top-aligned k-mers with labels 1 (present) / 0 (absent), taken
 * through the bf16-rounded weights the forward uses. */
int sbwt_mlp_grad(sbwt_mlp *nn, const uint64_t *d_kmers, const float *d_labels,
                  uint64_t batch, float *d_grad, float *d_loss, void *stream);
/* One Adam step (beta 0.9 / 0.999) on the batch's gradient; d_loss optional. */
int sbwt_mlp_train_step(sbwt_mlp *nn, const uint64_t *d_kmers, const float *d_labels,
                        uint64_t batch, float lr, float *d_loss, void *stream);
/* Logits of count k-mers on the matrix cores (reference != 0: the FP32
 * vector kernel with the same rounding, for cross-checking). */
int sbwt_mlp_logits(const sbwt_mlp *nn, const uint64_t *d_kmers, uint64_t count,
                    float *d_logits, int reference, void *stream);

/* Blocked Bloom filter over k-mers, sized for `expected` keys at
 * false-positive rate fpr. A key's bits lie in one 16-byte pair of a
 * 128-byte block, and the block is chosen by the key's minimizer (its
 * 16-mer, or k-mer for k < 16, of smallest hash), so the consecutive k-mers
 * of a read mostly share a cache line. Keys are top-aligned k-mers; only
 * their top 2k bits are hashed, so 32-character windows (sbwt_extract_kmers)
 * and k-mers with zeroed low bits (sbwt_enumerate_kmers) are the same key. */
int sbwt_bloom_create(uint32_t k, uint64_t expected, double fpr, sbwt_bloom **out);
int sbwt_bloom_free(sbwt_bloom *b);
uint64_t sbwt_bloom_bytes(const sbwt_bloom *b);
uint32_t sbwt_bloom_hashes(const sbwt_bloom *b);
/* Host copy of the filter's words (count = sbwt_bloom_bytes / 8). */
int sbwt_bloom_get_words(const sbwt_bloom *b, uint64_t *words, uint64_t count);
int sbwt_bloom_insert(sbwt_bloom *b, const uint64_t *d_keys, uint64_t count, void *stream);
int sbwt_bloom_query(const sbwt_bloom *b, const uint64_t *d_keys, uint64_t count,
                     uint8_t *d_out, void *stream);
/* The filter's answer for every k-mer of every row, read straight from the
 * packed rows (sbwt_pack): d_out[d_offsets[r] + j] for the k-mer at
 * character j of row r (d_offsets from sbwt_kmer_offsets for the filter's
 * k). The same answers as sbwt_bloom_query on the rows' extracted k-mers,
 * without materialising them. */
int sbwt_bloom_query_reads(const sbwt_bloom *b, const uint64_t *d_packed,
                           const uint64_t *d_row_starts, const uint64_t *d_offsets,
                           uint64_t rows, uint8_t *d_out, void *stream);

/* d_maybe[q] = 1 if logit(k-mer q) >= threshold or the Bloom filter (may be
 * NULL; built for the model's k) holds it; k-mers with 0 are absent from the
 * index for certain when the filter holds every indexed k-mer whose logit is
 * below threshold. Only the top 2k bits of each key are read. Two passes on
 * the stream: the matrix-core model, then the filter over the k-mers it
 * rejected. */
int sbwt_prefilter(const sbwt_mlp *nn, float threshold, const sbwt_bloom *b,
                   const uint64_t *d_kmers, uint64_t count, uint8_t *d_maybe,
                   void *stream);

/* ---- measurement ------------------------------------------------------ */

/* Random-access roof: `lanes` lanes each read `per_lane` (multiple of 8)
 * uniformly random lines of line_bytes from a 128-byte-aligned buffer of
 * `bytes`, 16 bytes per lane per line (full_line != 0: line_bytes/16 lanes
 * read one line together). Time it with events on `stream`; lines fetched =
 * lanes * per_lane / (full_line ? line_bytes/16 : 1). d_sink gets one u64 per
 * lane. No reference counterpart. */
int sbwt_gather_roof(const void *d_buf, uint64_t bytes, uint32_t line_bytes,
                     uint32_t full_line, uint32_t per_lane, uint64_t lanes,
                     uint64_t seed, uint64_t *d_sink, void *stream);

/* ---- host-side query ingest and result text ------------------------ */

/* Query text -> rows: one row per line, '>' lines and empty lines skipped
 * (KmerSearch::read_file, include/epicseq/kmer_search.hpp:319-342), a
 * trailing '\r' dropped. chars needs len bytes, row_starts row_cap entries
 * (rows + 1 are written). */
int sbwt_parse_lines(const char *text, uint64_t len, uint8_t *chars,
                     uint64_t *row_starts, uint64_t row_cap, uint64_t *rows_out);

/* FASTA / FASTQ text -> rows, one row per record. No reference counterpart:
 * the reference reads one sequence per line, so the k-mers that span a
 * multi-line record's line breaks are lost. The first non-empty line picks
 * the format:
 *   '>'  FASTA: a record's sequence lines, up to the next '>' line, joined;
 *   '@'  FASTQ: the sequence lines up to the '+' line, then quality lines
 *        holding as many characters as the sequence (skipped);
 *   else one row per line, as sbwt_parse_lines.
 * Lines end in '\n' or "\r\n"; empty lines are skipped. A record with no
 * sequence is an empty row. A malformed FASTQ record is SBWT_ERR_FORMAT.
 * chars needs len bytes, row_starts row_cap entries (rows + 1 are written). */
int sbwt_parse_records(const char *text, uint64_t len, uint8_t *chars,
                       uint64_t *row_starts, uint64_t row_cap, uint64_t *rows_out);

/* Result text (ResultsPersister::save_results_to_file,
 * include/epicseq/results_persister.hpp:82-96): each answer as a signed
 * decimal (~0 -> -1) and a space, '\n' after each row; offsets[rows+1] as
 * from sbwt_kmer_offsets. buf NULL: only *size_out is set. */
int sbwt_format_results(const uint64_t *results, const uint64_t *offsets,
                        uint64_t rows, char *buf, uint64_t cap,
                        uint64_t *size_out);

/* The same text written by the GPU: d_text[0 .. *d_size) gets the bytes
 * sbwt_format_results gives for the count answers d_results (device) with
 * row offsets d_offsets (device, rows + 1). If they need more than cap
 * bytes, nothing is written and *d_size still tells how many. d_scratch
 * holds sbwt_format_text_scratch(count) bytes. No host synchronisation.
 * Replaces the host printing of ResultsPersister::save_results_to_file
 * (include/epicseq/results_persister.hpp:82-96), the reference's slowest
 * stage. */
uint64_t sbwt_format_text_scratch(uint64_t count);
int sbwt_format_text(const uint64_t *d_results, const uint64_t *d_offsets,
                     uint64_t rows, uint64_t count, char *d_text, uint64_t cap,
                     uint64_t *d_size, void *d_scratch, uint64_t scratch_bytes,
                     void *stream);

/* Same text written to a file descriptor (1 = stdout). */
int sbwt_write_results(const uint64_t *results, const uint64_t *offsets,
                       uint64_t rows, int fd);

/* ---- batched read search with preallocated scratch ------------------ */

typedef struct sbwt_reads_plan sbwt_reads_plan;

/* Scratch for read searches of up to max_rows rows and max_chars characters
 * per call (packed string, non-ACGT mask, k-mer offsets, work prefix, item
 * map, scan temporaries) with segments of seg k-mers (0 = 256). Replaces
 * the per-call allocations of the reference's query pipeline
 * (KmerData::allocate_memory, include/epicseq/kmer_data.hpp:56-69). */
int sbwt_reads_plan_create(uint64_t max_rows, uint64_t max_chars, uint32_t seg,
                           sbwt_reads_plan **out);
int sbwt_reads_plan_free(sbwt_reads_plan *plan);

/* The whole read search of one batch: packing, k-mer offsets (into
 * d_offsets_out when not NULL, rows + 1 entries), work prefix, item map and
 * sbwt_search_reads with `flags`. d_out needs one entry per k-mer (at most
 * n_chars). Allocates nothing and does not synchronize, so a sequence of
 * batches can be captured in one HIP graph and replayed; calls sharing a
 * plan must be ordered on one stream. Each plan counts its own work items,
 * so calls on different plans may overlap on different streams, also on one
 * index (sbwt_search_reads calls on one index may not). */
int sbwt_reads_plan_search(sbwt_reads_plan *plan, const sbwt_index *ix,
                           const uint8_t *d_chars, uint64_t n_chars,
                           const uint64_t *d_row_starts, uint64_t rows,
                           uint64_t *d_out, uint64_t *d_offsets_out,
                           uint32_t flags, void *stream);

/* sbwt_reads_plan_search without a plan: the same query path (flags as
 * there, seg k-mers per work item, 32..4096, 0 = 256) on a scratch array
 * the library holds in each GPU's memory for the rest of the process
 * (about 45 MiB), so that the call allocates nothing. No counterpart in
 * the reference; it is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops. At most
 * SBWT_STATIC_MAX_ROWS rows and SBWT_STATIC_MAX_CHARS characters per call
 * (SBWT_ERR_ARG beyond). Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU race, so order them on one stream. */
#define SBWT_STATIC_MAX_ROWS 262144ull
#define SBWT_STATIC_MAX_CHARS 16777216ull
int sbwt_reads_search_static(const sbwt_index *ix, const uint8_t *d_chars,
                             uint64_t n_chars, const uint64_t *d_row_starts,
                             uint64_t rows, uint32_t seg, uint64_t *d_out,
                             uint64_t *d_offsets_out, uint32_t flags,
                             void *stream);

/* The found/absent bitmap of count answers: bit (bit_offset + q) of
 * d_bitmap is 1 iff d_results[q] != ~0 (words LSB-first; bits below
 * bit_offset and past the last answer are 0). Writes
 * ceil((bit_offset + count) / 64) words, so a caller can append batches
 * at any bit. d_found, when not NULL, receives the number of found
 * answers (a device word). No reference counterpart: the reference prints
 * each answer as text (include/epicseq/results_persister.hpp:82-96). */
int sbwt_found_bitmap(const uint64_t *d_results, uint64_t count,
                      uint32_t bit_offset, uint64_t *d_bitmap,
                      uint64_t *d_found, void *stream);

/* The `format` argument of sbwt_search_text: one input kind ORed with one
 * output kind. The binary outputs are a 16-byte header ("SBWTRES1", u32
 * kind: 1 nodes, 2 bitmap; u32 k) and then one block per batch: u64 rows,
 * u64 k-mers, the k-mer count of each row (rows x u64), then the answers
 * (k-mers x u64) or their found bits (ceil(k-mers / 64) x u64, as
 * sbwt_found_bitmap with bit_offset 0); all little-endian. */
#define SBWT_TEXT_LINES 0u      /* one row per line (sbwt_parse_lines) */
#define SBWT_TEXT_RECORDS 1u    /* FASTA / FASTQ records, or lines (sbwt_parse_records) */
#define SBWT_OUTPUT_TEXT 0u     /* the result text sbwt_write_results writes */
#define SBWT_OUTPUT_NODES 0x100u  /* binary blocks of answers */
#define SBWT_OUTPUT_BITMAP 0x200u /* binary blocks of found bits */

/* The whole query path over query text in host memory, pipelined: batches
 * of about batch_chars characters (0 = 64 MiB; rows never straddle
 * batches), three in flight. While the device searches batch b (on a reads
 * plan), batch b + 1 is parsed into pinned memory and copied in, and batch
 * b - 1's result text, printed on the GPU (sbwt_format_text), is copied
 * back and written to fd: the bytes sbwt_write_results writes (fd < 0: not
 * written). format: SBWT_TEXT_* | SBWT_OUTPUT_* (with a binary output,
 * each batch's answers or found bits are copied back and written instead of
 * the text). flags as sbwt_search_reads (not SBWT_COUNT). stats, when not
 * NULL: rows, k-mers, k-mers found. Returns when every batch is written. No
 * reference counterpart: the reference runs its stages one after another
 * (KmerSearch::search, include/epicseq/kmer_search.hpp:97-153). */
int sbwt_search_text(const sbwt_index *ix, const char *text, uint64_t len,
                     uint32_t format, uint32_t flags, uint64_t batch_chars,
                     int fd, uint64_t *stats);

/* Message of the calling thread's last failure ("" if none). */
const char *sbwt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
