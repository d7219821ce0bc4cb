/*
 * CPU oracle for the SBWT k-mer membership search.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker / the timed CPU baseline.
 *
 * It restates, in plain C, the reference algorithm of epicharri/sbwt_gpu:
 *   - bit vector file format          include/bit_vector.hpp:65-128
 *   - poppy / cum-poppy L0/L12        include/rank_data_structures.hpp:59-69,142-198
 *   - C array                         include/matrix_sbwt.hpp:150-155
 *   - rank                            include/gpu/rank_kernel.hpp:212-296,409-415
 *   - presearch                       include/gpu/presearch_kernel.hpp:19-37
 *   - search                          include/gpu/search_kernel.hpp:61-104
 *   - 2-bit packing                   include/gpu/create_kmer_string.hpp:18-42
 *   - k-mer start positions           include/gpu/positions_kernel.hpp:18-36
 *   - query line parsing              include/epicseq/kmer_search.hpp:319-342
 *   - result text                     include/epicseq/results_persister.hpp:82-96
 * plus a brute-force SBWT builder (the reference ships none; its bit vectors
 * come from an external tool) written from the SBWT definition of Alanko,
 * Puglisi and Vuohtoniemi: colexicographically sorted padded k-spectrum,
 * one subset of {A,C,G,T} per node, non-empty only at suffix-group starts.
 *
 * Bit order everywhere: bit i of a row lives in word i>>6 at bit 63-(i&63)
 * (MSB-first), which is the order the reference reaches after its byte swap
 * (rank_data_structures.hpp:98).
 */
#ifndef SBWT_ORACLE_H
#define SBWT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SBWT construction (brute force) -------------------------------- */
typedef struct orc_sbwt orc_sbwt;

/* seq: concatenated sequences, starts[nseq+1] offsets. Non-ACGT characters
 * (after case folding) break k-mers. k in [1,32]. */
orc_sbwt *orc_sbwt_build(const uint8_t *seq, const uint64_t *starts,
                         uint64_t nseq, uint32_t k);
void orc_sbwt_free(orc_sbwt *s);
uint64_t orc_sbwt_n(const orc_sbwt *s);
uint64_t orc_sbwt_nwords(const orc_sbwt *s);         /* padded words per row */
uint64_t orc_sbwt_kmers(const orc_sbwt *s);          /* |K| distinct k-mers */
/* rows: 4 buffers of nwords each (A,C,G,T), group: nwords (suffix-group starts) */
void orc_sbwt_rows(const orc_sbwt *s, uint64_t *a, uint64_t *c, uint64_t *g,
                   uint64_t *t, uint64_t *group);
void orc_sbwt_C(const orc_sbwt *s, uint64_t C[5]);
/* node labels: colex key (char at position p in bits 2p..2p+1, '$' = 0) and
 * the number of real (non-'$') characters */
void orc_sbwt_labels(const orc_sbwt *s, uint64_t *key, uint8_t *len);

/* ---- bit vector file format (bit_vector.hpp:65-128) ------------------ */
/* writes [u64 LE n][ceil(n/8) bytes], bit i = byte 8+(i>>3), bit 7-(i&7) */
int orc_bv_write(const char *path, const uint64_t *words_msb, uint64_t n);
/* returns n (0 on error); words must hold ceil(n/4096)*64 words */
uint64_t orc_bv_header(const char *path);
int orc_bv_read(const char *path, uint64_t *words_msb, uint64_t nwords);

/* ---- poppy rank (rank_data_structures.hpp, rank_kernel.hpp) --------- */
typedef struct orc_poppy orc_poppy;
enum { ORC_POPPY = 1, ORC_CUM_POPPY = 2 };
orc_poppy *orc_poppy_build(const uint64_t *words, uint64_t n, uint32_t sb,
                           int version);
void orc_poppy_free(orc_poppy *p);
uint64_t orc_poppy_n0(const orc_poppy *p);
uint64_t orc_poppy_n12(const orc_poppy *p);
const uint64_t *orc_poppy_L0(const orc_poppy *p);
const uint64_t *orc_poppy_L12(const orc_poppy *p);
uint64_t orc_poppy_ones(const orc_poppy *p);
uint64_t orc_rank_poppy(const orc_poppy *p, const uint64_t *words, uint64_t i);
uint64_t orc_rank_naive(const uint64_t *words, uint64_t i);

/* ---- search index (4 rows + poppy + C) ------------------------------ */
typedef struct orc_index orc_index;
/* rows are borrowed (must outlive the index) */
orc_index *orc_index_open(const uint64_t *a, const uint64_t *c,
                          const uint64_t *g, const uint64_t *t, uint64_t n,
                          uint32_t sb, int version);
void orc_index_free(orc_index *x);
void orc_index_C(const orc_index *x, uint64_t C[5]);
uint64_t orc_index_rank(const orc_index *x, int c, uint64_t i);

/* presearch_kernel.hpp:19-37: 4^m entries, m in [1,16] */
void orc_presearch(const orc_index *x, uint32_t m, uint64_t *left,
                   uint64_t *right);

/* search_kernel.hpp:61-104 over top-aligned k-mers (char 0 in bits 63:62).
 * m = 0 searches without presearch. threads <= 0: all OpenMP threads. */
void orc_search(const orc_index *x, const uint64_t *kmers, uint64_t count,
                uint32_t k, uint32_t m, const uint64_t *left,
                const uint64_t *right, uint64_t *out, int threads);

/* Same search, counting for each query the executed steps and the 128-byte
 * lines the two rank calls of each step touch under a rank layout whose
 * chunks hold `cols_per_chunk` columns in `chunk_bytes` bytes (each of the
 * four rows in its own array), plus 1 line for a presearch gather. */
void orc_search_count(const orc_index *x, const uint64_t *kmers,
                      uint64_t count, uint32_t k, uint32_t m,
                      const uint64_t *left, const uint64_t *right,
                      uint64_t cols_per_chunk, uint64_t chunk_bytes,
                      uint64_t *out, uint32_t *steps, uint32_t *lines);

/* ---- query ingest ---------------------------------------------------- */
/* create_kmer_string.hpp:22-42: nwords words from 32*nwords chars */
void orc_pack(const uint8_t *raw, uint64_t nwords, uint64_t *packed);
/* search_kernel.hpp:66-68: top-aligned k-mer starting at char `pos` */
uint64_t orc_extract(const uint64_t *packed, uint64_t pos);
/* positions_kernel.hpp:18-36 with rows shorter than k yielding no k-mers;
 * returns the number of positions written */
uint64_t orc_positions(const uint64_t *row_starts, uint64_t rows, uint32_t k,
                       uint64_t *out);
/* per-row first output index (rows+1 entries, exclusive scan of k-mer counts) */
void orc_kmer_offsets(const uint64_t *row_starts, uint64_t rows, uint32_t k,
                      uint64_t *offsets);
/* k-mers of every row, top-aligned; valid[j]=0 marks a window with a
 * non-ACGT char (only when n_breaks != 0; else such chars code as A, as the
 * reference's packing table does) */
uint64_t orc_read_kmers(const uint8_t *chars, const uint64_t *row_starts,
                        uint64_t rows, uint32_t k, int n_breaks,
                        uint64_t *kmers, uint8_t *valid);
/* kmer_search.hpp:319-342: parse lines, '>' lines skipped, one row per
 * non-empty line. Returns rows; chars/row_starts sized by caller (len, len+1).
 * Trailing '\r' is dropped and a last line without '\n' keeps its last char. */
uint64_t orc_parse_lines(const uint8_t *text, uint64_t len, uint8_t *chars,
                         uint64_t *row_starts);
/* results_persister.hpp:82-96: "<int64> " per k-mer, "\n" per row.
 * Returns bytes written (buf may be NULL to size). */
uint64_t orc_format_results(const uint64_t *results, const uint64_t *offsets,
                            uint64_t rows, char *buf);

#ifdef __cplusplus
}
#endif
#endif
