/*
 * CPU oracle for the SBWT k-mer membership search (test infrastructure only;
 * see oracle.h for what it restates and who may call it).
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

static u64 lowmask(u32 bits) { return bits >= 64 ? ~0ULL : ((1ULL << bits) - 1ULL); }

/* Strict code used by the builder and the N-breaking read path: A/C/G/T in
 * either case, -1 for anything else. */
static int code_strict(u8 ch) {
  switch (ch) {
  case 'A': case 'a': return 0;
  case 'C': case 'c': return 1;
  case 'G': case 'g': return 2;
  case 'T': case 't': return 3;
  default: return -1;
  }
}

/* The reference packing table, create_kmer_string.hpp:20: C/c->1, G/g->2,
 * T/t->3, every other byte (A, N, '\0', ...) -> 0. */
static int code_ref(u8 ch) {
  switch (ch) {
  case 'C': case 'c': return 1;
  case 'G': case 'g': return 2;
  case 'T': case 't': return 3;
  default: return 0;
  }
}

/* ===================================================================== */
/* SBWT construction (brute force)                                       */
/* ===================================================================== */

typedef struct { u64 key; u32 len; } node_t;

struct orc_sbwt {
  u32 k;
  u64 n, nwords, nkmers;
  u64 *rows[4];
  u64 *group;
  u64 C[5];
  u64 *key;
  u8 *len;
};

static int cmp_u64(const void *a, const void *b) {
  u64 x = *(const u64 *)a, y = *(const u64 *)b;
  return x < y ? -1 : (x > y);
}

/* Colex order of padded k-mers: compare the code (last char most
 * significant), and on equal codes the one with fewer real characters (more
 * leading '$', '$' < A) first. */
static int cmp_node(const void *a, const void *b) {
  const node_t *x = (const node_t *)a, *y = (const node_t *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->len < y->len ? -1 : (x->len > y->len);
}

static void set_bit(u64 *w, u64 i) { w[i >> 6] |= 1ULL << (63 - (i & 63)); }

/* index of (key,len) in sorted nodes, or -1 */
static int64_t find_node(const node_t *r, u64 n, u64 key, u32 len) {
  u64 lo = 0, hi = n;
  node_t q = {key, len};
  while (lo < hi) {
    u64 mid = lo + (hi - lo) / 2;
    if (cmp_node(&r[mid], &q) < 0) lo = mid + 1; else hi = mid;
  }
  if (lo < n && r[lo].key == key && r[lo].len == len) return (int64_t)lo;
  return -1;
}

orc_sbwt *orc_sbwt_build(const u8 *seq, const u64 *starts, u64 nseq, u32 k) {
  if (k < 1 || k > 32) return NULL;
  const u64 kmask = lowmask(2 * k);
  const u32 top = 2 * (k - 1);
  /* 1. all k-mers (colex key: char p in bits 2p), windows with non-ACGT skipped */
  u64 cap = 0;
  for (u64 s = 0; s < nseq; s++) {
    u64 L = starts[s + 1] - starts[s];
    if (L >= k) cap += L - k + 1;
  }
  u64 *keys = (u64 *)malloc((cap ? cap : 1) * sizeof(u64));
  u64 nk = 0;
  for (u64 s = 0; s < nseq; s++) {
    u64 run = 0, key = 0;
    for (u64 p = starts[s]; p < starts[s + 1]; p++) {
      int c = code_strict(seq[p]);
      if (c < 0) { run = 0; key = 0; continue; }
      key = ((key >> 2) | ((u64)c << top)) & kmask;
      if (++run >= k) keys[nk++] = key;
    }
  }
  qsort(keys, nk, sizeof(u64), cmp_u64);
  u64 u = 0;
  for (u64 i = 0; i < nk; i++)
    if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
  nk = u;

  /* 2. sources: no k-mer y with suffix_{k-1}(y) == prefix_{k-1}(x); then the
   * padded nodes $^{k-i} x[0..i-1], i = 0..k-1 (i = 0 is the root $^k, which
   * is always present so that column 0 is the root, matrix_sbwt.hpp:151). */
  const u64 pmask = lowmask(2 * (k - 1));
  node_t *pad = (node_t *)malloc(sizeof(node_t) * (1 + nk * (u64)k));
  u64 np = 0;
  pad[np].key = 0; pad[np].len = 0; np++;
  for (u64 i = 0; i < nk; i++) {
    u64 want = keys[i] & pmask; /* x[0..k-2] as a (k-1)-code */
    /* any y with (key_y >> 2) == want  <=>  key_y in [want<<2, want<<2 | 3] */
    u64 lo = 0, hi = nk, target = want << 2;
    while (lo < hi) {
      u64 mid = lo + (hi - lo) / 2;
      if (keys[mid] < target) lo = mid + 1; else hi = mid;
    }
    int has_pred = (lo < nk && (keys[lo] >> 2) == want);
    if (has_pred) continue;
    for (u32 len = 1; len < k; len++) {
      pad[np].key = (keys[i] & lowmask(2 * len)) << (2 * (k - len));
      pad[np].len = len;
      np++;
    }
  }
  qsort(pad, np, sizeof(node_t), cmp_node);
  u = 0;
  for (u64 i = 0; i < np; i++)
    if (u == 0 || cmp_node(&pad[i], &pad[u - 1]) != 0) pad[u++] = pad[i];
  np = u;

  /* 3. merge into the padded k-spectrum R */
  u64 n = nk + np;
  node_t *R = (node_t *)malloc(sizeof(node_t) * n);
  {
    u64 a = 0, b = 0, o = 0;
    while (a < nk || b < np) {
      node_t ka = {a < nk ? keys[a] : 0, k};
      if (b >= np || (a < nk && cmp_node(&ka, &pad[b]) < 0)) { R[o++] = ka; a++; }
      else R[o++] = pad[b++];
    }
  }
  free(keys);
  free(pad);

  orc_sbwt *s = (orc_sbwt *)calloc(1, sizeof(orc_sbwt));
  s->k = k;
  s->n = n;
  s->nkmers = nk;
  /* rows padded to a multiple of 4096 bits (bit_vector.hpp:103-104) */
  s->nwords = ((n + 4095) / 4096) * 64;
  for (int c = 0; c < 4; c++) s->rows[c] = (u64 *)calloc(s->nwords ? s->nwords : 1, sizeof(u64));
  s->group = (u64 *)calloc(s->nwords ? s->nwords : 1, sizeof(u64));
  s->key = (u64 *)malloc(sizeof(u64) * (n ? n : 1));
  s->len = (u8 *)malloc(n ? n : 1);

  /* 4. subsets: node j gets {c : suffix_{k-1}(R_j).c in R} iff it starts its
   * suffix group (j == 0 or its (k-1)-suffix differs from R_{j-1}'s). The
   * (k-1)-suffix is identified by (key >> 2, min(len, k-1)). */
  for (u64 j = 0; j < n; j++) {
    s->key[j] = R[j].key;
    s->len[j] = (u8)R[j].len;
    u32 sl = R[j].len < k - 1 ? R[j].len : k - 1;
    int first = 1;
    if (j > 0) {
      u32 psl = R[j - 1].len < k - 1 ? R[j - 1].len : k - 1;
      first = !((R[j - 1].key >> 2) == (R[j].key >> 2) && psl == sl);
    }
    if (!first) continue;
    set_bit(s->group, j);
    for (int c = 0; c < 4; c++) {
      u64 qk = (R[j].key >> 2) | ((u64)c << top);
      u32 ql = R[j].len + 1 < k ? R[j].len + 1 : k;
      if (find_node(R, n, qk, ql) >= 0) set_bit(s->rows[c], j);
    }
  }
  free(R);
  /* C array, matrix_sbwt.hpp:150-155 */
  s->C[0] = 1;
  for (int c = 0; c < 4; c++) {
    u64 ones = 0;
    for (u64 w = 0; w < s->nwords; w++) ones += (u64)__builtin_popcountll(s->rows[c][w]);
    s->C[c + 1] = s->C[c] + ones;
  }
  return s;
}

void orc_sbwt_free(orc_sbwt *s) {
  if (!s) return;
  for (int c = 0; c < 4; c++) free(s->rows[c]);
  free(s->group);
  free(s->key);
  free(s->len);
  free(s);
}
u64 orc_sbwt_n(const orc_sbwt *s) { return s->n; }
u64 orc_sbwt_nwords(const orc_sbwt *s) { return s->nwords; }
u64 orc_sbwt_kmers(const orc_sbwt *s) { return s->nkmers; }
void orc_sbwt_rows(const orc_sbwt *s, u64 *a, u64 *c, u64 *g, u64 *t, u64 *group) {
  u64 *dst[4] = {a, c, g, t};
  for (int x = 0; x < 4; x++)
    if (dst[x]) memcpy(dst[x], s->rows[x], s->nwords * sizeof(u64));
  if (group) memcpy(group, s->group, s->nwords * sizeof(u64));
}
void orc_sbwt_C(const orc_sbwt *s, u64 C[5]) { memcpy(C, s->C, 5 * sizeof(u64)); }
void orc_sbwt_labels(const orc_sbwt *s, u64 *key, u8 *len) {
  memcpy(key, s->key, s->n * sizeof(u64));
  memcpy(len, s->len, s->n);
}

/* ===================================================================== */
/* Bit vector files, bit_vector.hpp:65-128                               */
/* ===================================================================== */

int orc_bv_write(const char *path, const u64 *words, u64 n) {
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  u8 hdr[8];
  for (int i = 0; i < 8; i++) hdr[i] = (u8)(n >> (8 * i));
  fwrite(hdr, 1, 8, f);
  u64 nbytes = (n + 7) / 8;
  for (u64 b = 0; b < nbytes; b++) {
    /* byte b holds bits 8b..8b+7, first bit in the byte's MSB */
    u8 v = (u8)(words[b >> 3] >> (56 - 8 * (b & 7)));
    fputc(v, f);
  }
  fclose(f);
  return 0;
}

u64 orc_bv_header(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return 0;
  u8 hdr[8];
  u64 n = 0;
  if (fread(hdr, 1, 8, f) == 8)
    for (int i = 0; i < 8; i++) n |= ((u64)hdr[i]) << (8 * i);
  fclose(f);
  return n;
}

int orc_bv_read(const char *path, u64 *words, u64 nwords) {
  u64 n = orc_bv_header(path);
  if (n == 0) return 1;
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  fseek(f, 8, SEEK_SET);
  memset(words, 0, nwords * sizeof(u64));
  u64 nbytes = (n + 7) / 8;
  for (u64 b = 0; b < nbytes; b++) {
    int v = fgetc(f);
    if (v == EOF) { fclose(f); return 1; }
    /* the reference reads little-endian words and byte-swaps them
     * (rank_data_structures.hpp:98): byte b lands in bits 63-8(b&7)..56-8(b&7) */
    words[b >> 3] |= ((u64)(u8)v) << (56 - 8 * (b & 7));
  }
  /* bits past n in the last byte are data the reference keeps; we clear
   * them so rank beyond n is well defined */
  if (n & 63) words[n >> 6] &= ~(~0ULL >> (n & 63));
  fclose(f);
  return 0;
}

/* ===================================================================== */
/* Poppy rank structures, rank_data_structures.hpp:59-69,142-198          */
/* ===================================================================== */

struct orc_poppy {
  u32 sb, bb;
  int version;
  u32 log_hb;
  u64 n, n0, n12, ones;
  u64 *L0, *L12;
};

static u32 ilog2(u64 x) { u32 r = 0; while ((1ULL << r) < x) r++; return r; }

orc_poppy *orc_poppy_build(const u64 *words, u64 n, u32 sb, int version) {
  if (!(sb == 256 || sb == 512 || sb == 1024 || sb == 2048 || sb == 4096)) return NULL;
  if (version == ORC_CUM_POPPY && sb == 4096) return NULL; /* parameters.hpp:273-277 */
  orc_poppy *p = (orc_poppy *)calloc(1, sizeof(orc_poppy));
  p->sb = sb;
  p->bb = sb / 4;
  p->version = version;
  p->n = n;
  p->log_hb = sb == 4096 ? 31 : 32;
  p->n0 = 1 + (n >> p->log_hb);
  p->n12 = 1 + ((n + sb - 1) / sb);
  p->L0 = (u64 *)calloc(p->n0, sizeof(u64));
  p->L12 = (u64 *)calloc(p->n12, sizeof(u64));
  const u64 sb_per_hb = (1ULL << p->log_hb) / sb;
  const u32 wbb = p->bb / 64 ? p->bb / 64 : 1; /* words per basic block */
  const u32 big = (sb == 4096);
  u64 abs = 0, iw = 0, i12 = 0;
  for (u64 i0 = 0; i0 < p->n0; i0++) {
    p->L0[i0] = abs;
    u64 rel = 0;
    for (u64 r = 0; r < sb_per_hb; r++) {
      if (i12 >= p->n12 - 1) {
        p->L12[i12] = rel << (32 + big);
        p->ones = abs;
        return p;
      }
      u64 bbc[4];
      for (int b = 0; b < 4; b++) {
        bbc[b] = 0;
        for (u32 j = 0; j < wbb; j++) bbc[b] += (u64)__builtin_popcountll(words[iw + (u64)b * wbb + j]);
      }
      iw += 4ULL * wbb;
      if (version == ORC_POPPY) {
        p->L12[i12] = (rel << (32 + big)) | (bbc[0] << (20 + 2 * big)) | (bbc[1] << (10 + big)) | bbc[2];
      } else if (sb < 2048) {
        p->L12[i12] = (rel << 32) | (bbc[0] << 20) | ((bbc[0] + bbc[1]) << 10) | (bbc[0] + bbc[1] + bbc[2]);
      } else {
        p->L12[i12] = (rel << 32) | (bbc[0] << 22) | ((bbc[0] + bbc[1]) << 11) | (bbc[0] + bbc[1] + bbc[2]);
      }
      i12++;
      u64 cs = bbc[0] + bbc[1] + bbc[2] + bbc[3];
      abs += cs;
      rel += cs;
    }
  }
  p->ones = abs;
  return p;
}

void orc_poppy_free(orc_poppy *p) {
  if (!p) return;
  free(p->L0);
  free(p->L12);
  free(p);
}
u64 orc_poppy_n0(const orc_poppy *p) { return p->n0; }
u64 orc_poppy_n12(const orc_poppy *p) { return p->n12; }
const u64 *orc_poppy_L0(const orc_poppy *p) { return p->L0; }
const u64 *orc_poppy_L12(const orc_poppy *p) { return p->L12; }
u64 orc_poppy_ones(const orc_poppy *p) { return p->ones; }

/* popcount_before_basic_block, rank_kernel.hpp:253-296 */
static u64 before_bb(const orc_poppy *p, u64 i) {
  const u32 logsb = ilog2(p->sb), logbb = ilog2(p->bb);
  const u64 e = p->L12[i >> logsb];
  const u64 b = (i & (p->sb - 1)) >> logbb; /* basic block inside superblock */
  if (p->version == ORC_POPPY) {
    if (p->sb == 4096) {
      u64 f = (e & 0x1ffffffffULL) >> ((3 - b) * 11);
      return p->L0[i >> 31] + (e >> 33) + ((f & 0x7ff) + ((f >> 11) & 0x7ff) + (f >> 22));
    }
    u64 f = (e & 0x3fffffffULL) >> ((3 - b) * 10);
    return p->L0[i >> 32] + (e >> 32) + ((f & 0x3ff) + ((f >> 10) & 0x3ff) + ((f >> 20) & 0x3ff));
  }
  if (p->sb == 2048) {
    u64 f = (e & 0xffffffffULL) >> ((3 - b) * 11);
    return p->L0[i >> 32] + (e >> 32) + (f & 0x7ff);
  }
  u64 f = (e & 0x3fffffffULL) >> ((3 - b) * 10);
  return p->L0[i >> 32] + (e >> 32) + (f & 0x3ff);
}

/* popcount_inside_basic_block_normal, rank_kernel.hpp:212-251: the 1-bits
 * of i's basic block that precede i, MSB-first. The reference shifts by
 * 64-(i&63), which is 64 at word-aligned i (a shift PTX clamps to zero and
 * C leaves undefined); the mask ~(~0 >> r) is zero at r == 0 instead. */
static u64 inside_bb(const orc_poppy *p, const u64 *words, u64 i) {
  const u64 bbw = p->bb / 64;          /* words per basic block (>= 1) */
  const u64 first = (i / p->bb) * bbw; /* first word of i's basic block */
  const u64 iw = i >> 6;
  u64 cnt = 0;
  for (u64 w = first; w < iw; w++) cnt += (u64)__builtin_popcountll(words[w]);
  const u32 r = (u32)(i & 63);
  if (r) cnt += (u64)__builtin_popcountll(words[iw] & ~(~0ULL >> r));
  return cnt;
}

u64 orc_rank_poppy(const orc_poppy *p, const u64 *words, u64 i) {
  return before_bb(p, i) + inside_bb(p, words, i);
}

u64 orc_rank_naive(const u64 *words, u64 i) {
  u64 cnt = 0;
  for (u64 w = 0; w < (i >> 6); w++) cnt += (u64)__builtin_popcountll(words[w]);
  u32 r = (u32)(i & 63);
  if (r) cnt += (u64)__builtin_popcountll(words[i >> 6] & ~(~0ULL >> r));
  return cnt;
}

/* ===================================================================== */
/* Search index: matrix_sbwt.hpp:150-157 + rank + presearch + search      */
/* ===================================================================== */

struct orc_index {
  u64 n;
  const u64 *rows[4];
  orc_poppy *rank[4];
  u64 C[5];
};

orc_index *orc_index_open(const u64 *a, const u64 *c, const u64 *g, const u64 *t,
                          u64 n, u32 sb, int version) {
  orc_index *x = (orc_index *)calloc(1, sizeof(orc_index));
  x->n = n;
  const u64 *r[4] = {a, c, g, t};
  for (int i = 0; i < 4; i++) {
    x->rows[i] = r[i];
    x->rank[i] = orc_poppy_build(r[i], n, sb, version);
    if (!x->rank[i]) { orc_index_free(x); return NULL; }
  }
  x->C[0] = 1;
  for (int i = 0; i < 4; i++) x->C[i + 1] = x->C[i] + orc_poppy_ones(x->rank[i]);
  return x;
}

void orc_index_free(orc_index *x) {
  if (!x) return;
  for (int i = 0; i < 4; i++) orc_poppy_free(x->rank[i]);
  free(x);
}
void orc_index_C(const orc_index *x, u64 C[5]) { memcpy(C, x->C, sizeof(x->C)); }

static inline u64 xrank(const orc_index *x, int c, u64 i) {
  return orc_rank_poppy(x->rank[c], x->rows[c], i);
}
u64 orc_index_rank(const orc_index *x, int c, u64 i) { return xrank(x, c, i); }

void orc_presearch(const orc_index *x, u32 m, u64 *left, u64 *right) {
  const u64 P = 1ULL << (2 * m);
#pragma omp parallel for schedule(static)
  for (u64 idx = 0; idx < P; idx++) {
    int c = (int)((idx >> (2 * (m - 1))) & 3);
    u64 l = x->C[c], r = x->C[c + 1] - 1;
    for (int sh = (int)(2 * m) - 4; sh >= 0; sh -= 2) {
      c = (int)((idx >> sh) & 3);
      l = x->C[c] + xrank(x, c, l);
      r = x->C[c] + xrank(x, c, r + 1) - 1;
    }
    left[idx] = l;
    right[idx] = r;
  }
}

static inline u64 search_one(const orc_index *x, u64 kmer, u32 k, u32 m,
                             const u64 *left, const u64 *right, u32 *steps,
                             u32 *lines, u64 cpc, u64 cbytes) {
  u64 l, r;
  u32 st = 0, ln = 0;
  if (m) {
    u64 mm = kmer >> (64 - 2 * m);
    l = left[mm];
    r = right[mm];
    ln = 1;
  } else {
    l = 0;
    r = x->C[4] - 1;
  }
  for (u32 i = m; i < k; i++) {
    int c = (int)((kmer >> (62 - 2 * i)) & 3);
    if (lines) {
      u64 la = (l / cpc) * cbytes / 128, lb = ((r + 1) / cpc) * cbytes / 128;
      ln += (la == lb) ? 1 : 2;
    }
    l = x->C[c] + xrank(x, c, l);
    r = x->C[c] + xrank(x, c, r + 1) - 1;
    st++;
    if (l > r) break;
  }
  if (steps) *steps = st;
  if (lines) *lines = ln;
  return l > r ? ~0ULL : l;
}

void orc_search(const orc_index *x, const u64 *kmers, u64 count, u32 k, u32 m,
                const u64 *left, const u64 *right, u64 *out, int threads) {
#ifdef _OPENMP
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(static, 4096) num_threads(nt)
#endif
  for (u64 q = 0; q < count; q++)
    out[q] = search_one(x, kmers[q], k, m, left, right, NULL, NULL, 1, 1);
}

void orc_search_count(const orc_index *x, const u64 *kmers, u64 count, u32 k,
                      u32 m, const u64 *left, const u64 *right, u64 cpc,
                      u64 cbytes, u64 *out, u32 *steps, u32 *lines) {
#pragma omp parallel for schedule(static, 4096)
  for (u64 q = 0; q < count; q++)
    out[q] = search_one(x, kmers[q], k, m, left, right, &steps[q], &lines[q], cpc, cbytes);
}

/* ===================================================================== */
/* Query ingest and output                                               */
/* ===================================================================== */

void orc_pack(const u8 *raw, u64 nwords, u64 *packed) {
  for (u64 w = 0; w < nwords; w++) {
    u64 v = 0;
    for (u32 t = 0; t < 32; t++) v |= ((u64)code_ref(raw[32 * w + t])) << (62 - 2 * t);
    packed[w] = v;
  }
}

u64 orc_extract(const u64 *packed, u64 pos) {
  u64 b = pos << 1;
  return (packed[b >> 6] << (b & 63)) | ((packed[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}

void orc_kmer_offsets(const u64 *row_starts, u64 rows, u32 k, u64 *offsets) {
  u64 o = 0;
  for (u64 r = 0; r < rows; r++) {
    offsets[r] = o;
    u64 L = row_starts[r + 1] - row_starts[r];
    if (L >= k) o += L - k + 1;
  }
  offsets[rows] = o;
}

u64 orc_positions(const u64 *row_starts, u64 rows, u32 k, u64 *out) {
  u64 o = 0;
  for (u64 r = 0; r < rows; r++) {
    u64 L = row_starts[r + 1] - row_starts[r];
    if (L < k) continue;
    for (u64 i = 0; i + k <= L; i++) out[o++] = row_starts[r] + i;
  }
  return o;
}

u64 orc_read_kmers(const u8 *chars, const u64 *row_starts, u64 rows, u32 k,
                   int n_breaks, u64 *kmers, u8 *valid) {
  u64 o = 0;
  for (u64 r = 0; r < rows; r++) {
    u64 L = row_starts[r + 1] - row_starts[r];
    if (L < k) continue;
    const u8 *s = chars + row_starts[r];
    for (u64 i = 0; i + k <= L; i++) {
      u64 v = 0;
      int ok = 1;
      for (u32 t = 0; t < k; t++) {
        int c = code_strict(s[i + t]);
        if (c < 0) { ok = 0; c = code_ref(s[i + t]); }
        v |= ((u64)c) << (62 - 2 * t);
      }
      kmers[o] = v;
      if (valid) valid[o] = (u8)(n_breaks ? ok : 1);
      o++;
    }
  }
  return o;
}

u64 orc_parse_lines(const u8 *text, u64 len, u8 *chars, u64 *row_starts) {
  u64 rows = 0, nch = 0, p = 0;
  while (p < len) {
    u64 e = p;
    while (e < len && text[e] != '\n') e++;
    u64 le = e;
    if (le > p && text[le - 1] == '\r') le--;
    if (le > p && text[p] != '>') {
      row_starts[rows++] = nch;
      memcpy(chars + nch, text + p, le - p);
      nch += le - p;
    }
    p = e + 1;
  }
  row_starts[rows] = nch;
  return rows;
}

u64 orc_format_results(const u64 *res, const u64 *offsets, u64 rows, char *buf) {
  u64 o = 0;
  char tmp[32];
  for (u64 r = 0; r < rows; r++) {
    for (u64 j = offsets[r]; j < offsets[r + 1]; j++) {
      int w = snprintf(tmp, sizeof tmp, "%lld ", (long long)(int64_t)res[j]);
      if (buf) memcpy(buf + o, tmp, (size_t)w);
      o += (u64)w;
    }
    if (buf) buf[o] = '\n';
    o++;
  }
  return o;
}
